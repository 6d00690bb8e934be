!-----------------------------------------------------------------------
! ysx_bind -- iso_c_binding interface to libysx.so (include/ysx.h).
!
! The reference calls its solvers through Fortran module procedures and
! module globals (src/yspec.f90:247-473); a Fortran host binds the C-ABI
! entry points below instead.  Arrays are passed as assumed-size dummies:
! column-major, complex data as interleaved (re, im) pairs of c_double,
! exactly the memory layout of complex(c_double_complex).
!-----------------------------------------------------------------------
module ysx_bind
  use iso_c_binding
  implicit none

  interface
     integer(c_int) function ysx_create(ctx, device) bind(C, name='ysx_create')
       import :: c_int, c_ptr
       type(c_ptr) :: ctx
       integer(c_int), value :: device
     end function ysx_create

     subroutine ysx_destroy(ctx) bind(C, name='ysx_destroy')
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine ysx_destroy

     type(c_ptr) function ysx_last_error(ctx) bind(C, name='ysx_last_error')
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function ysx_last_error

     integer(c_int) function ysx_set_model(ctx, nknot, nic, noc, ians, tref, deck) &
          bind(C, name='ysx_set_model')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: nknot, nic, noc, ians
       real(c_double), value :: tref
       real(c_double) :: deck(*)
     end function ysx_set_model

     integer(c_int) function ysx_set_run(ctx, ip, fp, lat, lon) bind(C, name='ysx_set_run')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int) :: ip(*)
       real(c_double) :: fp(*), lat(*), lon(*)
     end function ysx_set_run

     integer(c_int) function ysx_run_info(ctx, iout, fout) bind(C, name='ysx_run_info')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int) :: iout(*)
       real(c_double) :: fout(*)
     end function ysx_run_info

     integer(c_int) function ysx_forward(ctx, seis) bind(C, name='ysx_forward')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: seis(*)
     end function ysx_forward

     integer(c_int) function ysx_spectra(ctx, ur, ut, up) bind(C, name='ysx_spectra')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: ur(*), ut(*), up(*)
     end function ysx_spectra

     integer(c_int) function ysx_seismograms(ctx, seis) bind(C, name='ysx_seismograms')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: seis(*)
     end function ysx_seismograms

     function c_strlen(s) bind(C, name='strlen')
       import :: c_ptr, c_size_t
       type(c_ptr), value :: s
       integer(c_size_t) :: c_strlen
     end function c_strlen
  end interface

contains

  ! stop with the library's message when a call fails (the reference stops
  ! with its own message at the same points)
  subroutine ysx_check(ctx, rc)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: n, i
    if (rc == 0) return
    n = int(c_strlen(ysx_last_error(ctx)))
    call c_f_pointer(ysx_last_error(ctx), msg, [n])
    text = ' '
    do i = 1, min(n, len(text))
       text(i:i) = msg(i)
    end do
    write(0, '(a)') trim(text)
    ! input the reference rejects (YSX_ERR_INPUT): it prints its stop text
    ! and ends with exit status 0 (src/yspec.f90:201-215); errors that have
    ! no counterpart there (HIP, arguments, unsupported runs) end with 1
    if (rc == 2) stop
    error stop 1
  end subroutine ysx_check

end module ysx_bind
