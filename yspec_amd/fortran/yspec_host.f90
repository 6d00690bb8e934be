!-----------------------------------------------------------------------
! yspec -- drop-in host program for the MI355X spectral path.
!
!   yspec <parameter file>
!
! Same command line, parameter file, deck format and per-receiver ASCII
! output as the reference program (src/yspec.f90): the parameter file is
! read with the same skip counts and list-directed reads
! (src/yspec.f90:79-137, src/module_util.f90:301-352), the deck with the
! same header and column layout (src/module_model.f90:194-222), and every
! receiver k gets the file <prefix>.<k> with rows  t, Z, N, E  in format
! 5e15.6 (src/yspec.f90:479-503).  All numerical work -- model transform,
! the (l, w) solves, the receiver sum, the filtered inverse FFT -- is done
! by libysx.so through ysx_bind.
!-----------------------------------------------------------------------
program yspec
  use iso_c_binding
  use ysx_bind
  implicit none

  integer, parameter :: dp = kind(1.0d0)
  integer, parameter :: iop = 7, iom = 8
  character(len=256) :: param_file, prefix, model_file, title, fname
  character(len=16) :: num
  logical :: there
  integer :: ios, i, k, nknot, nic, noc, ians, nr, ncol, nout, nt
  integer(c_int) :: ip(7), iout(8), rc
  real(c_double) :: fp(18), fout(8), tref
  real(c_double), allocatable :: deck(:,:), lat(:), lon(:), seis(:,:,:)
  real(dp) :: t, t_norm, vel_norm, dtn
  type(c_ptr) :: ctx

  call get_command_argument(1, param_file)
  inquire(file=trim(param_file), exist=there)
  if (.not. there) stop 'can''t find parameter file'
  open(iop, file=param_file, form='formatted', action='read', iostat=ios)
  if (ios /= 0) stop 'problem opening parameter file'

  call skip(iop, 4); read(iop, *) prefix
  call skip(iop, 2); read(iop, *) model_file
  ! ats grav out cor lmin lmax
  do i = 3, 6
     call skip(iop, 2); read(iop, *) ip(i)
  end do
  call skip(iop, 2); read(iop, *) ip(1)
  call skip(iop, 2); read(iop, *) ip(2)
  ! f1 f2 tout dt f11 f12 f21 f22 depth lat lon Mrr..Mpp receiver depth
  do i = 1, 18
     call skip(iop, 2); read(iop, *) fp(i)
  end do
  call skip(iop, 2); read(iop, *) nr
  call skip(iop, 2)
  allocate(lat(nr), lon(nr))
  do k = 1, nr
     read(iop, *) lat(k), lon(k)
  end do
  close(iop)
  ip(7) = nr

  ! deck model, raw columns as the file holds them
  open(iom, file=model_file, form='formatted', action='read', status='old', iostat=ios)
  if (ios /= 0) stop ' problem opening model file'
  read(iom, *) title
  read(iom, *) ians, tref
  read(iom, *) nknot, nic, noc
  allocate(deck(9, nknot))
  deck = 0.0_dp
  ncol = 6
  if (ians == 1) ncol = 9
  do i = 1, nknot
     read(iom, *) deck(1:ncol, i)
  end do
  close(iom)

  rc = ysx_create(ctx, 0_c_int)
  call ysx_check(ctx, rc)
  call ysx_check(ctx, ysx_set_model(ctx, nknot, nic, noc, ians, tref, deck))
  call ysx_check(ctx, ysx_set_run(ctx, ip, fp, lat, lon))
  call ysx_check(ctx, ysx_run_info(ctx, iout, fout))
  nt = iout(2)
  nout = iout(7)
  write(6, '(" yspec: ",i9," (l,w) pairs, ",i6," receivers, nt = ",i7)') iout(8), nr, nt

  allocate(seis(nt, 3, nr))
  call ysx_check(ctx, ysx_forward(ctx, seis))
  call ysx_destroy(ctx)

  ! time axis in seconds: t = (i-1)*dt in normalised units times t_norm,
  ! with t_norm formed as the reference forms it (src/module_model.f90:482-485;
  ! pi is the single-precision literal widened to double, src/nrtype.f90:20)
  vel_norm = 6.371e+6_dp*sqrt(real(3.141592653589793238462643383279502884197_4, dp) &
       *6.6723e-11_dp*5.515e+3_dp)
  t_norm = 1/(vel_norm/6.371e+6_dp)
  dtn = fout(3)
  do k = 1, nr
     write(num, '(i0)') k
     fname = trim(prefix)//'.'//trim(num)
     open(iop, file=trim(fname), form='formatted')
     do i = 1, nout
        t = (i-1)*dtn
        t = t*t_norm
        write(iop, 100) t, seis(i,1,k), seis(i,2,k), seis(i,3,k)
     end do
     close(iop)
  end do
100 format(5e15.6)

contains

  subroutine skip(io, n)
    integer, intent(in) :: io, n
    integer :: j
    do j = 1, n
       read(io, *)
    end do
  end subroutine skip

end program yspec
