!=====================================================================!
! ref_dump -- FP64 dump driver around the UNMODIFIED reference yspec  !
! modules (compiled from the reference's src/*.f90 by oracle/Makefile !
! into oracle/_ref/).  Test infrastructure only: it produces golden   !
! vectors and the timed CPU baseline; it is never linked into the     !
! product.                                                            !
!                                                                     !
! The reference prints its seismograms with 6 significant digits      !
! (src/yspec.f90:503), which cannot carry 1e-9 parity, so this driver  !
! calls the reference's own routines in the order src/yspec.f90 uses  !
! them and writes every intermediate result as raw little-endian FP64  !
! stream files:                                                       !
!   header.bin  run scalars and normalisation constants               !
!   model.bin   module_model arrays after read_model                  !
!   geom.bin    delaz / legendre outputs  (src/yspec.f90:144-174)     !
!   pairs.bin   one record per (l,w): source vectors, raw solver      !
!               outputs and the uu/vv/ww coefficients                 !
!   spec.bin    ur/ut/up spectra before the inverse FFT               !
!   seis.bin    ur/ut/up after ifft, exp(eps t), units and rotation   !
!                                                                     !
! usage:  ref_dump <param file> <out dir>                             !
!         ref_dump <param file> --time l1 [l2 ...]                     !
! The second form runs only the listed degrees (all frequencies, all  !
! receivers), writes nothing and prints the wall time of each degree; !
! it is the timed CPU baseline.                                       !
!=====================================================================!
program ref_dump
  use nrtype
  use module_util
  use module_model
  use module_solver
  use module_start
  use module_int, only : set_l, set_omega, set_ats
  use module_function
  use module_fourier
  implicit none

  integer(i4b), parameter :: iop = 7, iom = 8, iow = 9
  character(len=256) :: param_file, out_dir, pref_out, model_file, arg
  logical(lgt) :: timing, there
  integer(i4b) :: ats_in, grav_switch, out_switch, cor_switch, lmin, lmax
  integer(i4b) :: nr, i, j, k, l, m, nw, nt, i1, i2, is, ir, nout, mex, qex
  integer(i4b) :: ntime, it, itor, isph, c0, c1, crate
  integer(i4b), allocatable :: ltime(:)
  real(dp) :: f1, f2, tout, dt, f11, f12, f21, f22, rs, rr, lats, lons
  real(dp) :: depth_src, depth_rec, df, ep, dw, wmin, zeta, t, x, xp, xc
  real(dp) :: mm(6), sgn
  real(dp), allocatable :: latr(:), lonr(:), delta(:), azep(:), azst(:)
  real(dp), allocatable :: xa(:,:,:), xpa(:,:,:), xca(:,:,:)
  complex(dpc) :: w, ei, tmp1, tmp2
  complex(dpc) :: uu(5), vv(5), ww(5), ctmp1(4), ctmp2(4)
  complex(dpc) :: svt(5,2), svs(5,6), sr0(2)
  complex(dpc) :: s2(2,2), s6(6,4), ytor(2,2), ysph(6,4), yrad(2,2), prad(2,2)
  complex(dpc) :: s4(4,4), ysphng(4,4)
  complex(dpc) :: ytor_raw(2,2), ysph_raw(6,4), yrad_raw(2,2)
  complex(dpc), allocatable :: eia(:), ur(:,:), ut(:,:), up(:,:)
  real(dp) :: tsec

  call set_parameters

  call get_command_argument(1, param_file)
  call get_command_argument(2, out_dir)
  timing = (trim(out_dir) == '--time')
  if (timing) then
     ntime = command_argument_count() - 2
     allocate(ltime(max(ntime,1)))
     do i = 1, ntime
        call get_command_argument(2+i, arg)
        read(arg,*) ltime(i)
     end do
  end if

  inquire(file=trim(param_file), exist=there)
  if (.not. there) stop 'ref_dump: no parameter file'

  ! parameter file, read with the reference's own skip-count readers
  open(iop, file=param_file, form='formatted', action='read')
  call read_string(iop, pref_out, 4)
  call read_string(iop, model_file, 2)
  open(iom, file=model_file, form='formatted', action='read')
  call read_model(iom)
  close(iom)
  call read_int(iop, ats_in, 2)
  call read_int(iop, grav_switch, 2)
  call read_int(iop, out_switch, 2)
  call read_int(iop, cor_switch, 2)
  call read_int(iop, lmin, 2)
  call read_int(iop, lmax, 2)
  call read_float(iop, f1, 2)
  call read_float(iop, f2, 2)
  call read_float(iop, tout, 2)
  call read_float(iop, dt, 2)
  call read_float(iop, f11, 2)
  call read_float(iop, f12, 2)
  call read_float(iop, f21, 2)
  call read_float(iop, f22, 2)
  call read_float(iop, rs, 2)
  depth_src = rs
  call read_float(iop, lats, 2)
  call read_float(iop, lons, 2)
  do i = 1, 6
     call read_float(iop, mm(i), 2)
  end do
  call read_float(iop, rr, 2)
  depth_rec = rr
  call read_int(iop, nr, 2)
  allocate(latr(nr), lonr(nr))
  read(iop,*)
  read(iop,*)
  do i = 1, nr
     read(iop,*) latr(i), lonr(i)
  end do
  close(iop)

  ! receiver geometry and Legendre tables with the reference routines
  allocate(delta(nr), azep(nr), azst(nr))
  if (lats == 90.0_dp) lats = 89.98_dp
  do k = 1, nr
     call delaz(latr(k), lonr(k), lats, lons, delta(k), azep(k), azst(k))
  end do
  azep = twopi_d - azep
  azst = pi_d - azst
  allocate(xa(lmax+1,3,nr), xpa(lmax+1,3,nr), xca(lmax+1,3,nr))
  xa = 0.0_dp; xpa = 0.0_dp; xca = 0.0_dp
  do k = 1, nr
     call legendre(delta(k), 0, 0, xa(1,1:1,k), xpa(1,1:1,k), xca(1,1:1,k))
     if (lmax >= 1) call legendre(delta(k), 1, 1, xa(2,1:2,k), xpa(2,1:2,k), xca(2,1:2,k))
     do l = 2, lmax
        call legendre(delta(k), l, 2, xa(l+1,:,k), xpa(l+1,:,k), xca(l+1,:,k))
     end do
  end do
  allocate(eia(nr))
  eia = exp(ii*azst)

  ! normalisation (same conversions as the reference driver)
  f1 = f1*t_norm/1000.0_dp
  f2 = f2*t_norm/1000.0_dp
  dt = dt/t_norm
  rs = r(nknot) - rs*1000.0_dp/r_norm
  rr = r(nknot) - rr*1000.0_dp/r_norm
  tout = tout*60.0_dp/t_norm
  mm = mm/moment_norm
  f11 = f11/1000.0_dp*t_norm
  f12 = f12/1000.0_dp*t_norm
  f21 = f21/1000.0_dp*t_norm
  f22 = f22/1000.0_dp*t_norm
  mex = 5
  qex = 1
  call fcal(f1, f2, dt, tout, mex, qex, df, ep, nt, i1, i2)
  dw = twopi_d*df
  wmin = (i1-1)*dw
  nw = i2 - i1 + 1
  is = find_radius_index(rs, -1)
  ir = find_radius_index(rr, -1)
  if (is == 0 .or. ir == 0) stop 'ref_dump: source or receiver outside the model'

  allocate(ur(nw,nr), ut(nw,nr), up(nw,nr))
  ur = 0.0_dp; ut = 0.0_dp; up = 0.0_dp

  s2 = 0.0_dp; s2(1,1) = 1.0_dp; s2(2,2) = 1.0_dp
  s6 = 0.0_dp; s6(1,1) = 1.0_dp; s6(2,2) = 1.0_dp; s6(4,3) = 1.0_dp; s6(5,4) = 1.0_dp
  s4 = 0.0_dp
  do k = 1, 4
     s4(k,k) = 1.0_dp
  end do
  call set_ats_switch(ats_in)

  if (.not. timing) then
     call write_header()
     call write_model()
     open(iow, file=trim(out_dir)//'/geom.bin', access='stream', form='unformatted', status='replace')
     write(iow) delta, azep, azst, xa, xpa, xca
     close(iow)
     open(iow, file=trim(out_dir)//'/pairs.bin', access='stream', form='unformatted', status='replace')
     do l = lmin, lmax
        do j = 1, nw
           call one_pair(l, j)
        end do
     end do
     close(iow)
     open(iow, file=trim(out_dir)//'/spec.bin', access='stream', form='unformatted', status='replace')
     write(iow) ur, ut, up
     close(iow)
     call finish_seismograms()
     open(iow, file=trim(out_dir)//'/seis.bin', access='stream', form='unformatted', status='replace')
     write(iow) nt, nout
     write(iow) ur, ut, up
     close(iow)
  else
     call system_clock(count_rate=crate)
     do it = 1, ntime
        l = ltime(it)
        call system_clock(c0)
        do j = 1, nw
           call one_pair(l, j)
        end do
        call system_clock(c1)
        tsec = real(c1-c0,dp)/real(crate,dp)
        write(6,'("l ",i6," nw ",i6," seconds ",es16.8)') l, nw, tsec
     end do
  end if

contains

  subroutine one_pair(l, j)
    ! One (l, w_j) pair exactly as the reference loop body evaluates it
    ! (src/yspec.f90:254-421): source vectors, solvers, output-type
    ! scaling, coefficients and the receiver sum.
    integer(i4b), intent(in) :: l, j
    zeta = sqrt(real(l*(l+1)))
    w = wmin + (j-1)*dw - ii*ep
    ytor_raw = 0.0_dp; ysph_raw = 0.0_dp; yrad_raw = 0.0_dp
    svt = 0.0_dp; svs = 0.0_dp; sr0 = 0.0_dp
    uu = 0.0_dp; vv = 0.0_dp; ww = 0.0_dp
    itor = -1; isph = -1
    if (l == 0) then
       call source_vector_rad(w, is, rs, mm, sr0)
       if (grav_switch == 0) then
          call rad_solver_ng(w, rr, rs, ir, is, s2, yrad)
       else
          call rad_solver_wg(w, rr, rs, ir, is, s2, yrad, prad)
       end if
       yrad_raw = yrad
       uu(3) = (sr0(1)*yrad(1,1) + sr0(2)*yrad(1,2))/rr
       if (out_switch == 1) uu(3) = ii*w*uu(3)
       if (out_switch == 2) uu(3) = -w*w*uu(3)
       do k = 1, nr
          ur(j,k) = ur(j,k) + uu(3)*xa(1,1,k)
       end do
    else
       if (l > 1) then
          call source_vector_tor(l, w, is, rs, mm, svt)
          call tor_solver(l, w, rr, rs, ir, is, s2, ytor)
          call tor_start_level(l, real(w), merge(is, ir, rr >= rs), 1, itor)
          ytor_raw = ytor
          if (out_switch == 1) ytor = ii*w*ytor
          if (out_switch == 2) ytor = -w*w*ytor
          ww = (svt(:,1)*ytor(1,1) + svt(:,2)*ytor(1,2))/(zeta*rr)
          do k = 1, nr
             do m = -2, 2
                if (m == 0) cycle
                call table(l, m, k, x, xp, xc)
                ei = eia(k)**m
                ut(j,k) = ut(j,k) + ii*m*ww(m+3)*x*ei
                up(j,k) = up(j,k) - ww(m+3)*xp*ei
             end do
          end do
       end if
       call source_vector_sph(l, w, is, rs, mm, svs)
       ! solver choice and the 4-vector -> 6-vector layout of the
       ! reference driver (src/yspec.f90:342-374)
       if (fic) then
          if (grav_switch == 0) then
             call sph_solver_ng_fic(l, w, rr, rs, ir, is, s4, ysphng)
          else if (grav_switch == 1) then
             call sph_solver_cow_fic(l, w, rr, rs, ir, is, s4, ysphng)
          else
             call sph_solver_fic(l, w, rr, rs, ir, is, s6, ysph)
          end if
       else
          if (grav_switch == 0) then
             call sph_solver_ng(l, w, rr, rs, ir, is, s4, ysphng)
          else if (grav_switch == 1) then
             call sph_solver_cow(l, w, rr, rs, ir, is, s4, ysphng)
          else
             call sph_solver(l, w, rr, rs, ir, is, s6, ysph)
          end if
       end if
       if (grav_switch /= 2) then
          ysph(1,:) = ysphng(1,:)
          ysph(2,:) = ysphng(2,:)
          ysph(3,:) = 0.0_dp
          ysph(4,:) = ysphng(3,:)
          ysph(5,:) = ysphng(4,:)
          ysph(6,:) = 0.0_dp
       end if
       call sph_start_level(l, real(w), merge(is, ir, rr >= rs), 1, isph)
       ysph_raw = ysph
       if (out_switch == 1) then
          ysph(1:3,:) = ii*w*ysph(1:3,:)
       else if (out_switch == 2) then
          if (cor_switch == 0) then
             ysph(1:3,:) = -w*w*ysph(1:3,:)
          else
             ctmp1 = -((l+1)*ysph(3,:) + 2.0_dp*grav(ir)*ysph(1,:))/rr
             ctmp2 = zeta*(ysph(3,:) + grav(ir)*ysph(1,:))/rr
             ysph(1,:) = -w*w*ysph(1,:) + ctmp1
             ysph(2,:) = -w*w*ysph(2,:) + ctmp2
          end if
       end if
       uu = (ysph(1,1)*svs(:,1) + ysph(1,2)*svs(:,2) + ysph(1,3)*svs(:,4) + ysph(1,4)*svs(:,5))/rr
       vv = (ysph(2,1)*svs(:,1) + ysph(2,2)*svs(:,2) + ysph(2,3)*svs(:,4) + ysph(2,4)*svs(:,5))/(zeta*rr)
       do k = 1, nr
          do m = -2, 2
             call table(l, m, k, x, xp, xc)
             ei = eia(k)**m
             ur(j,k) = ur(j,k) + uu(m+3)*x*ei
             ut(j,k) = ut(j,k) + vv(m+3)*xp*ei
             up(j,k) = up(j,k) + ii*m*vv(m+3)*xc*ei
          end do
       end do
    end if
    if (.not. timing) then
       write(iow) l, j, itor, isph
       write(iow) svt, ytor_raw, svs, ysph_raw, sr0, yrad_raw, uu, vv, ww
    end if
  end subroutine one_pair

  subroutine table(l, m, k, x, xp, xc)
    ! Legendre table entry for order m (negative orders take (-1)**m of
    ! the |m| entry, src/yspec.f90:321-330).
    integer(i4b), intent(in) :: l, m, k
    real(dp), intent(out) :: x, xp, xc
    if (m >= 0) then
       sgn = 1.0_dp
    else
       sgn = (-1)**m
    end if
    x  = sgn*xa(l+1,abs(m)+1,k)
    xp = sgn*xpa(l+1,abs(m)+1,k)
    xc = sgn*xca(l+1,abs(m)+1,k)
  end subroutine table

  subroutine finish_seismograms()
    ! inverse FFT with the reference routine, then the reference's
    ! exp(eps t), dimensionalisation and rotation (src/yspec.f90:432-486)
    call ifft(nt, i1, i2, df, f11, f12, f21, f22, dt, ur, ut, up)
    nt = size(ur,1)
    do i = 1, nt
       t = (i-1)*dt
       ur(i,:) = ur(i,:)*exp(ep*t)
       ut(i,:) = ut(i,:)*exp(ep*t)
       up(i,:) = up(i,:)*exp(ep*t)
    end do
    if (out_switch == 0) then
       ur = ur*r_norm; ut = ut*r_norm; up = up*r_norm
    else if (out_switch == 1) then
       ur = ur*vel_norm; ut = ut*vel_norm; up = up*vel_norm
    else
       ur = ur*acl_norm; ut = ut*acl_norm; up = up*acl_norm
    end if
    do i = 1, nt
       do k = 1, nr
          tmp1 = ut(i,k)
          tmp2 = up(i,k)
          ut(i,k) = -tmp1*cos(azep(k)) + tmp2*sin(azep(k))
          up(i,k) =  tmp1*sin(azep(k)) + tmp2*cos(azep(k))
       end do
    end do
    nout = nt
    do i = 1, nt
       t = (i-1)*dt
       if (t > tout) then
          nout = i
          exit
       end if
    end do
  end subroutine finish_seismograms

  subroutine write_header()
    integer(i4b) :: flags(8), run(16)
    real(dp) :: norms(9), runf(22)
    flags = (/ nknot, nic, noc, nsl, merge(1,0,ifanis), merge(1,0,ocean), merge(1,0,fic), nlayer /)
    norms = (/ tref, r_norm, rho_norm, vel_norm, acl_norm, fre_norm, t_norm, moment_norm, pibigg /)
    run = (/ lmin, lmax, nw, nt, i1, i2, is, ir, nr, ats_in, grav_switch, out_switch, cor_switch, 0, 0, 0 /)
    runf(1:16) = (/ rs, rr, dt, tout, ep, df, wmin, dw, f11, f12, f21, f22, f1, f2, depth_src, depth_rec /)
    runf(17:22) = mm
    open(iow, file=trim(out_dir)//'/header.bin', access='stream', form='unformatted', status='replace')
    write(iow) flags, norms, run, runf
    close(iow)
  end subroutine write_header

  subroutine write_model()
    open(iow, file=trim(out_dir)//'/model.bin', access='stream', form='unformatted', status='replace')
    write(iow) r, rho, rho_cs, acon, acon_cs, ccon, ccon_cs, fcon, fcon_cs, &
         lcon, lcon_cs, ncon, ncon_cs, kappa, kappa_cs, mu, mu_cs, &
         qkappa, qkappa_cs, qmu, qmu_cs, rat, rat_cs, xlam, xlam_cs, &
         xa2, xa2_cs, grav, grav_cs, lambda, beta, alpha
    write(iow) layer_index
    close(iow)
  end subroutine write_model

end program ref_dump
