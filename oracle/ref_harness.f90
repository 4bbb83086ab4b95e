!-------------------------------------------------------------------------------
! ref_harness.f90 -- C-callable wrapper around the reference's own Fortran.
!
! TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
! the reference sources where they lie (/root/reference/src, never copied)
! into oracle/_ref/libykref.so.  It sets the module globals the reference
! reads (src/md_globals.f:36-172), then calls the reference's own
! accel (src/int_control.f:52) and newton (src/vel_verlet_2.f:18), so tests
! can compare the CPU restatement in yk_oracle.c, bit for bit, against the
! real thing on identical inputs.
!
! Array arguments use the reference's (3,n) column-major layout.
!-------------------------------------------------------------------------------
module ykref_harness
  use iso_c_binding
  use md_types
  use md_globals
  use md_comm
  implicit none
  include 'perf.h'
  logical, save :: timers_ready = .false.
contains

  ! kind: 0 pure-ion, 1 ion-mixture, 2 nucleon, 3 B-field (test)
  ! prm(1:17) = xl(3), xmuc, rccut, frp, xmass, rho, kT, zi, ai,
  !             aa, bb, cc, xpacket, rncut, bfield
  integer(c_int) function ykref_setup(kind, nn, coulomb_on, nuclear_on, prm, &
                                      xin, vin, zin, ain) bind(C, name='ykref_setup')
    integer(c_int), value :: kind, nn, coulomb_on, nuclear_on
    real(c_double), intent(in) :: prm(17)
    real(c_double), intent(in) :: xin(3, nn), vin(3, nn), zin(nn), ain(nn)
    if (.not. timers_ready) then
      call perf_init()
      timers_ready = .true.
    end if
    if (allocated(x)) deallocate(x)
    if (allocated(v)) deallocate(v)
    if (allocated(vold)) deallocate(vold)
    if (allocated(a)) deallocate(a)
    if (allocated(zii)) deallocate(zii)
    if (allocated(aii)) deallocate(aii)
    n = nn
    allocate(x(3, 0:n-1), v(3, 0:n-1), vold(3, 0:n-1), a(3, 0:n+2))
    allocate(zii(0:n-1), aii(0:n-1))
    x(:, 0:n-1) = xin
    v(:, 0:n-1) = vin
    vold = 0.0d0
    a = 0.0d0
    zii(0:n-1) = zin
    aii(0:n-1) = ain
    select case (kind)
    case (0)
      sim_type = 'pure-ion'
    case (1)
      sim_type = 'ion-mixture'
    case (2)
      sim_type = 'nucleon'
    case (3)
      sim_type = 'B-field'
    case default
      ykref_setup = -1
      return
    end select
    if (coulomb_on /= 0) then
      coulomb = 'screened-coulomb'
    else
      coulomb = 'none'
    end if
    if (nuclear_on /= 0) then
      nuclear = 'HPP'
    else
      nuclear = 'none'
    end if
    xl = prm(1:3)
    xmuc = prm(4)
    rccut = prm(5)
    frp = prm(6)
    xmass = prm(7)
    rho = prm(8)
    kT = prm(9)
    zi = prm(10)
    ai = prm(11)
    aa = prm(12)
    bb = prm(13)
    cc = prm(14)
    xpacket = prm(15)
    rncut = prm(16)
    bfield = prm(17)
    myrank = 0
    nprocs = 1
    time = 0.0d0
    ykref_setup = 0
  end function ykref_setup

  ! meas(1:10) = ev, vir(px-derived), px, pp11, pp12, pp13, pp22, pp23, pp33, ek
  subroutine pack_meas(meas)
    real(c_double), intent(out) :: meas(10)
    meas(1) = ev
    meas(2) = 0.0d0   ! vir is local to the reference's accel routines
    meas(3) = px
    meas(4) = pp(1, 1)
    meas(5) = pp(1, 2)
    meas(6) = pp(1, 3)
    meas(7) = pp(2, 2)
    meas(8) = pp(2, 3)
    meas(9) = pp(3, 3)
    meas(10) = ek
  end subroutine pack_meas

  subroutine ykref_accel(do_meas, aout, meas) bind(C, name='ykref_accel')
    integer(c_int), value :: do_meas
    real(c_double), intent(out) :: aout(3, n + 3)
    real(c_double), intent(out) :: meas(10)
    logical :: dm
    dm = (do_meas /= 0)
    pp = 0.0d0
    ev = 0.0d0
    px = 0.0d0
    ek = 0.0d0
    call accel(dm)
    aout = a(:, 0:n+2)
    call pack_meas(meas)
  end subroutine ykref_accel

  ! nsteps vel_verlet_2 steps with constant dt; measure on the last step if
  ! meas_last /= 0.  Positions/velocities/accelerations stay in the globals.
  subroutine ykref_newton(nsteps, dtin, meas_last, meas) bind(C, name='ykref_newton')
    integer(c_int), value :: nsteps, meas_last
    real(c_double), value :: dtin
    real(c_double), intent(out) :: meas(10)
    integer :: s
    logical :: dm
    dt = dtin
    do s = 1, nsteps
      dm = (meas_last /= 0) .and. (s == nsteps)
      call newton(dm)
    end do
    call pack_meas(meas)
  end subroutine ykref_newton

  ! nsteps of the reference's alternate integrators, vel_verlet_1 (which = 1)
  ! or vel_verlet_3 (which = 3).  Their sources are compiled with newton
  ! renamed to ykref_newton_vv1 / ykref_newton_vv3 (both define a newton
  ! without arguments, which would clash with vel_verlet_2's) and their
  ! argument-less call of accel renamed to ykref_accel_measuring below.
  subroutine ykref_newton_alt(which, nsteps, dtin, meas) bind(C, name='ykref_newton_alt')
    integer(c_int), value :: which, nsteps
    real(c_double), value :: dtin
    real(c_double), intent(out) :: meas(10)
    integer :: s
    external :: ykref_newton_vv1, ykref_newton_vv3
    dt = dtin
    do s = 1, nsteps
      pp = 0.0d0
      if (which == 1) then
        call ykref_newton_vv1()
      else
        call ykref_newton_vv3()
      end if
    end do
    call pack_meas(meas)
  end subroutine ykref_newton_alt

  subroutine ykref_set_accel(ain_) bind(C, name='ykref_set_accel')
    real(c_double), intent(in) :: ain_(3, n)
    a(:, 0:n-1) = ain_
  end subroutine ykref_set_accel

  subroutine ykref_get(xout, vout, aout) bind(C, name='ykref_get')
    real(c_double), intent(out) :: xout(3, n), vout(3, n), aout(3, n)
    xout = x(:, 0:n-1)
    vout = v(:, 0:n-1)
    aout = a(:, 0:n-1)
  end subroutine ykref_get

  real(c_double) function ykref_ttot() bind(C, name='ykref_ttot')
    real(dble) :: eka
    call ttot(eka)
    ykref_ttot = eka
  end function ykref_ttot

  ! The reference's velocity utilities on the globals (src/gen_utils.f:61-107);
  ! kT is the value ykref_setup set.
  subroutine ykref_tnorm() bind(C, name='ykref_tnorm')
    call tnorm()
  end subroutine ykref_tnorm

  subroutine ykref_cancel_vcom() bind(C, name='ykref_cancel_vcom')
    call cancel_vcom()
  end subroutine ykref_cancel_vcom

  ! The reference's strain (src/strain.f:19-82) with these rates, gage
  ! dimensions and timestep; returns the new box.
  subroutine ykref_strain(deps_in, xl0_in, dtin, xlout) bind(C, name='ykref_strain')
    real(c_double), intent(in) :: deps_in(3), xl0_in(3)
    real(c_double), value :: dtin
    real(c_double), intent(out) :: xlout(3)
    deps = deps_in
    xl0 = xl0_in
    dt = dtin
    call strain()
    xlout = xl
  end subroutine ykref_strain

  ! The measurement-phase step loop of src/md.F:124-133 on the globals: per
  ! step cancel_vcom (every ncom), tnorm (every ntnorm), strain, then newton
  ! measuring when mdmod(istep+1,nind) or mdmod(istep+1,nptensor) is 0.  One
  ! record per measuring step, as the .meas record of src/md.F:150:
  ! (istep, ev, ek, px, pp11, pp12, pp13, pp22, pp23, pp33).  xlout: the box
  ! after the loop.
  subroutine ykref_md_loop(nsteps, istep0, nind_, nptensor_, ncom_, ntnorm_, dtin, deps_in, &
                           xl0_in, maxrec, rec, nrec, xlout) bind(C, name='ykref_md_loop')
    integer(c_int), value :: nsteps, istep0, nind_, nptensor_, ncom_, ntnorm_, maxrec
    real(c_double), value :: dtin
    real(c_double), intent(in) :: deps_in(3), xl0_in(3)
    real(c_double), intent(out) :: rec(10, maxrec)
    integer(c_int), intent(out) :: nrec
    real(c_double), intent(out) :: xlout(3)
    integer :: i, istep
    logical :: dm
    dt = dtin
    deps = deps_in
    xl0 = xl0_in
    istep = istep0
    nrec = 0
    do i = 1, nsteps
      if (ymod(istep, ncom_) == 0) call cancel_vcom()
      if (ymod(istep, ntnorm_) == 0) call tnorm()
      call strain()
      dm = ymod(istep + 1, nind_) == 0 .or. ymod(istep + 1, nptensor_) == 0
      call newton(dm)
      istep = istep + 1
      if (dm .and. nrec < maxrec) then
        nrec = nrec + 1
        rec(1, nrec) = real(istep, c_double)
        rec(2, nrec) = ev
        rec(3, nrec) = ek
        rec(4, nrec) = px
        rec(5, nrec) = pp(1, 1)
        rec(6, nrec) = pp(1, 2)
        rec(7, nrec) = pp(1, 3)
        rec(8, nrec) = pp(2, 2)
        rec(9, nrec) = pp(2, 3)
        rec(10, nrec) = pp(3, 3)
      end if
    end do
    xlout = xl
  end subroutine ykref_md_loop

  ! mdmod of src/md.F:248-256 (an internal function of the main program
  ! there): -1 for a zero period, so that step never comes.
  integer function ymod(p, q)
    integer, intent(in) :: p, q
    if (q == 0) then
      ymod = -1
    else
      ymod = mod(p, q)
    end if
  end function ymod

  ! Reference timers (src/perf.h): pair-loop time and count, accel, newton.
  subroutine ykref_timers(t) bind(C, name='ykref_timers')
    real(c_double), intent(out) :: t(6)
    t(1) = t_calc_a
    t(2) = real(n_calc_a, c_double)
    t(3) = t_accel
    t(4) = real(n_accel, c_double)
    t(5) = t_newton
    t(6) = real(n_newton, c_double)
  end subroutine ykref_timers

  subroutine ykref_reset_timers() bind(C, name='ykref_reset_timers')
    t_calc_a = 0.0d0; ts_calc_a = 0.0d0; n_calc_a = 0
    t_accel = 0.0d0; ts_accel = 0.0d0; n_accel = 0
    t_newton = 0.0d0; ts_newton = 0.0d0; n_newton = 0
  end subroutine ykref_reset_timers

end module ykref_harness


! The accel the alternate integrators call (they pass no argument): the
! reference's accel with measurements on, so ev, px and pp are formed.
subroutine ykref_accel_measuring()
  implicit none
  call accel(.true.)
end subroutine ykref_accel_measuring
