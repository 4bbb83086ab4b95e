!-------------------------------------------------------------------------------
! md_yk_shim.f90 -- drop-in accel/newton for the reference's own Fortran host.
!
! Linked in place of the reference's int_control.o and vel_verlet_2.o (the
! INTOBJGPC / INTEGRATOR plug point, src/Makefile:33-34), these two external
! subroutines keep the reference's interface and side effects:
!
!   accel(do_measurements)    src/int_control.f:52-81: writes a, ev, px, pp
!   newton(do_measurements)   src/vel_verlet_2.f:18-120: writes x, v, vold, a,
!                             ek, pp and advances time
!
! and do the work on the GPU through libyukawa_hip.so (yk_bind).  Host code of
! the reference (cancel_vcom, tnorm, strain, checkpoint) may touch x and v
! between calls, so every call uploads x, v and a and downloads the results:
! exact drop-in semantics.  A host that keeps the state on the device calls
! yk_run instead.  The mixture kernel's last-particle behaviour
! (src/int_ion_mix.f:158, :201, :250) is reproduced unless YK_FIX_F1=1 is set
! in the environment; YK_PRECISION=mixed selects the mixed-precision kernel and
! YK_ALLPAIRS=1 the all-pairs tile kernel instead of the pair lists.
!
! Compiles against the host's md_types / md_globals / md_comm modules and
! perf.h; errors stop the run through md_exit (src/md_utils.f:158-183).
!-------------------------------------------------------------------------------
module md_yk_shim_state
  use iso_c_binding
  use yk_bind
  implicit none
  type(c_ptr), save :: handle = c_null_ptr
  integer, save :: handle_n = -1
  character(len=20), save :: handle_type = ''
contains

  subroutine shim_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= YK_OK) then
      write(6, '(a,a,i3,2a)') 'yk shim: ', where, rc, ': ', yk_error_message()
      call md_exit(1, '*** GPU pair-force library error ***')
    end if
  end subroutine shim_check

  subroutine shim_params(p)
    use md_globals
    type(yk_params_t), intent(out) :: p
    character(len=16) :: env
    integer :: ln, st
    select case (trim(sim_type))
    case ('pure-ion')
      p%sim_type = YK_PURE_ION
    case ('ion-mixture')
      p%sim_type = YK_ION_MIXTURE
    case ('nucleon')
      p%sim_type = YK_NUCLEON
    case ('B-field')
      p%sim_type = YK_BFIELD_TEST
    case default
      write(6, "('sim_type=***',a,'***')") trim(sim_type)
      call md_exit(1, '*** Program not compiled to do this simulation type ***')
    end select
    p%precision = YK_FP64
    call get_environment_variable('YK_PRECISION', env, ln, st)
    if (st == 0 .and. env(1:5) == 'mixed') p%precision = YK_MIXED
    p%flags = YK_FLAG_REF_F1
    call get_environment_variable('YK_FIX_F1', env, ln, st)
    if (st == 0 .and. env(1:1) == '1') p%flags = 0
    call get_environment_variable('YK_ALLPAIRS', env, ln, st)
    if (st == 0 .and. env(1:1) == '1') p%flags = ior(p%flags, YK_FLAG_ALLPAIRS)
    p%coulomb_on = merge(1, 0, coulomb == 'screened-coulomb')
    p%nuclear_on = merge(1, 0, nuclear == 'HPP')
    p%resort_every = 0
    p%xl = xl
    p%xmuc = xmuc
    p%rccut = rccut
    p%frp = frp
    p%vc = vc
    p%xmass = xmass
    p%rho = rho
    p%kT = kT
    p%zi = zi
    p%ai = ai
    p%aa = aa
    p%bb = bb
    p%cc = cc
    p%xpacket = xpacket
    p%rncut = rncut
    p%bfield = bfield
  end subroutine shim_params

  ! Create the device system on first use (or when n / sim_type change), then
  ! refresh parameters (the box may have been strained) and upload x, v.
  subroutine shim_sync_in()
    use md_globals
    type(yk_params_t) :: p
    call shim_params(p)
    if (.not. c_associated(handle) .or. handle_n /= n .or. handle_type /= sim_type) then
      if (c_associated(handle)) call yk_destroy(handle)
      call shim_check(yk_create(p, int(n, c_int32_t), zii, aii, handle), 'yk_create')
      handle_n = n
      handle_type = sim_type
    else
      call shim_check(yk_set_params(handle, p), 'yk_set_params')
    end if
    call shim_check(yk_put_state(handle, x, v), 'yk_put_state')
  end subroutine shim_sync_in

  ! Scalars as the reference leaves them in md_globals (upper triangle of pp
  ! from accel; the full symmetric tensor over the volume from newton).
  subroutine shim_take(m, do_measurements, from_newton)
    use md_globals
    type(yk_meas_t), intent(in) :: m
    logical, intent(in) :: do_measurements, from_newton
    pp = 0.0d0
    if (.not. do_measurements) then
      ev = 0.0d0
      px = 0.0d0
      if (from_newton) ek = 0.0d0
      return
    end if
    ev = m%ev
    px = m%px
    pp(1, 1) = m%pp(1); pp(1, 2) = m%pp(2); pp(1, 3) = m%pp(3)
    pp(2, 2) = m%pp(4); pp(2, 3) = m%pp(5); pp(3, 3) = m%pp(6)
    if (from_newton) then
      ek = m%ek
      pp(2, 1) = pp(1, 2); pp(3, 1) = pp(1, 3); pp(3, 2) = pp(2, 3)
    end if
  end subroutine shim_take

end module md_yk_shim_state


!*******************************************************************************
subroutine accel(do_measurements)
  use md_types
  use md_globals
  use md_comm
  use md_yk_shim_state
  implicit none
  include 'perf.h'
  logical do_measurements
  type(yk_meas_t) :: m
  call starttimer()   ! accel
  call starttimer()   ! calc_a
  call shim_sync_in()
  call shim_check(yk_accel(handle, merge(1_c_int32_t, 0_c_int32_t, do_measurements), m), &
                  'yk_accel')
  call shim_check(yk_get_state(handle, x, v, a), 'yk_get_state')
  call stoptimer(1, t_calc_a, ts_calc_a, n_calc_a)
  call shim_take(m, do_measurements, .false.)
  call stoptimer(1, t_accel, ts_accel, n_accel)
end subroutine accel


!*******************************************************************************
subroutine newton(do_measurements)
  use md_types
  use md_globals
  use md_comm
  use md_yk_shim_state
  implicit none
  include 'perf.h'
  logical do_measurements
  type(yk_meas_t) :: m
  call starttimer()   ! newton
  vold(:, 0:n-1) = v(:, 0:n-1)
  call shim_sync_in()
  call shim_check(yk_put_accel(handle, a), 'yk_put_accel')
  call shim_check(yk_newton(handle, dt, merge(1_c_int32_t, 0_c_int32_t, do_measurements), m), &
                  'yk_newton')
  call shim_check(yk_get_state(handle, x, v, a), 'yk_get_state')
  call shim_take(m, do_measurements, .true.)
  time = time + dt
  call stoptimer(1, t_newton, ts_newton, n_newton)
end subroutine newton
