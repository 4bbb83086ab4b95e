!-------------------------------------------------------------------------------
! yk_bind.f90 -- iso_c_binding interface to libyukawa_hip.so (include/yk.h).
!
! A Fortran host keeps control of the run and calls the MI355X pair-force /
! velocity-Verlet kernels through these interfaces.  The derived types are
! interoperable copies of yk_params / yk_meas / yk_timing; logicals cross the
! boundary as integer(c_int32_t); arrays are the reference's (3,n) layout.
!
! Entry points and the reference routines they replace:
!   yk_accel   accel(do_measurements)   src/int_control.f:52-81
!   yk_newton  newton(do_measurements)  src/vel_verlet_2.f:18-120
!   yk_run     nsteps of newton with the md.F measurement cadence (src/md.F:129)
!   yk_tnorm, yk_cancel_vcom  tnorm, cancel_vcom  src/gen_utils.f:61-107
!   yk_strain  strain                   src/strain.f:19-82
!   yk_run_md  the measurement-phase step loop of src/md.F:124-133
!   yk_batch_* many independent ion systems advanced together (no reference
!              counterpart: the reference runs one system per process)
!-------------------------------------------------------------------------------
module yk_bind
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: YK_PURE_ION = 0, YK_ION_MIXTURE = 1, &
                                           YK_NUCLEON = 2, YK_BFIELD_TEST = 3
  integer(c_int32_t), parameter, public :: YK_FP64 = 0, YK_MIXED = 1
  integer(c_int32_t), parameter, public :: YK_FLAG_REF_F1 = 1, YK_FLAG_KDK = 2, &
                                           YK_FLAG_TIMING = 4, YK_FLAG_ALLPAIRS = 8, &
                                           YK_FLAG_SW = 16, YK_FLAG_TILE_LISTS = 32
  integer(c_int), parameter, public :: YK_OK = 0
  real(c_double), parameter, public :: YK_XNOSTRAIN = 1000.0d0

  type, bind(C), public :: yk_params_t
    integer(c_int32_t) :: sim_type = YK_PURE_ION
    integer(c_int32_t) :: precision = YK_FP64
    integer(c_int32_t) :: flags = 0
    integer(c_int32_t) :: coulomb_on = 1
    integer(c_int32_t) :: nuclear_on = 1
    integer(c_int32_t) :: resort_every = 0
    real(c_double) :: xl(3) = 0.0_c_double
    real(c_double) :: xmuc = 0.0_c_double
    real(c_double) :: rccut = 1.0e50_c_double
    real(c_double) :: frp = 1.0_c_double
    real(c_double) :: vc = 197.327_c_double / 137.036_c_double
    real(c_double) :: xmass = 931.0_c_double
    real(c_double) :: rho = 0.0_c_double
    real(c_double) :: kT = 0.0_c_double
    real(c_double) :: zi = 0.0_c_double
    real(c_double) :: ai = 0.0_c_double
    real(c_double) :: aa = 110.0_c_double
    real(c_double) :: bb = -26.0_c_double
    real(c_double) :: cc = 24.0_c_double
    real(c_double) :: xpacket = 1.25_c_double
    real(c_double) :: rncut = 0.0_c_double
    real(c_double) :: bfield = 0.0_c_double
  end type yk_params_t

  type, bind(C), public :: yk_meas_t
    real(c_double) :: ev = 0, ek = 0, px = 0, vir = 0
    real(c_double) :: pp(6) = 0
  end type yk_meas_t

  type, bind(C), public :: yk_timing_t
    real(c_double) :: force_ms = 0
    integer(c_int64_t) :: force_launches = 0
    real(c_double) :: step_ms = 0
    integer(c_int64_t) :: steps = 0
    integer(c_int64_t) :: pairs_swept = 0
  end type yk_timing_t

  type, bind(C), public :: yk_loop_t
    real(c_double) :: dt = 0
    integer(c_int32_t) :: nsteps = 0, istep0 = 0, nind = 0, nptensor = 0, ncom = 0, ntnorm = 0
    real(c_double) :: deps(3) = YK_XNOSTRAIN
    real(c_double) :: xl0(3) = 0
  end type yk_loop_t

  public :: yk_abi_version, yk_create, yk_destroy, yk_set_params, yk_put_state, &
            yk_put_accel, yk_get_state, yk_accel, yk_newton, yk_run, yk_ttot, &
            yk_tnorm, yk_cancel_vcom, yk_strain, yk_run_md, &
            yk_get_timing, yk_reset_timing, yk_accel_noalloc, yk_error_message
  public :: yk_batch_create, yk_batch_destroy, yk_batch_footprint, yk_batch_add, &
            yk_batch_remove, yk_batch_set_tuning, yk_batch_set_resort, yk_batch_run, &
            yk_batch_status, yk_batch_get, yk_batch_last_run, yk_batch_error_message

  interface
    integer(c_int32_t) function yk_abi_version() bind(C, name='yk_abi_version')
      import :: c_int32_t
    end function

    type(c_ptr) function yk_last_error_c() bind(C, name='yk_last_error')
      import :: c_ptr
    end function

    integer(c_int) function yk_create(params, n, zii, aii, sys) bind(C, name='yk_create')
      import :: c_int, c_int32_t, c_double, c_ptr, yk_params_t
      type(yk_params_t), intent(in) :: params
      integer(c_int32_t), value :: n
      real(c_double), intent(in) :: zii(*), aii(*)
      type(c_ptr), intent(out) :: sys
    end function

    subroutine yk_destroy(sys) bind(C, name='yk_destroy')
      import :: c_ptr
      type(c_ptr), value :: sys
    end subroutine

    integer(c_int) function yk_set_params(sys, params) bind(C, name='yk_set_params')
      import :: c_int, c_ptr, yk_params_t
      type(c_ptr), value :: sys
      type(yk_params_t), intent(in) :: params
    end function

    integer(c_int) function yk_put_state(sys, x, v) bind(C, name='yk_put_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sys
      real(c_double), intent(in) :: x(3, *), v(3, *)
    end function

    integer(c_int) function yk_put_accel(sys, a) bind(C, name='yk_put_accel')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sys
      real(c_double), intent(in) :: a(3, *)
    end function

    integer(c_int) function yk_get_state(sys, x, v, a) bind(C, name='yk_get_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sys
      real(c_double), intent(out) :: x(3, *), v(3, *), a(3, *)
    end function

    integer(c_int) function yk_accel(sys, do_measurements, meas) bind(C, name='yk_accel')
      import :: c_int, c_int32_t, c_ptr, yk_meas_t
      type(c_ptr), value :: sys
      integer(c_int32_t), value :: do_measurements
      type(yk_meas_t), intent(out) :: meas
    end function

    integer(c_int) function yk_newton(sys, dt, do_measurements, meas) bind(C, name='yk_newton')
      import :: c_int, c_int32_t, c_ptr, c_double, yk_meas_t
      type(c_ptr), value :: sys
      real(c_double), value :: dt
      integer(c_int32_t), value :: do_measurements
      type(yk_meas_t), intent(out) :: meas
    end function

    integer(c_int) function yk_run(sys, dt, nsteps, nind, records) bind(C, name='yk_run')
      import :: c_int, c_int32_t, c_ptr, c_double, yk_meas_t
      type(c_ptr), value :: sys
      real(c_double), value :: dt
      integer(c_int32_t), value :: nsteps, nind
      type(yk_meas_t), intent(inout) :: records(*)
    end function

    integer(c_int) function yk_ttot(sys, eka) bind(C, name='yk_ttot')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sys
      real(c_double), intent(out) :: eka
    end function

    integer(c_int) function yk_tnorm(sys) bind(C, name='yk_tnorm')
      import :: c_int, c_ptr
      type(c_ptr), value :: sys
    end function

    integer(c_int) function yk_cancel_vcom(sys) bind(C, name='yk_cancel_vcom')
      import :: c_int, c_ptr
      type(c_ptr), value :: sys
    end function

    integer(c_int) function yk_strain(sys, dt, deps, xl0, xl_out) bind(C, name='yk_strain')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sys
      real(c_double), value :: dt
      real(c_double), intent(in) :: deps(3), xl0(3)
      real(c_double), intent(out) :: xl_out(3)
    end function

    integer(c_int) function yk_run_md(sys, loop, records, rec_steps, max_records, nrec, &
                                      xl_out) bind(C, name='yk_run_md')
      import :: c_int, c_int32_t, c_ptr, c_double, yk_loop_t, yk_meas_t
      type(c_ptr), value :: sys
      type(yk_loop_t), intent(in) :: loop
      type(yk_meas_t), intent(inout) :: records(*)
      integer(c_int32_t), intent(inout) :: rec_steps(*)
      integer(c_int32_t), value :: max_records
      integer(c_int32_t), intent(out) :: nrec
      real(c_double), intent(out) :: xl_out(3)
    end function

    integer(c_int) function yk_get_timing(sys, t) bind(C, name='yk_get_timing')
      import :: c_int, c_ptr, yk_timing_t
      type(c_ptr), value :: sys
      type(yk_timing_t), intent(out) :: t
    end function

    integer(c_int) function yk_reset_timing(sys) bind(C, name='yk_reset_timing')
      import :: c_int, c_ptr
      type(c_ptr), value :: sys
    end function

    ! accel on arrays already in device memory, with no allocation (no
    ! reference counterpart; for callers that cannot allocate): x, zii, aii, a
    ! and meas are device addresses (zii/aii may be c_null_ptr for
    ! 'pure-ion'), stream a hipStream_t or c_null_ptr; returns before the
    ! device work is done.
    integer(c_int) function yk_accel_noalloc(params, n, x, zii, aii, a, do_measurements, &
                                             meas, stream) bind(C, name='yk_accel_noalloc')
      import :: c_int, c_int32_t, c_ptr, yk_params_t
      type(yk_params_t), intent(in) :: params
      integer(c_int32_t), value :: n, do_measurements
      type(c_ptr), value :: x, zii, aii, a, meas, stream
    end function

    type(c_ptr) function yk_batch_last_error_c() bind(C, name='yk_batch_last_error')
      import :: c_ptr
    end function

    integer(c_int) function yk_batch_create(capacity, max_replicas, max_records, b) &
        bind(C, name='yk_batch_create')
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      integer(c_int64_t), value :: capacity
      integer(c_int32_t), value :: max_replicas, max_records
      type(c_ptr), intent(out) :: b
    end function

    subroutine yk_batch_destroy(b) bind(C, name='yk_batch_destroy')
      import :: c_ptr
      type(c_ptr), value :: b
    end subroutine

    integer(c_int) function yk_batch_footprint(b, n, slots, bytes) &
        bind(C, name='yk_batch_footprint')
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: b
      integer(c_int32_t), value :: n
      integer(c_int64_t), intent(out) :: slots, bytes
    end function

    integer(c_int) function yk_batch_add(b, params, n, zii, aii, x, v, dt, steps, id) &
        bind(C, name='yk_batch_add')
      import :: c_int, c_int32_t, c_double, c_ptr, yk_params_t
      type(c_ptr), value :: b
      type(yk_params_t), intent(in) :: params
      integer(c_int32_t), value :: n, steps
      real(c_double), intent(in) :: zii(*), aii(*), x(*), v(*)
      real(c_double), value :: dt
      integer(c_int32_t), intent(out) :: id
    end function

    integer(c_int) function yk_batch_remove(b, id) bind(C, name='yk_batch_remove')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: b
      integer(c_int32_t), value :: id
    end function

    integer(c_int) function yk_batch_set_tuning(b, nbuckets, n_upper, chunk_tiles, waves, &
                                                list_items) bind(C, name='yk_batch_set_tuning')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: b
      integer(c_int32_t), value :: nbuckets
      integer(c_int32_t), intent(in) :: n_upper(*), chunk_tiles(*), waves(*), list_items(*)
    end function

    integer(c_int) function yk_batch_set_resort(b, every) bind(C, name='yk_batch_set_resort')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: b
      integer(c_int32_t), value :: every
    end function

    integer(c_int) function yk_batch_run(b, nsteps, nind) bind(C, name='yk_batch_run')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: b
      integer(c_int32_t), value :: nsteps, nind
    end function

    integer(c_int) function yk_batch_status(b, id, steps_left, nrec) &
        bind(C, name='yk_batch_status')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: b
      integer(c_int32_t), value :: id
      integer(c_int32_t), intent(out) :: steps_left, nrec
    end function

    integer(c_int) function yk_batch_get(b, id, x, v, a, init, records, rec_cap, nrec) &
        bind(C, name='yk_batch_get')
      import :: c_int, c_int32_t, c_double, c_ptr, yk_meas_t
      type(c_ptr), value :: b
      integer(c_int32_t), value :: id, rec_cap
      real(c_double), intent(inout) :: x(*), v(*), a(*)
      type(yk_meas_t), intent(inout) :: init, records(*)
      integer(c_int32_t), intent(out) :: nrec
    end function

    integer(c_int) function yk_batch_last_run(b, ms, steps, items) &
        bind(C, name='yk_batch_last_run')
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: b
      real(c_double), intent(out) :: ms
      integer(c_int64_t), intent(out) :: steps, items
    end function

    integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! yk_last_error() as a Fortran string.
  function yk_error_message() result(msg)
    character(len=:), allocatable :: msg
    msg = c_string(yk_last_error_c())
  end function yk_error_message

  ! yk_batch_last_error() as a Fortran string.
  function yk_batch_error_message() result(msg)
    character(len=:), allocatable :: msg
    msg = c_string(yk_batch_last_error_c())
  end function yk_batch_error_message

  function c_string(p) result(msg)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: chars(:)
    integer :: k, m
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    m = int(c_strlen(p))
    call c_f_pointer(p, chars, [m])
    allocate(character(len=m) :: msg)
    do k = 1, m
      msg(k:k) = chars(k)
    end do
  end function c_string

end module yk_bind
