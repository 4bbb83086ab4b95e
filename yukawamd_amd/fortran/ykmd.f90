!-------------------------------------------------------------------------------
! ykmd.f90 -- MD run driver over the GPU library, with the state resident on
! the device.
!
! Reads the reference's run-parameter file (runmd.in, namelist
! run_parameters with the names of src/input1.F:28-53, so the same file
! serves both programs), derives the run parameters as src/input1.F and
! src/initconfig.f do, takes the charges and masses from the spec_list, a
! species file ('ZAfa', 'ZAfb') or the start file, builds the initial state of
! a start='random' or start='nuclear' run with the reference's generator
! (irnd = 0 or 4: MT19937; yk_host.f90) or reads a start file of any type the
! reference reads (yk_config.f90), generating Maxwell velocities when it has
! none, and runs the schedule of src/md.F:66-205 on the GPU:
! initial accel and ttot, nwgroup warm-up groups (nwsteps steps, then
! cancel_vcom and tnorm), ngroup measurement groups of ntot*nind steps with
! the ncom / ntnorm / strain / nind / nptensor cadences (one yk_run_md call
! per group), block statistics, the final accel and ttot.
!
! Output, in the reference's formats: the log (md.<tend>.log: the parameter
! echo, the 'e', 'p' and 's' lines and the initial / final lines), the binary measurement
! records (md.<tend>.meas, src/md.F:151), the initial configuration
! (md.<tstart>.xv8b, src/initconfig.f:385-390), the configurations saved
! every nout steps of the measurement phase (save_config, src/md_utils.f:61-114:
! the binary types of src/rw_xvb.f and src/rw_xva.f, one file per
! configuration or appended to md.traj.<tend>.<ftype>; the formatted types of
! src/rw_xvfa.f and src/rw_xvfb.f into the file named by start, as the
! reference writes them) and the checkpoints
! (md.<time>.xv8b, src/md_utils.f:19-38) every nckpt steps, after the warm-up
! and at the end of the run (src/md.F:93, :104, :165, :176).  A run that saves
! or checkpoints is split into device runs between those steps; their headers
! carry what the reference's globals hold then: the energies, pressure and
! tensor of the step just done, all zero after a step that does not measure
! (every accel of the reference starts them from zero, src/int_ion_mix.f:132-137,
! src/int_ion_pure.f:243-248, src/int_nn.f:334-339; src/vel_verlet_2.f:107-110).
!
! YK_PRECISION=mixed selects the mixed-precision kernels; the mixture
! kernel's last-particle behaviour of the reference (src/int_ion_mix.f:158)
! is reproduced unless YK_FIX_F1=1.  `ykmd --init-only` stops after writing
! the initial configuration and the log's parameter echo (no device needed).
!-------------------------------------------------------------------------------
program ykmd
  use iso_c_binding
  use yk_bind
  use yk_host
  use yk_config
  implicit none

  ! run parameters (names and defaults of src/md_globals.f:36-135)
  character(20) :: sim_type = ''
  character(256) :: start = 'md.in.xv82'
  real(dp) :: tstart = -1.0d0, dt = 0.0d0
  integer :: nwgroup = 0, nwsteps = 0, ngroup = 0, ntot = 0, nind = 0, nptensor = 0
  integer :: ncom = 0, ntnorm = 0, ngofr = 0, nbin = 0, nsbin = 0, nckpt = 0, nout = 0
  real(dp) :: qmin = 0.0d0, dq = 0.0d0
  character(6) :: ftype = 'x4b'
  logical :: append = .true.
  character(256) :: spec_file = ''
  type(species_t) :: spec_list(MAXSPEC)
  real(dp) :: rho = 0.0d0, aspect(3) = 0.0d0, deps(3) = XNOSTRAIN, stretch(3) = 1.0d0
  real(dp) :: rccut = 1.0d50, kT = 0.0d0, xmass = 931.0d0
  character(20) :: coulomb = '', nuclear = ''
  real(dp) :: xlambda = -1.0d0, frp = 1.0d0
  real(dp) :: bfield = 0.0d0, efield = 0.0d0, q0 = 0.0d0, w0 = 0.0d0, tref = 0.0d0
  real(dp) :: xpacket = 1.25d0, aa = 110.0d0, bb = -26.0d0, cc = 24.0d0, rncut = 0.0d0
  integer :: irnd = 0, iseed = 0
  real(dp) :: rmax = 0.0d0
  namelist /run_parameters/ sim_type, start, tstart, dt, nwgroup, nwsteps, ngroup, ntot, &
                            nind, nptensor, ncom, ntnorm, ngofr, nbin, nsbin, qmin, dq, &
                            nckpt, nout, ftype, append, spec_file, spec_list, rho, aspect, &
                            deps, stretch, rccut, kT, xmass, coulomb, xlambda, frp, bfield, &
                            efield, q0, w0, tref, nuclear, xpacket, aa, bb, cc, rncut, &
                            irnd, iseed, rmax

  integer :: n, nspec, i, j, k, ig, ios, irtn, istep, nrec, r, u8, u9, nc, kd
  real(dp), allocatable :: x(:, :), v(:, :), a(:, :), zii(:), aii(:)
  real(dp) :: xlf, xl0(3), xl(3), ztot, xmuc, time, tend, ev, ek, px, e, pp(3, 3)
  real(dp) :: evave, ev2ave, pave, p2ave, eva, ev2a, pa, p2a, dev, dp_, devave, dpave
  real(dp) :: x1, x2, rhof
  real(dp) :: hev, hek, hpx, hpp(3, 3)   ! ev, ek, px, pp of src/md_globals.f now
  character(40) :: suffix
  character(20) :: sim_type_in
  character(20) :: xtype
  character(6) :: spec_type
  logical :: spec_read, za_in_start, v_generated
  character(256) :: env, arg
  type(yk_params_t) :: p
  type(yk_meas_t) :: m
  type(yk_meas_t), allocatable :: recs(:)
  integer(c_int32_t), allocatable :: rsteps(:)
  type(yk_loop_t) :: lp
  type(c_ptr) :: h
  integer(c_int32_t) :: nrec32
  real(c_double) :: box(3), eka

  open(10, file='runmd.in', status='old', form='formatted', iostat=ios)
  if (ios /= 0) call fail('cannot open runmd.in')
  read(10, nml=run_parameters, iostat=ios)
  if (ios /= 0) call fail('error reading namelist run_parameters from runmd.in')
  close(10)

  sim_type_in = sim_type
  if (nout /= 0) then
    select case (trim(ftype))
    case ('onp', 'md1', 'xfa', 'xvfa', 'xvZAfa', 'xfb', 'xvfb', 'xZAfb', 'xvZAfb', &
          'md0', 'x4a', 'xv4a', 'x8a', 'xv8a', 'xva8a', 'x4b', 'xv4b', 'x8b', 'xv8b', 'xva8b')
    case default
      call fail('ftype not supported: '//trim(ftype))
    end select
  end if

  ! species (src/input1.F:128-157): a spec_list in runmd.in wins over a
  ! spec_file; its species are counted here, in list order
  n = 0
  nspec = 0
  ztot = 0.0d0
  if (maxval(spec_list(:)%n) > 0) then
    spec_file = 'internal'
    do k = 1, MAXSPEC
      if (spec_list(k)%n == 0 .and. spec_list(k)%z == 0.0d0 .and. spec_list(k)%a == 0.0d0) exit
      if (spec_list(k)%n > 0) n = n + spec_list(k)%n
      nspec = k
    end do
  end if

  ! random-number seed (src/input1.F:174-177): MT19937 only
  if (irnd /= 0 .and. irnd /= 4) &
    call fail('only irnd = 0 or 4 (Mersenne Twister) is supported')
  call mt_seed(iseed)

  ! charges and masses (src/initconfig.f:104-163): from the list, from a
  ! species file, or later from a start file that holds them
  spec_read = .false.
  spec_type = ''
  if (trim(spec_file) == 'internal') then
    allocate(zii(0:n - 1), aii(0:n - 1))
    j = 0
    do k = 1, nspec
      ztot = ztot + spec_list(k)%n * spec_list(k)%z
      do i = 1, spec_list(k)%n
        zii(j) = spec_list(k)%z
        aii(j) = spec_list(k)%a
        j = j + 1
      end do
    end do
  else if (trim(spec_file) /= '') then
    spec_type = file_type(spec_file)
    select case (trim(spec_type))
    case ('ZAfa', 'ZAfb')
      call read_species(spec_file, spec_type, n, zii, aii, ztot, irtn)
      if (irtn < 0) call fail('cannot read spec_file '//trim(spec_file))
      spec_read = .true.
    end select
  end if

  ! initial configuration (src/initconfig.f:172-345)
  select case (trim(start))
  case ('random', 'nuclear')
    xtype = trim(start)
  case default
    xtype = file_type(start)
  end select
  za_in_start = .false.
  pp = 0.0d0
  ev = 0.0d0
  ek = 0.0d0
  px = 0.0d0
  time = 0.0d0
  irtn = 0
  select case (trim(xtype))
  case ('random', 'nuclear')
    if (n <= 0 .or. .not. allocated(zii)) call fail('no particles: give a spec_list or spec_file')
    call random_box(n, rho, aspect, xlf, xl0)
    xl = xl0
    if (tstart < 0.0d0) call fail('tstart not specified')
    allocate(x(3, 0:n - 1))
    if (trim(xtype) == 'random') then
      call random_positions(n, xl, x)
    else   ! a sphere about the box centre (src/initconfig.f:227-253)
      rmax = min(rmax, min(xl(1), xl(2), xl(3)) * 0.5)
      call sphere_positions(n, xl, rmax, x)
    end if
    irtn = GOT_X
  case ('onp', 'md1', 'xfa', 'xvfa', 'md0', 'x4a', 'xv4a', 'x8a', 'xv8a', 'xvZAfa')
    ! revision a (no box in the file): the box follows n, rho and the aspect
    call read_config(start, xtype, n, time, xl, xl0, ev, ek, px, pp, x, v, zii, aii, ztot, irtn)
    if (irtn < 0) call fail('cannot read start file '//trim(start))
    za_in_start = trim(xtype) == 'xvZAfa'
    if (tstart < 0.0d0) tstart = time
    if (aspect(1) * aspect(2) * aspect(3) == 0.0d0) aspect = 1.0d0
    xlf = (real(real(n), dp) / (aspect(1) * aspect(2) * aspect(3) * rho))**(1.0d0 / 3.0d0)
    xl0 = aspect * xlf
    xl = stretch * xl0
  case ('x4b', 'x8b', 'xv4b', 'xv8b', 'xfb', 'xvfb', 'xZAfb', 'xvZAfb')
    ! revision b: the box, time and n from the file; density and gage box follow
    call read_config(start, xtype, n, time, xl, xl0, ev, ek, px, pp, x, v, zii, aii, ztot, irtn)
    if (irtn < 0) call fail('cannot read start file '//trim(start))
    za_in_start = trim(xtype) == 'xZAfb' .or. trim(xtype) == 'xvZAfb'
    if (tstart < 0.0d0) tstart = time
    rho = real(real(n), dp) / (xl(1) * xl(2) * xl(3))
    if (aspect(1) * aspect(2) * aspect(3) == 0.0d0) then
      xlf = minval(xl)
      aspect = xl / xlf
      xl0 = xl
    else
      xlf = (real(real(n), dp) / (aspect(1) * aspect(2) * aspect(3) * rho))**(1.0d0 / 3.0d0)
      xl0 = aspect * xlf
    end if
    stretch = xl / xl0
  case default
    ! also 'xva8a', whose reference reader takes the particles from the time
    ! record and stops there
    call fail('input configuration type not supported: '//trim(start))
  end select
  if (.not. allocated(zii) .or. .not. allocated(aii)) &
    call fail('no charges and masses: give a spec_list, a spec_file or a start file with them')
  if (size(zii) /= n) call fail('the start file holds another particle count than the species')
  allocate(a(3, 0:n - 1))
  a = 0.0d0
  ! velocities from a Maxwell distribution when the start has none
  ! (src/initconfig.f:357-373)
  v_generated = iand(irtn, GOT_V) == 0
  if (v_generated) then
    if (allocated(v)) deallocate(v)
    allocate(v(3, 0:n - 1))
    call maxwell_velocities(n, aii, kT, xmass, v)
    call host_cancel_vcom(n, v, aii, xmass, kT)
  end if
  time = tstart
  ! the initial configuration, unless it came from an 'xv8b' file
  ! (src/initconfig.f:385-390): the header of the globals at this point
  if (trim(xtype) /= 'xv8b') then
    write(suffix, '(i5.5,i6.6)') int(time / 1000000.d0), int(mod(time, 1000000.d0))
    call write_xvb('md.'//trim(suffix)//'.xv8b', 'xv8b', sim_type, time, xl, ev, &
                   ek, px, pp, n, x, v, a)
  end if

  ! derived parameters (src/input1.F:184-196, :222, :255-311)
  xl = stretch * xl0
  if (trim(sim_type) == 'ion') then
    if (nspec == 1) then
      sim_type = 'pure-ion'
    else
      sim_type = 'ion-mixture'
    end if
  end if
  xmuc = 0.0d0
  if (trim(coulomb) == 'screened-coulomb') then
    if (xlambda > 0.0d0) then
      xmuc = 1.0d0 / xlambda
    else
      xmuc = thomas_fermi_mu((ztot * rho) / n)
    end if
  end if
  if (trim(nuclear) == 'HPP' .and. rncut == 0.0d0) rncut = auto_rncut(aa, bb, cc, xpacket)

  ! log file (src/md.F:44-55)
  tend = tstart + dt * (nwgroup * nwsteps + ngroup * ntot * nind)
  write(suffix, '(i5.5,i6.6)') int(tend / 1000000.d0), nint(mod(tend, 1000000.d0))
  open(newunit=u8, file='md.'//trim(suffix)//'.log', form='formatted', status='unknown')
  call echo_parameters(u8)

  call get_command_argument(1, arg)
  if (trim(arg) == '--init-only') stop


  ! the device system
  select case (trim(sim_type))
  case ('pure-ion')
    p%sim_type = YK_PURE_ION
    p%zi = zii(0)
    p%ai = aii(0)
  case ('ion-mixture')
    p%sim_type = YK_ION_MIXTURE
  case ('nucleon')
    p%sim_type = YK_NUCLEON
  case default
    call fail('sim_type not supported: '//trim(sim_type))
  end select
  p%coulomb_on = merge(1, 0, trim(coulomb) == 'screened-coulomb')
  p%nuclear_on = merge(1, 0, trim(nuclear) == 'HPP')
  p%xl = xl
  p%xmuc = xmuc
  p%rccut = rccut
  p%frp = frp
  p%vc = ALPHA * HBARC
  p%xmass = xmass
  p%rho = rho
  p%kT = kT
  p%aa = aa
  p%bb = bb
  p%cc = cc
  p%xpacket = xpacket
  p%rncut = rncut
  p%bfield = bfield
  call get_environment_variable('YK_PRECISION', env)
  if (trim(env) == 'mixed') p%precision = YK_MIXED
  call get_environment_variable('YK_FIX_F1', env)
  if (trim(env) /= '1') p%flags = ior(p%flags, YK_FLAG_REF_F1)
  call check(yk_create(p, int(n, c_int32_t), zii, aii, h), 'yk_create')
  call check(yk_put_state(h, x, v), 'yk_put_state')


  ! initial state (src/md.F:66-77)
  call check(yk_accel(h, 1_c_int32_t, m), 'yk_accel')
  call check(yk_ttot(h, eka), 'yk_ttot')
  ek = eka
  e = m%ev + ek
  write(u8, 10050) '  Initial Vtot, Ttot, E, P = ', m%ev, ek, e, m%px
10050 format(/, a, 4(2x, es16.9))
  hev = m%ev
  hek = eka
  hpx = m%px
  hpp = 0.0d0

  ! warm-up groups (src/md.F:80-101)
  istep = 0
  if (nwgroup >= 1) write(u8, '(/a)') '  Warmup phase'
  allocate(recs(max(1, ntot * nind)), rsteps(max(1, ntot * nind)))
  do ig = 1, nwgroup
    kd = 0
    do while (kd < nwsteps)
      nc = min(nwsteps - kd, steps_to(istep, nckpt))
      call check(yk_run(h, dt, int(nc, c_int32_t), 0_c_int32_t, recs), 'yk_run')
      do k = 1, nc
        time = time + dt
      end do
      istep = istep + nc
      kd = kd + nc
      hev = 0.0d0          ! accel(.false.) and newton(.false.)
      hek = 0.0d0
      hpx = 0.0d0
      hpp = 0.0d0
      if (mdmod(istep, nckpt) == 0) call checkpoint()
    end do
    call check(yk_cancel_vcom(h), 'yk_cancel_vcom')
    call check(yk_tnorm(h), 'yk_tnorm')
    write(u8, 10082) istep, time
  end do
10082 format(2x, i7, 1x, f14.2, '   warm up')
  if (istep > 0 .and. mdmod(istep, nckpt) /= 0) call checkpoint()

  ! measurement groups (src/md.F:105-168)
  if (ngroup >= 1) then
    write(u8, 10090)
    write(suffix, '(i5.5,i6.6)') int(tend / 1000000.d0), nint(mod(tend, 1000000.d0))
    open(newunit=u9, file='md.'//trim(suffix)//'.meas', form='unformatted', status='unknown')
  end if
10090 format('  Measurement phase', /, &
             'e   step     time            Vtot              Ttot               E              pressure', /, &
             'p   step     time            P_xx              P_xy              P_xz              P_yy              P_yz              P_zz')
  evave = 0.0d0
  ev2ave = 0.0d0
  pave = 0.0d0
  p2ave = 0.0d0
  istep = 0
  nrec = 0
  ev = m%ev
  ek = eka
  px = m%px
  hpp = 0.0d0
  do ig = 1, ngroup
    eva = 0.0d0
    ev2a = 0.0d0
    pa = 0.0d0
    p2a = 0.0d0
    kd = 0
    do while (kd < ntot * nind)
      ! one device run up to the next configuration output or checkpoint
      nc = min(ntot * nind - kd, steps_to(istep, nout), steps_to(istep, nckpt))
      lp%dt = dt
      lp%nsteps = int(nc, c_int32_t)
      lp%istep0 = int(istep, c_int32_t)
      lp%nind = int(nind, c_int32_t)
      lp%nptensor = int(nptensor, c_int32_t)
      lp%ncom = int(ncom, c_int32_t)
      lp%ntnorm = int(ntnorm, c_int32_t)
      lp%deps = deps
      lp%xl0 = xl0
      call check(yk_run_md(h, lp, recs, rsteps, int(size(recs), c_int32_t), nrec32, box), &
                 'yk_run_md')
      nrec = nrec32
      xl = box
      r = 0
      do k = 1, nc
        time = time + dt
        istep = istep + 1
        if (mdmod(istep, nind) /= 0 .and. mdmod(istep, nptensor) /= 0) then
          hev = 0.0d0
          hek = 0.0d0
          hpx = 0.0d0
          hpp = 0.0d0
          cycle
        end if
        r = r + 1
        if (r > nrec .or. rsteps(r) /= istep) call fail('measurement records out of step')
        ev = recs(r)%ev
        ek = recs(r)%ek
        px = recs(r)%px
        hev = ev
        hek = ek
        hpx = px
        hpp(1, 1) = recs(r)%pp(1); hpp(1, 2) = recs(r)%pp(2); hpp(1, 3) = recs(r)%pp(3)
        hpp(2, 1) = recs(r)%pp(2); hpp(2, 2) = recs(r)%pp(4); hpp(2, 3) = recs(r)%pp(5)
        hpp(3, 1) = recs(r)%pp(3); hpp(3, 2) = recs(r)%pp(5); hpp(3, 3) = recs(r)%pp(6)
        if (mdmod(istep, nind) == 0) then
          eva = eva + ev
          ev2a = ev2a + ev**2
          pa = pa + px
          p2a = p2a + px**2
          e = ek + ev
          write(u8, 10101) 'e', istep, time, ev, ek, e, px
        end if
        if (mdmod(istep, nptensor) == 0) write(u8, 10101) 'p', istep, time, recs(r)%pp
        write(u9) istep, time, ev, ek, px, recs(r)%pp
      end do
      kd = kd + nc
      if (mdmod(istep, nout) == 0) call save_config()
      if (mdmod(istep, nckpt) == 0) call checkpoint()
    end do
    call block_stats(eva, ev2a, ntot, dev, evave, ev2ave)
    call block_stats(pa, p2a, ntot, dp_, pave, p2ave)
    write(u8, 10110) ig, time, eva, dev, pa, dp_
  end do
10101 format(a1, 1x, i7, 1x, f14.2, 6(2x, es16.9))
10110 format('s', i4, 1x, f14.2, 2x, 1pe16.9, 2x, 1pe16.9, 2x, 16x, 2x, 1pe16.9, 2x, 1pe16.9)

  ! run averages and the end of the run (src/md.F:170-205)
  x1 = 0.0d0
  x2 = 0.0d0
  if (ngroup >= 1) then
    call block_stats(evave, ev2ave, ngroup, devave, x1, x2)
    call block_stats(pave, p2ave, ngroup, dpave, x1, x2)
    write(u8, 10200) '  average potential = ', evave, ' +/-', devave
    write(u8, 10200) '  average pressure  = ', pave, ' +/-', dpave
    close(u9)
  end if
10200 format(2(a, 1pe16.9))
  ! end-of-run checkpoint (src/md.F:176)
  if (istep > 0 .and. mdmod(istep, nckpt) /= 0) call checkpoint()
  call check(yk_accel(h, 1_c_int32_t, m), 'yk_accel')
  call check(yk_ttot(h, eka), 'yk_ttot')
  ek = eka
  e = m%ev + ek
  stretch = xl / xl0
  aspect = stretch * aspect
  rhof = real(real(n), dp) / (xl(1) * xl(2) * xl(3))
  write(u8, 10206) xl, aspect, stretch, rhof
10206 format('  Final dimensions  =', 3(2x, f17.9), ' fm', / &
             '  Final aspect      =', 3(2x, f17.9), / &
             '  Final stretch     =', 3(2x, f17.9), / &
             '  Final density     =', 1x, es15.8, ' fm-3')
  write(u8, 10210) '  Final Vtot, Ttot, E, P   = ', m%ev, ek, e, m%px
10210 format(a, 4(2x, es16.9))
  close(u8)
  call yk_destroy(h)

contains

  ! The parameter echo that opens the reference's log (src/input1.F:82-373,
  ! src/initconfig.f:197-200, :216, :260, :350, :366-368), with the values this
  ! run uses; the first line names this program and the OpenMP line is left out.
  subroutine echo_parameters(u)
    integer, intent(in) :: u
    character(8) :: date
    character(10) :: daytime
    character(5) :: zone
    character(16) :: how
    real(dp) :: lam
    integer :: ks
    character(6), parameter :: rngname(0:4) = [character(6) :: 'N/A', 'ran11', 'rannyu', &
                                                'ran3', 'grnd']
    write(u, '(2x,a)') 'ykmd F90+HIP'
    call date_and_time(date, daytime, zone)
    write(u, '(''  Start date/time = '',a8,2x,a10,2x,a5)') date, daytime, zone
    write(u, '(''  Number of processes:  '',i6)') 1
    write(u, '(a)') '  ----------------------------------------------------------'
    write(u, '(/,''  SIMULATION TYPE:  '',a /''  ------------------------'')') trim(sim_type_in)
    if (trim(spec_file) == 'internal') then
      write(u, '(a)') '  Species  #particles    Z        A'
      do ks = 1, nspec
        if (spec_list(ks)%n > 0) write(u, '(3x,i4,4x,i9,2x,f7.2,2x,f7.2)') ks, spec_list(ks)%n, &
                                       spec_list(ks)%z, spec_list(ks)%a
      end do
      write(u, '(''   Total:  '',i9)') n
    end if
    if (irnd > 0) write(u, '(/''  irnd  = '',i9,2x,''('',a,'')''/''  iseed = '',i9)') &
      irnd, trim(rngname(irnd)), iseed
    if (spec_read) then   ! src/initconfig.f:143-161
      write(u, '(''  spec file:  '',a)') trim(spec_file)
      write(u, '(''     n = '',i9)') n
    end if
    write(u, '(/,''  INITIAL CONFIGURATION'',/''  ---------------------'')')
    select case (trim(xtype))
    case ('random')
      write(u, '(a)') '  Generating random coordinates in box'
    case ('nuclear')
      write(u, '(''  Generating random coordinates in sphere''/''  radius     = '',f8.4,'' fm'')') rmax
    case default
      write(u, '(''  file          = '',a)') trim(start)
      if (za_in_start) write(u, '(''  This file also includes charges and masses.'', / &
                                 &''     n          = '',i10)') n
    end select
    if (v_generated) write(u, '(a)') '  Generating velocities from Maxwell distribution' 
    write(u, '(''  start time    =  '',f14.2,'' fm/c''/''  time step     =        '',f8.2,'' fm/c'')') &
      tstart, dt
    write(u, '(/,''  rho               ='',es16.8,'' fm-3''/''  L0                ='',2x,f17.9, &
             &'' fm''/''  gage dimensions   ='',3(2x,f17.9),'' fm''/''  gage aspect       ='', &
             &3(6x,f13.9)/''  strainrates       ='')', advance='no') rho, xlf, xl0, aspect
    do ks = 1, 3
      if (deps(ks) == XNOSTRAIN) then
        write(u, '(19x)', advance='no')
      else
        write(u, '(2x,es17.9)', advance='no') deps(ks)
      end if
    end do
    write(u, '('' c/fm'')')
    write(u, '(''  config dimensions ='',3(2x,f17.9),'' fm''/''  config stretch    ='',3(2x,f17.9))') &
      xl, stretch
    write(u, '(/,''  kT                = '',f10.6,'' MeV''/''  nucleon mass      = '',f8.3,'' MeV'')') &
      kT, xmass
    if (ngofr > 0) write(u, '(a)') '  *** g(r) and S(q) disabled in this version ***'
    write(u, '(/''  INTERACTIONS''/''  ------------'')')
    if (trim(coulomb) == 'screened-coulomb') then
      write(u, '(2x,a)') trim(coulomb)
      if (xlambda > 0.0d0) then
        lam = xlambda
        how = '(defined)'
      else
        lam = 1.0d0 / xmuc
        how = '(Thomas-Fermi)'
      end if
      write(u, '(''     screening length = '',f13.8,'' fm  '',a/''     ion form factor  = '',f13.8/ &
               &''     rccut            = '',f13.8,'' fm'')') lam, trim(how), frp, rccut
    end if
    if (trim(nuclear) == 'HPP') then
      write(u, '(2x,a)') trim(nuclear)
      write(u, '(''     aa       = '',f8.3,'' MeV''/''     bb       = '',f8.3,'' MeV''/ &
               &''     cc       = '',f8.3,'' MeV''/''     Lambda   = '',f8.3,'' fm^2''/ &
               &''     rncut    = '',f8.3,'' fm'')') aa, bb, cc, xpacket, rncut
    end if
    if (bfield > 0.0d0) write(u, '(''  B-field:''/''     B0   = '',1pe10.3,'' Gauss'')') bfield
    if (efield > 0.0d0) write(u, '(a)') '  *** external E field disabled in this version ***'
    write(u, '(/,''  TIME STEP INTERVALS''/''  -------------------''/''  nwgroup, nwsteps:      '', &
             &i8,1x,i8/''  ngroup, ntot, nind:    '',i8,1x,i8,1x,i8/''  ntnorm:                '', &
             &i8/''  ncom:                  '',i8/''  nptensor:              '',i8)') &
      nwgroup, nwsteps, ngroup, ntot, nind, ntnorm, ncom, nptensor
    write(u, '(''  nout, ftype, append:   '',i8,7x,a5,5x,l1/''  nckpt:                 '',i8)') &
      nout, ftype, append, nckpt
  end subroutine echo_parameters

  ! Steps from istep to the next step whose mdmod with q is 0 (q = 0: never).
  integer function steps_to(istep, q)
    integer, intent(in) :: istep, q
    if (q == 0) then
      steps_to = huge(0)
    else
      steps_to = abs(q) - mod(istep, abs(q))
    end if
  end function steps_to

  ! checkpoint (src/md_utils.f:19-38): md.<time>.xv8b
  subroutine checkpoint()
    character(40) :: sfx
    call check(yk_get_state(h, x, v, a), 'yk_get_state')
    write(sfx, '(i5.5,i6.6)') int(time / 1000000.d0), nint(mod(time, 1000000.d0))
    call write_xvb('md.'//trim(sfx)//'.xv8b', 'xv8b', sim_type, time, xl, hev, hek, hpx, hpp, &
                   n, x, v, a)
  end subroutine checkpoint

  ! save_config(ftype, append) (src/md_utils.f:61-114): one file per
  ! configuration, md.<time>.<ftype>, or appended to md.traj.<tend>.<ftype>
  subroutine save_config()
    character(40) :: sfx
    character(300) :: fname
    call check(yk_get_state(h, x, v, a), 'yk_get_state')
    if (append) then
      write(sfx, '(i5.5,i6.6)') int(tend / 1000000.d0), nint(mod(tend, 1000000.d0))
      fname = 'md.traj.'//trim(sfx)//'.'//trim(ftype)
    else
      write(sfx, '(i5.5,i6.6)') int(time / 1000000.d0), nint(mod(time, 1000000.d0))
      fname = 'md.'//trim(sfx)//'.'//trim(ftype)
    end if
    select case (trim(ftype))
    case ('x4b', 'xv4b', 'x8b', 'xv8b')
      call write_xvb(trim(fname), ftype, sim_type, time, xl, hev, hek, hpx, hpp, n, x, v, a, &
                     append)
    case ('md0', 'x4a', 'xv4a', 'x8a', 'xv8a', 'xva8a')
      call write_xva(trim(fname), ftype, time, n, x, v, a, append)
    ! the formatted types go to the file named by start, as the reference
    ! has it (src/md_utils.f:83-104); the revision-a ones always appended
    case ('onp', 'md1', 'xfa', 'xvfa')
      call write_config_fa(start, ftype, .true., time, n, x, v, zii, aii)
    case ('xvZAfa')
      call write_config_fa(start, ftype, append, time, n, x, v, zii, aii)
    case ('xfb', 'xvfb', 'xZAfb', 'xvZAfb')
      call write_config_fb(start, ftype, append, sim_type, time, xl0, xl, hev, hek, hpx, hpp, &
                           n, x, v, zii, aii)
    end select   ! 'xva8b': nothing (no case of src/md_utils.f:82-105)
  end subroutine save_config

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    if (rc /= YK_OK) call fail(what//': '//yk_error_message())
  end subroutine check

  subroutine fail(msg)
    character(*), intent(in) :: msg
    write(0, '(a,a)') 'ykmd: ', msg
    error stop 1
  end subroutine fail

end program ykmd
