!-------------------------------------------------------------------------------
! yk_host.f90 -- host-side pieces of an MD run around the GPU library:
! the initial state of a 'random' start, the derived run parameters, block
! statistics and the revision-b configuration files.  Used by ykmd.f90.
!
!   mt_seed / mt_uniform   Mersenne Twister MT19937 (Matsumoto and Nishimura,
!                          ACM TOMACS 8 (1998) 3) with the state filled by the
!                          multiplier 69069 from the seed, as the reference's
!                          sgrnd/grnd (src/randnums.f:239-323): uniform in
!                          [0, 1] as the 32-bit output over 2^32 - 1
!   gauss_polar            Marsaglia's polar method (Marsaglia and Bray, SIAM
!                          Review 6 (1964) 260): a pair per accepted point, the
!                          second returned by the next call, with the
!                          reference's arithmetic (src/randnums.f:5-25)
!   random_positions       x(k,i) = xl(k) * u, particle by particle
!                          (src/initconfig.f:218-223)
!   sphere_positions       start='nuclear': uniform in a sphere of radius
!                          rmax about the box centre, three draws per
!                          particle (src/initconfig.f:227-253)
!   maxwell_velocities     v(k,i) = sqrt(kT/(A_i xmass)) * g
!                          (src/initconfig.f:366-372)
!   host_ttot / host_tnorm / host_cancel_vcom
!                          the velocity utilities in particle order
!                          (src/gen_utils.f:40-107), for the initial state
!                          (the step loop uses the device versions)
!   random_box             edge lengths of a random start
!                          (src/initconfig.f:209-214)
!   thomas_fermi_mu        screening when no length is given (src/input1.F:262-273)
!   auto_rncut             nuclear cutoff when rncut = 0 (src/input1.F:292-311)
!   block_stats            statt (src/md.F:228-243)
!   write_xvb / read_xvb   revision-b configuration files (header records,
!                          then the particles), types 'x4b', 'xv4b', 'x8b',
!                          'xv8b', 'xva8b' (src/rw_xvb.f:148-203); read_xvb
!                          reads 'xv8b'
!   write_xva              revision-a configuration files (time, then the
!                          particles; no header), types 'md0', 'x4a', 'xv4a',
!                          'x8a', 'xv8a', 'xva8a' (src/rw_xva.f:169-223)
!
! Run with one thread of the reference when comparing: its OpenMP reductions
! sum in thread order.
!-------------------------------------------------------------------------------
module yk_host
  use iso_c_binding
  implicit none
  private

  integer, parameter, public :: dp = kind(1.0d0)
  integer, parameter, public :: MAXSPEC = 1000
  real(dp), parameter, public :: XNOSTRAIN = 1000.0d0
  ! physical constants of src/md_globals.f:24-28
  real(dp), parameter, public :: HBARC = 197.327d0, ALPHA = 1.0d0 / 137.036d0, &
                                 ME = 0.510999d0, PI = 3.14159265358979324d0

  type, public :: species_t
    integer :: n = 0
    real(dp) :: z = 0.0d0
    real(dp) :: a = 0.0d0
  end type species_t

  ! MT19937 state: 32-bit words held in 64-bit integers (no signed overflow)
  integer, parameter :: MT_N = 624, MT_M = 397
  integer(8), parameter :: MASK32 = 4294967295_8, UPPER = 2147483648_8, LOWER = 2147483647_8
  integer(8), parameter :: MATRIX_A = 2567483615_8, TEMPER_B = 2636928640_8, &
                           TEMPER_C = 4022730752_8
  integer(8), save :: mt(0:MT_N - 1) = 0
  integer, save :: mti = MT_N + 1
  ! second value of the last accepted polar point
  logical, save :: have_spare = .false.
  real(dp), save :: spare = 0.0d0

  public :: mt_seed, mt_uniform, gauss_polar, random_positions, sphere_positions, &
            maxwell_velocities, &
            host_ttot, host_tnorm, host_cancel_vcom, random_box, thomas_fermi_mu, &
            auto_rncut, block_stats, mdmod, write_xvb, write_xva, read_xvb

contains

  subroutine mt_seed(seed)
    integer, intent(in) :: seed
    integer :: k
    mt(0) = iand(int(seed, 8), MASK32)
    do k = 1, MT_N - 1
      mt(k) = iand(69069_8 * mt(k - 1), MASK32)
    end do
    mti = MT_N
    have_spare = .false.
  end subroutine mt_seed

  real(dp) function mt_uniform()
    integer(8) :: y
    integer :: k
    if (mti >= MT_N) then
      if (mti == MT_N + 1) call mt_seed(4357)
      do k = 0, MT_N - 1
        y = ior(iand(mt(k), UPPER), iand(mt(mod(k + 1, MT_N)), LOWER))
        mt(k) = ieor(mt(mod(k + MT_M, MT_N)), ishft(y, -1))
        if (iand(y, 1_8) == 1_8) mt(k) = ieor(mt(k), MATRIX_A)
      end do
      mti = 0
    end if
    y = mt(mti)
    mti = mti + 1
    y = ieor(y, ishft(y, -11))
    y = ieor(y, iand(ishft(y, 7), TEMPER_B))
    y = ieor(y, iand(ishft(y, 15), TEMPER_C))
    y = ieor(y, ishft(y, -18))
    mt_uniform = real(y, dp) / 4294967295.0d0
  end function mt_uniform

  real(dp) function gauss_polar()
    real(dp) :: v1, v2, rsq, fac
    if (have_spare) then
      have_spare = .false.
      gauss_polar = spare
      return
    end if
    do
      v1 = 2.0d0 * mt_uniform() - 1.0d0
      v2 = 2.0d0 * mt_uniform() - 1.0d0
      rsq = v1 * v1 + v2 * v2
      if (rsq < 1.0d0 .and. rsq /= 0.0d0) exit
    end do
    fac = sqrt(-2.0d0 * log(rsq) / rsq)
    spare = v1 * fac
    have_spare = .true.
    gauss_polar = v2 * fac
  end function gauss_polar

  subroutine random_positions(n, xl, x)
    integer, intent(in) :: n
    real(dp), intent(in) :: xl(3)
    real(dp), intent(out) :: x(3, 0:n - 1)
    integer :: i, k
    do i = 0, n - 1
      do k = 1, 3
        x(k, i) = xl(k) * mt_uniform()
      end do
    end do
  end subroutine random_positions

  ! r = u1**0.3333333 rmax (the exponent a default-real literal, as in the
  ! reference), cos(theta) = 2 u2 - 1, phi = 2 pi u3, about the box centre.
  subroutine sphere_positions(n, xl, rmax, x)
    integer, intent(in) :: n
    real(dp), intent(in) :: xl(3), rmax
    real(dp), intent(out) :: x(3, 0:n - 1)
    integer :: i
    real(dp) :: xrr, costheta, sintheta, xphi
    do i = 0, n - 1
      xrr = mt_uniform()**.3333333 * rmax
      costheta = 2. * mt_uniform() - 1.
      sintheta = sqrt(1.d0 - costheta * costheta)
      xphi = 2.d0 * PI * mt_uniform()
      x(1, i) = xrr * sintheta * cos(xphi) + 0.5 * xl(1)
      x(2, i) = xrr * sintheta * sin(xphi) + 0.5 * xl(2)
      x(3, i) = xrr * costheta + 0.5 * xl(3)
    end do
  end subroutine sphere_positions

  subroutine maxwell_velocities(n, aii, kT, xmass, v)
    integer, intent(in) :: n
    real(dp), intent(in) :: aii(0:n - 1), kT, xmass
    real(dp), intent(out) :: v(3, 0:n - 1)
    integer :: i, k
    real(dp) :: velfac
    do i = 0, n - 1
      velfac = sqrt(kT / (aii(i) * xmass))
      do k = 1, 3
        v(k, i) = velfac * gauss_polar()
      end do
    end do
  end subroutine maxwell_velocities

  real(dp) function host_ttot(n, v, aii, xmass)
    integer, intent(in) :: n
    real(dp), intent(in) :: v(3, 0:n - 1), aii(0:n - 1), xmass
    real(dp) :: s
    integer :: i
    s = 0.0d0
    do i = 0, n - 1
      s = s + aii(i) * (v(1, i) * v(1, i) + v(2, i) * v(2, i) + v(3, i) * v(3, i))
    end do
    host_ttot = (0.5d0 * xmass * s) / real(real(n), dp)   ! float(n), src/gen_utils.f:53
  end function host_ttot

  subroutine host_tnorm(n, v, aii, xmass, kT)
    integer, intent(in) :: n
    real(dp), intent(inout) :: v(3, 0:n - 1)
    real(dp), intent(in) :: aii(0:n - 1), xmass, kT
    real(dp) :: fac
    fac = sqrt(host_ttot(n, v, aii, xmass) / (1.5d0 * kT))
    v = v / fac
  end subroutine host_tnorm

  subroutine host_cancel_vcom(n, v, aii, xmass, kT)
    integer, intent(in) :: n
    real(dp), intent(inout) :: v(3, 0:n - 1)
    real(dp), intent(in) :: aii(0:n - 1), xmass, kT
    real(dp) :: mtot, ptot(3)
    integer :: i
    mtot = 0.0d0
    ptot = 0.0d0
    do i = 0, n - 1
      mtot = mtot + aii(i)
      ptot = ptot + aii(i) * v(:, i)
    end do
    ptot = ptot / mtot
    do i = 0, n - 1
      v(:, i) = v(:, i) - ptot
    end do
    call host_tnorm(n, v, aii, xmass, kT)
  end subroutine host_cancel_vcom

  ! fundamental edge and gage dimensions of a random start; aspect = 1 when
  ! any ratio is 0
  subroutine random_box(n, rho, aspect, xlf, xl0)
    integer, intent(in) :: n
    real(dp), intent(in) :: rho
    real(dp), intent(inout) :: aspect(3)
    real(dp), intent(out) :: xlf, xl0(3)
    if (aspect(1) * aspect(2) * aspect(3) == 0.0d0) aspect = 1.0d0
    xlf = (real(real(n), dp) / (aspect(1) * aspect(2) * aspect(3) * rho))**(1.0d0 / 3.0d0)
    xl0 = aspect * xlf
  end subroutine random_box

  ! inverse Thomas-Fermi screening length of an electron gas of density rhoe
  real(dp) function thomas_fermi_mu(rhoe)
    real(dp), intent(in) :: rhoe
    real(dp) :: xkf, ef, qtf
    xkf = (3.0d0 * (PI**2) * rhoe)**(1.0d0 / 3.0d0) * HBARC
    ef = sqrt(xkf**2 + ME**2)
    qtf = 2.0d0 * sqrt((ALPHA / PI) * xkf * ef)
    thomas_fermi_mu = qtf / HBARC
  end function thomas_fermi_mu

  ! smallest r (a step of 0.1 in units of sqrt(2 Lambda)) at which the HPP
  ! force falls below 2^-53 of 1e-4
  real(dp) function auto_rncut(aa, bb, cc, xpacket)
    real(dp), intent(in) :: aa, bb, cc, xpacket
    real(dp) :: eps, r, r0, step
    eps = (sqrt(2.0d0 * xpacket) / (2.0d0 * (-bb + cc))) * 0.0001d0 * (2.0d0**(-53))
    r = sqrt(-log((-bb - cc) / (2.0d0 * aa)))
    r0 = r
    step = 1.0d0
    do
      do while (r * exp(-r**2) > eps)
        r0 = r
        r = r + step
      end do
      if (step < 0.11d0) exit
      step = step / 10.0d0
      r = r0
    end do
    auto_rncut = r * sqrt(2.0d0 * xpacket)
  end function auto_rncut

  ! block mean and standard error of the mean from n samples' sum and sum of
  ! squares; adds the block mean and its square to the run totals
  subroutine block_stats(xa, xa2, n, dx, xave, x2ave)
    real(dp), intent(inout) :: xa, xa2, xave, x2ave
    integer, intent(in) :: n
    real(dp), intent(out) :: dx
    xa = xa / real(real(n), dp)
    xa2 = xa2 / real(real(n), dp)
    if (n > 1) then
      dx = sqrt(abs(xa2 - xa * xa) / real(real(n - 1), dp))
    else
      dx = 0.0d0
    end if
    xave = xave + xa
    x2ave = x2ave + xa * xa
  end subroutine block_stats

  ! p mod q, and -1 (never 0) for q = 0
  integer function mdmod(p, q)
    integer, intent(in) :: p, q
    if (q == 0) then
      mdmod = -1
    else
      mdmod = mod(p, q)
    end if
  end function mdmod

  ! Revision-b configuration file: type, code name and version, date, sim
  ! type, (time, box, ev, ek, px, pp(3,3), n), then the particle record.
  ! Opens a configuration file for one more configuration: a new file, or the
  ! end of an existing one when append (a trajectory file).
  subroutine open_config(file, append, u)
    character(*), intent(in) :: file
    logical, intent(in) :: append
    integer, intent(out) :: u
    if (append) then
      open(newunit=u, file=file, status='unknown', form='unformatted', position='append')
    else
      open(newunit=u, file=file, status='unknown', form='unformatted')
    end if
  end subroutine open_config

  subroutine write_xvb(file, ftype, sim_type, time, xl, ev, ek, px, pp, n, x, v, a, append)
    character(*), intent(in) :: file, ftype, sim_type
    real(dp), intent(in) :: time, xl(3), ev, ek, px, pp(3, 3)
    integer, intent(in) :: n
    real(dp), intent(in) :: x(3, 0:n - 1), v(3, 0:n - 1), a(3, 0:n - 1)
    logical, intent(in), optional :: append
    character(6) :: ft
    character(10) :: cname
    character(8) :: cver, date
    character(10) :: daytime
    character(5) :: zone
    character(20) :: st
    integer :: u, i
    logical :: app
    app = .false.
    if (present(append)) app = append
    ft = ftype
    cname = 'MD'
    cver = '6.2.0'
    st = sim_type
    call date_and_time(date, daytime, zone)
    call open_config(file, app, u)
    write(u) ft
    write(u) cname, cver
    write(u) date, daytime, zone
    write(u) st
    write(u) time, xl(1), xl(2), xl(3), ev, ek, px, pp, n
    select case (trim(ft))
    case ('x4b')
      write(u) (real(x(1, i)), real(x(2, i)), real(x(3, i)), i = 0, n - 1)
    case ('xv4b')
      write(u) (real(x(1, i)), real(x(2, i)), real(x(3, i)), &
                real(v(1, i)), real(v(2, i)), real(v(3, i)), i = 0, n - 1)
    case ('x8b')
      write(u) (x(1, i), x(2, i), x(3, i), i = 0, n - 1)
    case ('xv8b')
      write(u) (x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i), i = 0, n - 1)
    case ('xva8b')
      write(u) (x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i), &
                a(1, i), a(2, i), a(3, i), i = 0, n - 1)
    end select
    close(u)
  end subroutine write_xvb

  subroutine write_xva(file, ftype, time, n, x, v, a, append)
    character(*), intent(in) :: file, ftype
    real(dp), intent(in) :: time
    integer, intent(in) :: n
    real(dp), intent(in) :: x(3, 0:n - 1), v(3, 0:n - 1), a(3, 0:n - 1)
    logical, intent(in) :: append
    integer :: u, i
    call open_config(file, append, u)
    select case (trim(ftype))
    case ('md0')
      write(u) (x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i), i = 0, n - 1)
    case ('x4a')
      write(u) real(time)
      write(u) (real(x(1, i)), real(x(2, i)), real(x(3, i)), i = 0, n - 1)
    case ('xv4a')
      write(u) real(time)
      write(u) (real(x(1, i)), real(x(2, i)), real(x(3, i)), &
                real(v(1, i)), real(v(2, i)), real(v(3, i)), i = 0, n - 1)
    case ('x8a')
      write(u) time
      write(u) (x(1, i), x(2, i), x(3, i), i = 0, n - 1)
    case ('xv8a')
      write(u) time
      write(u) (x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i), i = 0, n - 1)
    case ('xva8a')
      write(u) time
      write(u) (x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i), &
                a(1, i), a(2, i), a(3, i), i = 0, n - 1)
    end select
    close(u)
  end subroutine write_xva

  ! Reads a revision-b 'xv8b' file: box, time and n, then positions and
  ! velocities (allocated here).  irtn = 0 on success.
  subroutine read_xvb(file, time, xl, n, x, v, irtn)
    character(*), intent(in) :: file
    real(dp), intent(out) :: time, xl(3)
    integer, intent(out) :: n, irtn
    real(dp), allocatable, intent(out) :: x(:, :), v(:, :)
    character(6) :: ft
    character(10) :: cname, daytime
    character(8) :: cver, date
    character(5) :: zone
    character(20) :: st
    real(dp) :: ev, ek, px, pp(3, 3)
    integer :: u, i, ios
    irtn = 1
    open(newunit=u, file=file, status='old', form='unformatted', iostat=ios)
    if (ios /= 0) return
    read(u) ft
    read(u) cname, cver
    read(u) date, daytime, zone
    read(u) st
    read(u) time, xl(1), xl(2), xl(3), ev, ek, px, pp, n
    if (trim(ft) /= 'xv8b') then
      close(u)
      irtn = 2
      return
    end if
    allocate(x(3, 0:n - 1), v(3, 0:n - 1))
    read(u) (x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i), i = 0, n - 1)
    close(u)
    irtn = 0
  end subroutine read_xvb

end module yk_host
