!-------------------------------------------------------------------------------
! yk_config.f90 -- species and configuration files of an MD run, read and
! written as the reference reads and writes them, for the run driver ykmd.f90.
!
!   file_type          a file name's extension as the type of the file: the
!                      characters after the last '.', 'none' without one,
!                      'long' beyond six (src/gen_utils.f:18-33)
!   read_species       spec_file types 'ZAfa' (one "k Z A" line per particle,
!                      k counted from 1) and 'ZAfb' (runs "count Z A"), after
!                      '#' comment lines and a particle-count line
!                      (src/rw_xvfa.f:84-98, src/rw_xvfb.f:11-62)
!   read_config        every start-file type of src/initconfig.f:254-345:
!                      formatted revision a 'onp', 'md1', 'xfa', 'xvfa',
!                      'xvZAfa' (src/rw_xvfa.f:11-118); binary revision a
!                      'md0', 'x4a', 'xv4a', 'x8a', 'xv8a' (src/rw_xva.f:39-
!                      164); binary revision b 'x4b', 'x8b', 'xv4b', 'xv8b'
!                      (src/rw_xvb.f:36-143); formatted revision b 'xfb',
!                      'xvfb', 'xZAfb', 'xvZAfb' (src/rw_xvfb.f:127-205)
!   write_config_fa    formatted revision a 'md1', 'xfa', 'xvfa', 'xvZAfa'
!                      (src/rw_xvfa.f:122-178)
!   write_config_fb    formatted revision b 'xfb', 'xvfb', 'xZAfb', 'xvZAfb'
!                      (src/rw_xvfb.f:210-288)
!
! The reading routines return what the file held as the reference's bit
! field: 16 positions, 256 velocities, 65536 charges, 1048576 masses.  Two of
! the reference's readers cannot read what its writers produce: the
! formatted revision-b reader takes the header as six list-directed records
! (type; code name and version; date, time and zone; simulation type; the
! gage box; time, box, ev, ek, px, pp(6), n), where the writer puts '# '
! comment lines; and the revision-a 'xva8a' reader takes the particles from
! the time record.  The first is restated as the reader has it; 'xva8a' is
! refused, as the reference's run ends at that read.
!-------------------------------------------------------------------------------
module yk_config
  use yk_host, only: dp
  implicit none
  private

  integer, parameter, public :: GOT_X = 16, GOT_V = 256, GOT_Z = 65536, GOT_A = 1048576

  public :: file_type, read_species, read_config, write_config_fa, write_config_fb

contains

  character(6) function file_type(name)
    character(*), intent(in) :: name
    integer :: i, j
    j = len_trim(name)
    i = scan(name, '.', .true.)
    if (i == 0) then
      file_type = 'none'
    else if (j - i > 6) then
      file_type = 'long'
    else
      file_type = name(i + 1:j)
    end if
  end function file_type

  ! Skips '#' comment lines; returns the first other line.
  subroutine first_data_line(u, line)
    integer, intent(in) :: u
    character(256), intent(out) :: line
    line(1:1) = '#'
    do while (line(1:1) == '#')
      read(u, '(a)') line
    end do
  end subroutine first_data_line

  ! Species file of type ftype ('ZAfa' or 'ZAfb').  n: the particle count,
  ! set from the file when 0 on entry ('ZAfa') or always ('ZAfb'); zii, aii
  ! allocated here; ztot: the total charge ('ZAfb' sums it, 'ZAfa' leaves it
  ! as it was).  irtn: the bit field, -1 for an unknown type or no file.
  subroutine read_species(file, ftype, n, zii, aii, ztot, irtn)
    character(*), intent(in) :: file, ftype
    integer, intent(inout) :: n
    real(dp), allocatable, intent(inout) :: zii(:), aii(:)
    real(dp), intent(inout) :: ztot
    integer, intent(out) :: irtn
    character(256) :: line
    integer :: u, ios, i, j, k
    real(dp) :: xzi, xai
    irtn = -1
    open(newunit=u, file=trim(file), status='old', form='formatted', iostat=ios)
    if (ios /= 0) return
    select case (trim(ftype))
    case ('ZAfa')
      call first_data_line(u, line)
      read(line, '(i10)') k
      if (n == 0) n = k
      if (.not. allocated(zii)) allocate(zii(0:n - 1))
      if (.not. allocated(aii)) allocate(aii(0:n - 1))
      do i = 0, n - 1
        read(u, *) k, zii(k - 1), aii(k - 1)
      end do
      irtn = GOT_Z + GOT_A
    case ('ZAfb')
      call first_data_line(u, line)
      read(line, *) n
      if (.not. allocated(zii)) allocate(zii(0:n - 1))
      if (.not. allocated(aii)) allocate(aii(0:n - 1))
      i = 0
      ztot = 0.0d0
      do
        read(u, *, iostat=ios) k, xzi, xai
        if (ios /= 0) exit
        do j = 1, k
          zii(i) = xzi
          aii(i) = xai
          i = i + 1
        end do
        ztot = ztot + k * xzi
      end do
      irtn = GOT_Z + GOT_A
    end select
    close(u)
  end subroutine read_species

  ! Bytes of a file (for the revision-a types' particle count).
  integer(8) function file_bytes(file)
    character(*), intent(in) :: file
    integer(8) :: sz
    inquire(file=trim(file), size=sz)
    file_bytes = sz
  end function file_bytes

  ! Start file of type ftype.  n: the particle count (0 on entry: from the
  ! file where the type allows); time, xl, xl0, ev, ek, px, pp: what a
  ! revision-b file's header holds (left as they were otherwise); x, v, zii,
  ! aii: allocated here when absent; ztot: summed by the revision-b types
  ! with charges.  irtn: the bit field; -1 for an unknown or refused type,
  ! -2 when the file cannot be opened.
  subroutine read_config(file, ftype, n, time, xl, xl0, ev, ek, px, pp, x, v, zii, aii, &
                         ztot, irtn)
    character(*), intent(in) :: file, ftype
    integer, intent(inout) :: n
    real(dp), intent(inout) :: time, xl(3), xl0(3), ev, ek, px, pp(3, 3), ztot
    real(dp), allocatable, intent(inout) :: x(:, :), v(:, :), zii(:), aii(:)
    integer, intent(out) :: irtn
    character(256) :: line
    character(6) :: ft
    character(10) :: cname, daytime
    character(8) :: cver, date
    character(5) :: zone
    character(20) :: st
    real(dp) :: dummy
    real :: time4
    real, allocatable :: x4(:, :), v4(:, :)
    integer :: u, ios, i, k
    irtn = -1
    select case (trim(ftype))
    case ('onp', 'md1', 'xfa', 'xvfa', 'xvZAfa', 'xfb', 'xvfb', 'xZAfb', 'xvZAfb')
      open(newunit=u, file=trim(file), status='old', form='formatted', iostat=ios)
    case ('md0', 'x4a', 'xv4a', 'x8a', 'xv8a', 'x4b', 'x8b', 'xv4b', 'xv8b')
      open(newunit=u, file=trim(file), status='old', form='unformatted', iostat=ios)
    case default
      return
    end select
    if (ios /= 0) then
      irtn = -2
      return
    end if
    irtn = 0
    select case (trim(ftype))
    ! formatted revision a -------------------------------------------------
    case ('onp')
      if (.not. allocated(x)) allocate(x(3, 0:n - 1))
      do i = 0, n - 1
        read(u, *, iostat=ios) x(1, i), x(2, i), x(3, i), dummy
        if (ios /= 0) exit
      end do
      irtn = GOT_X
    case ('md1')
      if (.not. allocated(x)) allocate(x(3, 0:n - 1))
      if (.not. allocated(v)) allocate(v(3, 0:n - 1))
      do i = 0, n - 1
        read(u, *, iostat=ios) x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i)
        if (ios /= 0) exit
      end do
      irtn = GOT_X + GOT_V
    case ('xfa')
      if (.not. allocated(x)) allocate(x(3, 0:n - 1))
      read(u, *) time
      do i = 0, n - 1
        read(u, *, iostat=ios) x(1, i), x(2, i), x(3, i)
        if (ios /= 0) exit
      end do
      irtn = GOT_X
    case ('xvfa')
      if (.not. allocated(x)) allocate(x(3, 0:n - 1))
      if (.not. allocated(v)) allocate(v(3, 0:n - 1))
      read(u, *) time
      do i = 0, n - 1
        read(u, *, iostat=ios) x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i)
        if (ios /= 0) exit
      end do
      irtn = GOT_X + GOT_V
    case ('xvZAfa')
      call first_data_line(u, line)
      read(line, '(i10)') k
      if (n == 0) n = k
      if (.not. allocated(zii)) allocate(zii(0:n - 1))
      if (.not. allocated(aii)) allocate(aii(0:n - 1))
      if (.not. allocated(x)) allocate(x(3, 0:n - 1))
      if (.not. allocated(v)) allocate(v(3, 0:n - 1))
      read(u, *) time
      do i = 0, n - 1
        read(u, *, iostat=ios) x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i), &
                               zii(i), aii(i)
        if (ios /= 0) exit
      end do
      irtn = GOT_X + GOT_V + GOT_Z + GOT_A
    ! binary revision a ----------------------------------------------------
    case ('md0')
      if (n == 0) n = int((file_bytes(file) - 8) / (6 * 8))
      if (.not. allocated(x)) allocate(x(3, 0:n - 1))
      if (.not. allocated(v)) allocate(v(3, 0:n - 1))
      read(u) (x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i), i = 0, n - 1)
      irtn = GOT_X + GOT_V
    case ('x4a', 'xv4a')
      if (n == 0) then
        if (trim(ftype) == 'x4a') then
          n = int((file_bytes(file) - 12 - 8) / (3 * 4))
        else
          n = int((file_bytes(file) - 12 - 8) / (6 * 4))
        end if
      end if
      allocate(x4(3, 0:n - 1))
      if (.not. allocated(x)) allocate(x(3, 0:n - 1))
      read(u) time4
      time = time4
      if (trim(ftype) == 'x4a') then
        read(u) (x4(1, i), x4(2, i), x4(3, i), i = 0, n - 1)
        irtn = GOT_X
      else
        allocate(v4(3, 0:n - 1))
        if (.not. allocated(v)) allocate(v(3, 0:n - 1))
        read(u) (x4(1, i), x4(2, i), x4(3, i), v4(1, i), v4(2, i), v4(3, i), i = 0, n - 1)
        v = v4
        irtn = GOT_X + GOT_V
      end if
      x = x4
    case ('x8a')
      if (n == 0) n = int((file_bytes(file) - 16 - 8) / (3 * 8))
      if (.not. allocated(x)) allocate(x(3, 0:n - 1))
      read(u) time
      read(u) (x(1, i), x(2, i), x(3, i), i = 0, n - 1)
      irtn = GOT_X
    case ('xv8a')
      if (n == 0) n = int((file_bytes(file) - 16 - 8) / (6 * 8))
      if (.not. allocated(x)) allocate(x(3, 0:n - 1))
      if (.not. allocated(v)) allocate(v(3, 0:n - 1))
      read(u) time
      read(u) (x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i), i = 0, n - 1)
      irtn = GOT_X + GOT_V
    ! binary revision b ----------------------------------------------------
    case ('x4b', 'x8b', 'xv4b', 'xv8b')
      read(u) ft
      read(u) cname, cver
      read(u) date, daytime, zone
      read(u) st
      read(u) time, xl(1), xl(2), xl(3), ev, ek, px, pp, n
      if (.not. allocated(x)) allocate(x(3, 0:n - 1))
      select case (trim(ftype))
      case ('x4b')
        allocate(x4(3, 0:n - 1))
        read(u) (x4(1, i), x4(2, i), x4(3, i), i = 0, n - 1)
        x = x4
        irtn = GOT_X
      case ('xv4b')
        allocate(x4(3, 0:n - 1), v4(3, 0:n - 1))
        if (.not. allocated(v)) allocate(v(3, 0:n - 1))
        read(u) (x4(1, i), x4(2, i), x4(3, i), v4(1, i), v4(2, i), v4(3, i), i = 0, n - 1)
        x = x4
        v = v4
        irtn = GOT_X + GOT_V
      case ('x8b')
        read(u) (x(1, i), x(2, i), x(3, i), i = 0, n - 1)
        irtn = GOT_X
      case ('xv8b')
        if (.not. allocated(v)) allocate(v(3, 0:n - 1))
        read(u) (x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i), i = 0, n - 1)
        irtn = GOT_X + GOT_V
      end select
    ! formatted revision b: six list-directed header records ---------------
    case ('xfb', 'xvfb', 'xZAfb', 'xvZAfb')
      read(u, *) ft
      read(u, *) cname, cver
      read(u, *) date, daytime, zone
      read(u, *) st
      read(u, *) xl0(1), xl0(2), xl0(3)
      read(u, *) time, xl(1), xl(2), xl(3), ev, ek, px, pp, n
      if (.not. allocated(x)) allocate(x(3, 0:n - 1))
      select case (trim(ftype))
      case ('xfb')
        do i = 0, n - 1
          read(u, *, iostat=ios) x(1, i), x(2, i), x(3, i)
          if (ios /= 0) exit
        end do
        irtn = GOT_X
      case ('xvfb')
        if (.not. allocated(v)) allocate(v(3, 0:n - 1))
        do i = 0, n - 1
          read(u, *, iostat=ios) x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i)
          if (ios /= 0) exit
        end do
        irtn = GOT_X + GOT_V
      case ('xZAfb')
        if (.not. allocated(zii)) allocate(zii(0:n - 1))
        if (.not. allocated(aii)) allocate(aii(0:n - 1))
        ztot = 0.0d0
        do i = 0, n - 1
          read(u, *, iostat=ios) x(1, i), x(2, i), x(3, i), zii(i), aii(i)
          if (ios /= 0) exit
          ztot = ztot + zii(i)
        end do
        irtn = GOT_X + GOT_Z + GOT_A
      case ('xvZAfb')
        if (.not. allocated(v)) allocate(v(3, 0:n - 1))
        if (.not. allocated(zii)) allocate(zii(0:n - 1))
        if (.not. allocated(aii)) allocate(aii(0:n - 1))
        ztot = 0.0d0
        do i = 0, n - 1
          read(u, *, iostat=ios) x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i), &
                                 zii(i), aii(i)
          if (ios /= 0) exit
          ztot = ztot + zii(i)
        end do
        irtn = GOT_X + GOT_V + GOT_Z + GOT_A
      end select
    end select
    close(u)
  end subroutine read_config

  ! Opens a formatted file for one more configuration (appended when append).
  subroutine open_formatted(file, append, u)
    character(*), intent(in) :: file
    logical, intent(in) :: append
    integer, intent(out) :: u
    if (append) then
      open(newunit=u, file=trim(file), status='unknown', form='formatted', position='append')
    else
      open(newunit=u, file=trim(file), status='unknown', form='formatted')
    end if
  end subroutine open_formatted

  ! Formatted revision a (src/rw_xvfa.f:122-178); 'onp' writes nothing.
  subroutine write_config_fa(file, ftype, append, time, n, x, v, zii, aii)
    character(*), intent(in) :: file, ftype
    logical, intent(in) :: append
    real(dp), intent(in) :: time
    integer, intent(in) :: n
    real(dp), intent(in) :: x(3, 0:n - 1), v(3, 0:n - 1), zii(0:n - 1), aii(0:n - 1)
    integer :: u, i
    call open_formatted(file, append, u)
    select case (trim(ftype))
    case ('md1')
      write(u, '(es24.16,es24.16,es24.16,es24.16,es24.16,es24.16)') &
        (x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i), i = 0, n - 1)
    case ('xfa')
      write(u, '(f15.3)') time
      write(u, '(es24.16,es24.16,es24.16,es24.16,es24.16,es24.16)') &
        (x(1, i), x(2, i), x(3, i), i = 0, n - 1)
    case ('xvfa')
      write(u, '(f15.3)') time
      write(u, '(es24.16,es24.16,es24.16,es24.16,es24.16,es24.16)') &
        (x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i), i = 0, n - 1)
    case ('xvZAfa')
      write(u, '(f15.3)') time
      write(u, '(es24.16,es24.16,es24.16,es24.16,es24.16,es24.16,1x,f8.3,1x,f8.3)') &
        (x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i), zii(i), aii(i), i = 0, n - 1)
    end select
    close(u)
  end subroutine write_config_fa

  ! Formatted revision b (src/rw_xvfb.f:210-288): a '# ' header, then the
  ! particles.
  subroutine write_config_fb(file, ftype, append, sim_type, time, xl0, xl, ev, ek, px, pp, &
                             n, x, v, zii, aii)
    character(*), intent(in) :: file, ftype, sim_type
    logical, intent(in) :: append
    real(dp), intent(in) :: time, xl0(3), xl(3), ev, ek, px, pp(3, 3)
    integer, intent(in) :: n
    real(dp), intent(in) :: x(3, 0:n - 1), v(3, 0:n - 1), zii(0:n - 1), aii(0:n - 1)
    character(8) :: date
    character(10) :: daytime
    character(5) :: zone
    integer :: u, i
    call open_formatted(file, append, u)
    call date_and_time(date, daytime, zone)
    write(u, '(''# '',a,2x,a,2x,a)') trim(ftype)
    write(u, '(''# '',a,2x,a,2x,a)') 'MD', '6.2.0'
    write(u, '(''# '',a,2x,a,2x,a)') trim(date), trim(daytime), trim(zone)
    write(u, '(''# '',a,2x,a,2x,a)') trim(sim_type)
    write(u, '(''#'',/,''# '',3(1x,f12.6))') xl0(1), xl0(2), xl0(3)
    write(u, '(''#'',/,''# '',3(1x,f12.6))') time
    write(u, '(''#'',/,''# '',3(1x,f12.6))') xl(1), xl(2), xl(3)
    write(u, '(''#'',/,''# '',3(1x,es15.8))') ev, ek, px
    write(u, '(''#'',/,2(''# '',3(1x,es15.8),/),''# '',3(1x,es15.8))') pp
    write(u, '(''# '',i10)') n
    select case (trim(ftype))
    case ('xfb')
      write(u, '(es24.16,es24.16,es24.16)') (x(1, i), x(2, i), x(3, i), i = 0, n - 1)
    case ('xvfb')
      write(u, '(es24.16,es24.16,es24.16,es24.16,es24.16,es24.16)') &
        (x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i), i = 0, n - 1)
    case ('xZAfb')
      write(u, '(es24.16,es24.16,es24.16,1x,f8.3,1x,f8.3)') &
        (x(1, i), x(2, i), x(3, i), zii(i), aii(i), i = 0, n - 1)
    case ('xvZAfb')
      write(u, '(es24.16,es24.16,es24.16,es24.16,es24.16,es24.16,1x,f8.3,1x,f8.3)') &
        (x(1, i), x(2, i), x(3, i), v(1, i), v(2, i), v(3, i), zii(i), aii(i), i = 0, n - 1)
    end select
    close(u)
  end subroutine write_config_fb

end module yk_config
