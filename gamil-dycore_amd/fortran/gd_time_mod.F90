! gd_time_mod -- the time manager and output alerts of the reference
! (time_mod.F90:48-215, io_mod.F90:86-216) without its un-vendored datetime
! and hash-table libraries.
!
! Times are seconds relative to the first start time of the run (the value of
! start_time when time_init runs), held in real(8); calendar dates are
! proleptic Gregorian, converted with the days-from-civil / civil-from-days
! algorithms of H. Hinnant ("chrono-Compatible Low-Level Date Algorithms").
!
! Reference behaviour kept here:
!   - start_time = 0001-01-01T00:00 when run_days/hours/minutes is set (the
!     datetime default date), else the namelist start_time; end_time likewise
!     from the run length (run_seconds is not part of it) or end_time
!     (time_mod.F90:57-68);
!   - the loop runs while curr_time < end_time (time_mod.F90:138-142);
!   - an alert rings when last_time + period <= curr_time; the next
!     time_advance sets its last_time to the time at which it rang
!     (time_mod.F90:95-110, 196-215); alerts start at the start time of
!     time_init, so after a restart (time_reset_start_time) they ring at once;
!   - IO periods "<value> days|hours|minutes|seconds|steps" or "once" (period
!     0: rings at every output check) (io_mod.F90:192-216);
!   - time values are elapsed seconds over time_units_in_seconds, with
!     'seconds' -> 60 as in io_mod.F90:93-94;
!   - file time stamps: '%Y-%m-%dT%H_%M_%S' at the start, ISO 8601
!     '%Y-%m-%dT%H:%M:%S' after the first time_advance (time_mod.F90:79-80,
!     116).
module gd_time_mod

  use gd_params_mod

  implicit none

  private

  public time_init
  public time_reset_start_time
  public time_advance
  public time_is_finished
  public time_add_alert
  public time_is_alerted
  public time_alert_would_ring
  public time_units_in_seconds
  public io_period_seconds
  public format_time
  public parse_time
  public time_elapsed_seconds
  public time_value
  public curr_time_format
  public start_time_format
  public time_step
  public time_curr
  public time_end

  integer, parameter :: max_alerts = 8

  type alert_type
    character(30) :: name = ''
    real(r8) :: period = 0
    real(r8) :: last_time = 0
    logical :: ring = .false.
  end type alert_type

  type(alert_type) alerts(max_alerts)
  integer :: n_alerts = 0

  integer(8) :: epoch = 0            ! absolute seconds (since 1970-01-01) of time 0
  real(r8) :: time_start = 0         ! relative to epoch
  real(r8) :: time_curr = 0
  real(r8) :: time_end = 0
  real(r8) :: elapsed_seconds = 0
  integer :: time_step = 0
  character(30) :: start_time_format = ''
  character(30) :: curr_time_format = ''

contains

  ! Days since 1970-01-01 of a proleptic Gregorian date.
  integer(8) function days_from_civil(y0, m, d) result(res)

    integer, intent(in) :: y0, m, d

    integer(8) y, era, yoe, doy, doe, mp

    y = y0
    if (m <= 2) y = y - 1
    if (y >= 0) then
      era = y / 400
    else
      era = (y - 399) / 400
    end if
    yoe = y - era * 400
    if (m > 2) then
      mp = m - 3
    else
      mp = m + 9
    end if
    doy = (153 * mp + 2) / 5 + d - 1
    doe = yoe * 365 + yoe / 4 - yoe / 100 + doy
    res = era * 146097 + doe - 719468

  end function days_from_civil

  subroutine civil_from_days(z0, y, m, d)

    integer(8), intent(in) :: z0
    integer, intent(out) :: y, m, d

    integer(8) z, era, doe, yoe, doy, mp

    z = z0 + 719468
    if (z >= 0) then
      era = z / 146097
    else
      era = (z - 146096) / 146097
    end if
    doe = z - era * 146097
    yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365
    doy = doe - (365 * yoe + yoe / 4 - yoe / 100)
    mp = (5 * doy + 2) / 153
    d = int(doy - (153 * mp + 2) / 5 + 1)
    if (mp < 10) then
      m = int(mp + 3)
    else
      m = int(mp - 9)
    end if
    y = int(yoe + era * 400)
    if (m <= 2) y = y + 1

  end subroutine civil_from_days

  ! kind 1: '%Y-%m-%dT%H_%M_%S'; kind 2: '%Y-%m-%dT%H:%M:%S'
  function format_time(t, kind) result(res)

    real(r8), intent(in) :: t
    integer, intent(in) :: kind
    character(30) res

    integer(8) s, days, sod
    integer y, m, d, hh, mm, ss
    character sep

    s = epoch + int(floor(t + 1.0e-6_r8), 8)
    days = s / 86400
    sod = s - days * 86400
    if (sod < 0) then
      sod = sod + 86400
      days = days - 1
    end if
    call civil_from_days(days, y, m, d)
    hh = int(sod / 3600)
    mm = int(mod(sod, 3600_8) / 60)
    ss = int(mod(sod, 60_8))
    sep = merge('_', ':', kind == 1)
    write(res, '(I4.4, "-", I2.2, "-", I2.2, "T", I2.2, A1, I2.2, A1, I2.2)') y, m, d, hh, sep, mm, sep, ss

  end function format_time

  ! Inverse of format_time for either separator; returns .false. on bad text.
  logical function parse_time(text, t) result(ok)

    character(*), intent(in) :: text
    real(r8), intent(out) :: t

    integer y, m, d, hh, mm, ss, ios
    character(30) s

    s = adjustl(text)
    ok = .false.
    t = 0
    if (len_trim(s) < 19) return
    if (s(5:5) /= '-' .or. s(8:8) /= '-' .or. s(11:11) /= 'T') return
    read(s, '(I4, 1X, I2, 1X, I2, 1X, I2, 1X, I2, 1X, I2)', iostat=ios) y, m, d, hh, mm, ss
    if (ios /= 0) return
    if (m < 1 .or. m > 12 .or. d < 1 .or. d > 31 .or. hh > 23 .or. mm > 59 .or. ss > 60) return
    t = real(days_from_civil(y, m, d) * 86400_8 + hh * 3600 + mm * 60 + ss - epoch, r8)
    ok = .true.

  end function parse_time

  subroutine time_init()

    integer(8) base

    base = days_from_civil(1, 1, 1) * 86400_8
    if (run_days > 0 .or. run_hours > 0 .or. run_minutes > 0) then
      epoch = base
      time_end = run_days * 86400.0_r8 + run_hours * 3600.0_r8 + run_minutes * 60.0_r8
    else if (sum(start_time) > 0) then
      epoch = days_from_civil(start_time(1), max(start_time(2), 1), max(start_time(3), 1)) * 86400_8 + &
              start_time(4) * 3600_8 + start_time(5) * 60_8
      time_end = 0
      if (sum(end_time) > 0) &
        time_end = real(days_from_civil(end_time(1), max(end_time(2), 1), max(end_time(3), 1)) * 86400_8 + &
                        end_time(4) * 3600_8 + end_time(5) * 60_8 - epoch, r8)
    else
      epoch = base
      time_end = 0
    end if
    time_start = 0
    time_curr = 0
    elapsed_seconds = 0
    time_step = 0
    n_alerts = 0
    start_time_format = format_time(time_start, 1)
    curr_time_format = start_time_format

  end subroutine time_init

  subroutine time_reset_start_time(t)

    real(r8), intent(in) :: t

    time_start = t
    time_curr = t
    start_time_format = format_time(time_start, 1)
    curr_time_format = start_time_format

  end subroutine time_reset_start_time

  subroutine time_advance()

    integer k

    do k = 1, n_alerts
      if (alerts(k)%ring) then
        alerts(k)%last_time = time_curr
        alerts(k)%ring = .false.
      end if
    end do
    time_step = time_step + 1
    elapsed_seconds = elapsed_seconds + time_step_size
    time_curr = time_curr + time_step_size
    curr_time_format = format_time(time_curr, 2)

  end subroutine time_advance

  logical function time_is_finished() result(res)

    res = time_curr >= time_end - 1.0e-6_r8

  end function time_is_finished

  real(r8) function time_elapsed_seconds() result(res)

    res = elapsed_seconds

  end function time_elapsed_seconds

  ! Seconds per time unit of the history time axis (io_mod.F90:86-98);
  ! returns a negative value for an invalid unit.
  real(r8) function time_units_in_seconds(units) result(res)

    character(*), intent(in) :: units

    select case (trim(units))
    case ('days')
      res = 86400.0_r8
    case ('hours')
      res = 3600.0_r8
    case ('seconds')
      res = 60.0_r8
    case default
      res = -1.0_r8
    end select

  end function time_units_in_seconds

  real(r8) function time_value(units) result(res)

    character(*), intent(in) :: units

    res = elapsed_seconds / time_units_in_seconds(units)

  end function time_value

  ! IO period in seconds (io_mod.F90:192-214); ok = .false. for an invalid period.
  real(r8) function io_period_seconds(period, ok) result(res)

    character(*), intent(in) :: period
    logical, intent(out) :: ok

    character(30) p, value_text, units
    integer k, ios

    ok = .true.
    res = 0
    p = adjustl(period)
    if (trim(p) == 'once') return
    k = index(trim(p), ' ')
    if (k == 0) then
      ok = .false.
      return
    end if
    value_text = p(1:k-1)
    units = adjustl(p(k+1:))
    k = index(trim(units), ' ')
    if (k > 0) units = units(1:k-1)
    read(value_text, *, iostat=ios) res
    if (ios /= 0) then
      ok = .false.
      return
    end if
    select case (trim(units))
    case ('days')
      res = res * 86400
    case ('hours')
      res = res * 3600
    case ('minutes')
      res = res * 60
    case ('seconds')
      res = res
    case ('steps')
      res = res * time_step_size
    case default
      ok = .false.
    end select

  end function io_period_seconds

  ! Alerts begin at the start time of time_init (time_mod.F90:181-183).
  subroutine time_add_alert(name, seconds)

    character(*), intent(in) :: name
    real(r8), intent(in) :: seconds

    n_alerts = n_alerts + 1
    alerts(n_alerts)%name = name
    alerts(n_alerts)%period = seconds
    alerts(n_alerts)%last_time = 0
    alerts(n_alerts)%ring = .false.

  end subroutine time_add_alert

  integer function find_alert(name) result(k)

    character(*), intent(in) :: name

    do k = 1, n_alerts
      if (alerts(k)%name == name) return
    end do
    k = 0

  end function find_alert

  logical function time_is_alerted(name) result(res)

    character(*), intent(in) :: name

    integer k

    k = find_alert(name)
    res = .false.
    if (k == 0) return
    if (alerts(k)%last_time + alerts(k)%period <= time_curr + 1.0e-6_r8) then
      alerts(k)%ring = .true.
      res = .true.
    end if

  end function time_is_alerted

  ! time_is_alerted without ringing: whether any alert would ring now.
  logical function time_alert_would_ring() result(res)

    integer k

    res = .false.
    do k = 1, n_alerts
      if (alerts(k)%last_time + alerts(k)%period <= time_curr + 1.0e-6_r8) res = .true.
    end do

  end function time_alert_would_ring

end module gd_time_mod
