! dycore_mod -- the reference's dycore facade (src/dycore_mod.F90:22-25:
! dycore_init, dycore_restart, dycore_run, dycore_final) over the MI355X
! library.  The numerics run on the device through the C ABI; this module
! keeps the reference's conventions: module-global state written by the IC
! code (state(1)%u/v/gd and static%ghs become state_u/v/gd and static_ghs,
! halo'd (-1:N+2,-1:M+2) like parallel_mod allocates them), one log line per
! step, [Error] + stop 1 on failure (log_mod.F90:66-79), and the history and
! restart files at the alerts of the time manager (dycore_mod.F90:175-182).
module dycore_mod

  use iso_c_binding
  use gd_params_mod
  use gd_capi_mod
  use gd_time_mod

  implicit none

  private

  public dycore_init
  public dycore_restart
  public dycore_run
  public dycore_final
  public dycore_describe
  public dycore_schedule
  public state_u, state_v, state_gd, static_ghs
  public final_mass, final_energy, final_beta, members

  real(r8), allocatable, target :: state_u(:,:,:)    ! (-1:N+2, -1:M+2, member)
  real(r8), allocatable, target :: state_v(:,:,:)
  real(r8), allocatable, target :: state_gd(:,:,:)
  real(r8), allocatable, target :: static_ghs(:,:,:)
  real(r8) :: final_mass = 0, final_energy = 0, final_beta = 0
  integer :: members = 1

  type(c_ptr) :: handle = c_null_ptr
  integer(c_int32_t), allocatable, target :: cutoff(:)
  integer :: n_cutoff = 0

  ! steps per gd_step call when no output falls in between
  integer, parameter :: chunk = 720

  ! NUL-terminated attribute strings handed to gd_io_info
  character(kind=c_char, len=512), target :: s_desc, s_author, s_units, s_start, s_scheme, &
                                             s_split, s_rtime

contains

  subroutine log_error(msg)

    character(*), intent(in) :: msg

    write(6, '("[Error]: ", A)') trim(msg)
    stop 1

  end subroutine log_error

  subroutine log_notice(msg)

    character(*), intent(in) :: msg

    write(6, '("[Notice]: ", A)') trim(msg)

  end subroutine log_notice

  ! Mapping of the string options (dycore_mod.F90:78-107) and the filter table.
  subroutine build_config(cfg)

    type(gd_config), intent(out) :: cfg

    integer k, n

    cfg%num_lon = num_lon
    cfg%num_lat = num_lat
    cfg%time_step_size = time_step_size
    select case (split_scheme)
    case ('csp1')
      cfg%split_scheme = 1
    case ('csp2')
      cfg%split_scheme = 2
    case ('isp')
      cfg%split_scheme = 3
    case default
      cfg%split_scheme = 0
    end select
    cfg%subcycles = subcycles
    cfg%qcon_modified = merge(1, 0, qcon_modified)
    select case (uv_adv_scheme)
    case ('center_diff')
      cfg%uv_adv_scheme = 0
    case ('upwind')
      cfg%uv_adv_scheme = 1
    case ('weno')
      cfg%uv_adv_scheme = 2
    case default
      call log_error('Unknown uv_adv_scheme ' // trim(uv_adv_scheme) // '!')
    end select
    cfg%uv_adv_upwind_lon_beta = uv_adv_upwind_lon_beta
    cfg%uv_adv_upwind_lat_beta = uv_adv_upwind_lat_beta
    cfg%use_diffusion = merge(1, 0, use_diffusion)
    cfg%diffusion_order = diffusion_order
    cfg%diffusion_coef = diffusion_coef
    cfg%use_zonal_tend_filter = merge(1, 0, use_zonal_tend_filter)
    if (allocated(cutoff)) deallocate(cutoff)
    if (zonal_tend_filter_lat0 > 0.0_r8) then
      allocate(cutoff(max(num_lat, 1)))
      n = gd_cutoff_from_lat0(int(num_lon, c_int32_t), int(num_lat, c_int32_t), &
                              real(zonal_tend_filter_lat0, c_double), cutoff, int(num_lat, c_int32_t))
      if (n < 0) call log_error('Bad zonal_tend_filter_lat0!')
      n_cutoff = n
    else
      n_cutoff = 0
      do k = 1, max_cutoff
        if (zonal_tend_filter_cutoff_wavenumber(k) /= 0) n_cutoff = k
      end do
      allocate(cutoff(max(n_cutoff, 1)))
      cutoff = 0
      if (n_cutoff > 0) cutoff(1:n_cutoff) = zonal_tend_filter_cutoff_wavenumber(1:n_cutoff)
    end if
    cfg%n_cutoff = n_cutoff
    cfg%cutoff = c_loc(cutoff)
    select case (precision)
    case ('fp64')
      cfg%precision = 0
    case ('mixed')
      cfg%precision = 1
    case ('mixed_bf16')
      cfg%precision = 2
    case ('mixed_mfma')
      cfg%precision = 3
    case default
      call log_error('Unknown precision ' // trim(precision) // '!')
    end select
    cfg%device = device

  end subroutine build_config

  subroutine dycore_describe()

    type(gd_config) cfg

    call build_config(cfg)
    write(6, '("case_name        = ", A)') trim(case_name)
    write(6, '("grid             = ", I0, " x ", I0)') num_lon, num_lat
    write(6, '("time_step_size   = ", F0.3)') time_step_size
    write(6, '("steps            = ", I0)') total_steps()
    write(6, '("split_scheme     = ", I0, " (", A, "), subcycles = ", I0)') cfg%split_scheme, &
      trim(split_scheme), subcycles
    write(6, '("qcon_modified    = ", L1)') qcon_modified
    write(6, '("uv_adv_scheme    = ", I0)') cfg%uv_adv_scheme
    write(6, '("use_diffusion    = ", L1, ", order ", I0, ", coef ", ES12.5)') use_diffusion, &
      diffusion_order, diffusion_coef
    write(6, '("filter entries   = ", I0)') n_cutoff
    if (n_cutoff > 0) write(6, '("filter cutoffs   = ", *(I0, 1X))') cutoff(1:n_cutoff)
    write(6, '("precision        = ", A)') trim(precision)
    write(6, '("members          = ", I0)') ensemble_members

  end subroutine dycore_describe

  ! dycore_init (dycore_mod.F90:60-111)
  subroutine dycore_init()

    type(gd_config) cfg
    integer(c_int) rc

    if (case_name == '') call log_error('case_name is not set!')
    if (time_scheme /= 'predict_correct') call log_error('Unknown time_scheme ' // trim(time_scheme) // '!')
    if (num_lon < 8 .or. num_lat < 5) call log_error('num_lon/num_lat not set!')
    members = max(1, ensemble_members)
    call build_config(cfg)
    rc = gd_create(cfg, int(members, c_int32_t), handle)
    if (rc /= GD_OK) call log_error(gd_error_text(c_null_ptr))
    allocate(state_u(-1:num_lon+2, -1:num_lat+2, members))
    allocate(state_v(-1:num_lon+2, -1:num_lat+2, members))
    allocate(state_gd(-1:num_lon+2, -1:num_lat+2, members))
    allocate(static_ghs(-1:num_lon+2, -1:num_lat+2, members))
    state_u = 0.0_r8
    state_v = 0.0_r8
    state_gd = 0.0_r8
    static_ghs = 0.0_r8
    if (split_scheme /= 'csp1' .and. split_scheme /= 'csp2' .and. split_scheme /= 'isp') &
      call log_notice('No fast-slow split.')
    call init_time_and_alerts()
    call log_notice('Dycore module is initialized.')

  end subroutine dycore_init

  ! time_init, io_init and the output alerts of history_init / restart_init
  ! (dycore_mod.F90:68-73, io_mod.F90:86-98, 192-216).
  subroutine init_time_and_alerts()

    real(r8) p
    logical ok

    call time_init()
    if (time_units_in_seconds(time_units) < 0) call log_error('Invalid time_units ' // trim(time_units) // '!')
    p = io_period_seconds(history_periods(1), ok)
    if (.not. ok) call log_error('Invalid IO period ' // trim(history_periods(1)) // '!')
    call time_add_alert('hist0.output', p)
    p = io_period_seconds(restart_period, ok)
    if (.not. ok) call log_error('Invalid IO period ' // trim(restart_period) // '!')
    call time_add_alert('restart.output', p)

  end subroutine init_time_and_alerts

  ! dycore_restart (dycore_mod.F90:113-117, restart_mod.F90:37-55): state and
  ! ghs from restart_file, start time reset to its restart_time.
  subroutine dycore_restart()

    integer(c_int) rc
    integer(c_int64_t) ld
    character(kind=c_char) rt(64)
    real(c_double) el
    real(r8) t
    character(64) text
    integer k

    if (members > 1) call log_error('restart_file with ensemble_members > 1 is not provided!')
    ld = num_lon + 4
    rt = c_null_char
    rc = gd_restart_read(c_string(restart_file), int(num_lon, c_int32_t), int(num_lat, c_int32_t), &
                         state_u(1, 1, 1), state_v(1, 1, 1), state_gd(1, 1, 1), static_ghs(1, 1, 1), &
                         ld, rt, 64_c_int32_t, el)
    if (rc /= GD_OK) call log_error(f_string(gd_io_last_error(), 512))
    text = ''
    do k = 1, 64
      if (rt(k) == c_null_char) exit
      text(k:k) = rt(k)
    end do
    if (.not. parse_time(text, t)) call log_error('Bad restart_time "' // trim(text) // '" in ' // trim(restart_file))
    call time_reset_start_time(t)
    call log_notice('Reset time to ' // trim(curr_time_format) // '.')
    call fill_halo(state_u(:, :, 1))
    call fill_halo(state_v(:, :, 1))
    call fill_halo(state_gd(:, :, 1))
    call fill_halo(static_ghs(:, :, 1))

  end subroutine dycore_restart

  ! Advance the time manager over the steps of the next gd_step call: up to
  ! the end of the run, the next step at which an alert rings, or chunk steps.
  integer function next_chunk() result(n)

    n = 0
    do
      call time_advance()
      n = n + 1
      if (time_is_finished() .or. time_alert_would_ring() .or. n == chunk) exit
    end do

  end function next_chunk

  function file_prefix(m) result(res)

    integer, intent(in) :: m
    character(512) res

    res = trim(case_name)
    if (members > 1) write(res, '(A, ".m", I3.3)') trim(case_name), m
    if (output_dir /= '') res = trim(output_dir) // '/' // trim(res)

  end function file_prefix

  ! <case_name>.h0 + '.h0' (io_mod.F90:182-184, history_mod.F90:34) and
  ! <case_name>.r, each followed by the current time stamp (io_mod.F90:417-421).
  function history_path(m) result(res)

    integer, intent(in) :: m
    character(512) res

    res = trim(file_prefix(m)) // '.h0.h0.' // trim(curr_time_format) // '.nc'

  end function history_path

  function restart_path(m) result(res)

    integer, intent(in) :: m
    character(512) res

    res = trim(file_prefix(m)) // '.r.' // trim(curr_time_format) // '.nc'

  end function restart_path

  subroutine set_c(dst, src)

    character(kind=c_char, len=512), intent(out) :: dst
    character(*), intent(in) :: src

    dst = trim(src) // c_null_char

  end subroutine set_c

  function io_info(restart) result(info)

    logical, intent(in) :: restart
    type(gd_io_info) info

    if (restart) then
      call set_c(s_desc, 'Restart file')
    else
      call set_c(s_desc, case_desc)
    end if
    call set_c(s_author, 'N/A')
    call set_c(s_units, time_units)
    call set_c(s_start, start_time_format)
    call set_c(s_scheme, time_scheme)
    call set_c(s_split, split_scheme)
    call set_c(s_rtime, curr_time_format)
    info%desc = c_loc(s_desc)
    info%author = c_loc(s_author)
    info%time_units = c_loc(s_units)
    info%start_time = c_loc(s_start)
    info%time_value = time_value(time_units)
    info%time_step_size = time_step_size
    info%time_scheme = c_loc(s_scheme)
    info%split_scheme = c_loc(s_split)
    info%subcycles = subcycles
    info%restart_time = c_loc(s_rtime)
    info%elapsed_seconds = time_elapsed_seconds()

  end function io_info

  ! output (dycore_mod.F90:175-182): history and restart files of every
  ! member when their alerts ring; mass/energy(member) are the diagnostics of
  ! the current state.
  subroutine output(mass, energy)

    real(r8), intent(in) :: mass(:), energy(:)

    logical hist, rest
    integer(c_int) rc
    integer(c_int64_t) ld
    integer m
    type(gd_io_info) info
    real(r8), allocatable :: div(:,:), vor(:,:)

    hist = time_is_alerted('hist0.output')
    rest = time_is_alerted('restart.output')
    if (.not. write_output .or. .not. (hist .or. rest)) return
    ld = num_lon + 4
    allocate(div(num_lon + 4, num_lat), vor(num_lon + 4, num_lat))
    do m = 1, members
      rc = gd_get_state(handle, int(m - 1, c_int32_t), state_u(1, 1, m), state_v(1, 1, m), &
                        state_gd(1, 1, m), ld)
      if (rc /= GD_OK) call log_error(gd_error_text(handle))
      if (hist) then
        rc = gd_get_diag_fields(handle, int(m - 1, c_int32_t), div, vor, ld)
        if (rc /= GD_OK) call log_error(gd_error_text(handle))
        info = io_info(.false.)
        rc = gd_history_write(c_string(history_path(m)), info, int(num_lon, c_int32_t), &
                              int(num_lat, c_int32_t), state_u(1, 1, m), state_v(1, 1, m), &
                              state_gd(1, 1, m), static_ghs(1, 1, m), div, vor, ld, energy(m), mass(m))
        if (rc /= GD_OK) call log_error(f_string(gd_io_last_error(), 512))
      end if
      if (rest) then
        info = io_info(.true.)
        rc = gd_restart_write(c_string(restart_path(m)), info, int(num_lon, c_int32_t), &
                              int(num_lat, c_int32_t), state_u(1, 1, m), state_v(1, 1, m), &
                              state_gd(1, 1, m), static_ghs(1, 1, m), ld)
        if (rc /= GD_OK) call log_error(f_string(gd_io_last_error(), 512))
      end if
    end do

  end subroutine output

  ! The output schedule of a run without the dynamics: one line per file the
  ! run would write ("history|restart <step> <time value> <path>") and the
  ! step count.  Uses the same time manager calls as dycore_run.
  subroutine dycore_schedule()

    integer n, done

    members = max(1, ensemble_members)
    call init_time_and_alerts()
    if (is_restart_run) call log_notice('Schedule of a restart run starts at time 0 of the namelist.')
    done = 0
    call schedule_check(done)
    do while (.not. time_is_finished())
      n = next_chunk()
      done = done + n
      call schedule_check(done)
    end do
    write(6, '("steps = ", I0)') done

  end subroutine dycore_schedule

  subroutine schedule_check(step)

    integer, intent(in) :: step

    logical hist, rest
    integer m

    hist = time_is_alerted('hist0.output')
    rest = time_is_alerted('restart.output')
    if (.not. write_output) return
    do m = 1, members
      if (hist) write(6, '("history ", I0, 1X, ES24.17, 1X, A)') step, time_value(time_units), &
        trim(history_path(m))
      if (rest) write(6, '("restart ", I0, 1X, ES24.17, 1X, A)') step, time_value(time_units), &
        trim(restart_path(m))
    end do

  end subroutine schedule_check

  subroutine fill_halo(x)

    real(r8), intent(inout) :: x(-1:, -1:)

    x(-1, :) = x(num_lon - 1, :)
    x(0, :) = x(num_lon, :)
    x(num_lon + 1, :) = x(1, :)
    x(num_lon + 2, :) = x(2, :)

  end subroutine fill_halo

  ! The autotuner (no counterpart in the reference): launch shapes that
  ! cannot change a result, timed from the state just handed over, which it
  ! leaves as it was.
  subroutine tune_launch_shapes()

    integer(c_int) rc
    integer(c_int32_t) fused, upd_wpm
    real(c_double) ms
    character(kind=c_char, len=512) path

    call set_c(path, autotune_cache)
    rc = gd_autotune(handle, int(min(autotune_steps, 512), c_int32_t), path, fused, upd_wpm, ms)
    if (rc /= GD_OK) call log_error(gd_error_text(handle))
    write(6, '("autotune: fused = ", I0, ", update workgroups per member = ", I0, ", ", F10.4, " ms per step")') &
      fused, upd_wpm, ms

  end subroutine tune_launch_shapes

  subroutine log_step(step, mass, energy, beta)

    integer, intent(in) :: step
    real(r8), intent(in) :: mass, energy, beta

    integer(8) t
    integer d, hh, mm, ss

    t = nint(step * time_step_size, 8)
    d = int(t / 86400)
    hh = int(mod(t, 86400_8) / 3600)
    mm = int(mod(t, 3600_8) / 60)
    ss = int(mod(t, 60_8))
    write(6, '(" => ", I4.4, "T", I2.2, ":", I2.2, ":", I2.2, 3(1X, ES27.20))') d, hh, mm, ss, &
      mass, energy, beta

  end subroutine log_step

  ! dycore_run (dycore_mod.F90:119-142): iap_transform + diag_run + output of
  ! the initial state, then the time loop (time_integrate, time_advance,
  ! diag_run, output, log_step) on the device in calls of up to chunk steps
  ! that end where an output alert rings.
  subroutine dycore_run()

    integer(c_int) rc
    integer(c_int32_t) nan_step
    integer m, k, done, n
    integer(c_int64_t) ld
    real(r8), allocatable :: mass(:), energy(:), beta(:)

    ld = num_lon + 4
    do m = 1, members
      rc = gd_set_state(handle, int(m - 1, c_int32_t), state_u(1, 1, m), state_v(1, 1, m), &
                        state_gd(1, 1, m), static_ghs(1, 1, m), ld)
      if (rc /= GD_OK) call log_error(gd_error_text(handle))
    end do
    allocate(mass(chunk * members), energy(chunk * members), beta(chunk * members))
    if (autotune_steps > 0) call tune_launch_shapes()
    rc = gd_diag(handle, mass, energy)
    if (rc /= GD_OK) call log_error(gd_error_text(handle))
    call output(mass(1:members), energy(1:members))
    call log_step(0, mass(1), energy(1), 1.0_r8)
    final_mass = mass(1)
    final_energy = energy(1)
    final_beta = 1.0_r8

    done = 0
    do while (.not. time_is_finished())
      n = next_chunk()
      rc = gd_step(handle, int(n, c_int32_t), mass, energy, beta, nan_step)
      if (rc == GD_ERR_NAN) then
        do k = 1, nan_step - 1
          if (mod(done + k, log_every) == 0) call log_step(done + k, mass((k - 1) * members + 1), &
                                                           energy((k - 1) * members + 1), beta((k - 1) * members + 1))
        end do
        call log_error('Total mass or energy is NaN!')
      else if (rc /= GD_OK) then
        call log_error(gd_error_text(handle))
      end if
      call output(mass((n - 1) * members + 1:n * members), energy((n - 1) * members + 1:n * members))
      do k = 1, n
        if (mod(done + k, log_every) == 0 .or. time_is_finished() .and. k == n) &
          call log_step(done + k, mass((k - 1) * members + 1), energy((k - 1) * members + 1), &
                        beta((k - 1) * members + 1))
      end do
      final_mass = mass((n - 1) * members + 1)
      final_energy = energy((n - 1) * members + 1)
      final_beta = beta((n - 1) * members + 1)
      done = done + n
    end do

    do m = 1, members
      rc = gd_get_state(handle, int(m - 1, c_int32_t), state_u(1, 1, m), state_v(1, 1, m), &
                        state_gd(1, 1, m), ld)
      if (rc /= GD_OK) call log_error(gd_error_text(handle))
      call fill_halo(state_u(:, :, m))
      call fill_halo(state_v(:, :, m))
      call fill_halo(state_gd(:, :, m))
    end do

  end subroutine dycore_run

  ! dycore_final (dycore_mod.F90:144-157)
  subroutine dycore_final()

    integer(c_int) rc

    if (c_associated(handle)) rc = gd_destroy(handle)
    handle = c_null_ptr
    if (allocated(state_u)) deallocate(state_u, state_v, state_gd, static_ghs)
    call log_notice('Dycore module is finalized.')

  end subroutine dycore_final

end module dycore_mod
