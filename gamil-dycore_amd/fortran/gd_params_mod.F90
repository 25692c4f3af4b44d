! gd_params_mod -- the &dycore_params namelist of params_mod.F90:13-98
! (reference) as a superset, read the way params_read does (params_mod.F90:102-115).
!
! Superset keys:
!   zonal_tend_filter_lat0  > 0: build the polar filter table from the rule
!                           c_k = floor(num_lon/2 cos(lat_k)/cos(lat0))
!   precision               'fp64' | 'mixed' | 'mixed_mfma' | 'mixed_bf16' (the last two
!                           filter the momentum rows on the matrix cores, FP32 or BF16 inputs)
!   ensemble_members        number of device-resident runs of this config
!   device                  HIP device ordinal
!   log_every               print the per-step diagnostics line every n steps
!   autotune_steps          > 0: time that many steps of each launch shape that cannot
!                           change a result and keep the fastest (gd_autotune) before the run
!   autotune_cache          file of earlier autotuner results ('' : none)
!   write_output            .false.: no history or restart files (timing runs)
!   output_dir              directory of the history and restart files
!                           (default: the working directory, as the reference)
!   days                    accepted as run_days (run/namelist.rh_test:6)
!   use_zonal_reduce, reduce_adv_lon, zonal_reduce_factors,
!   use_reduce_tend_smooth  keys of run/namelist.jz_test:16-19 that the
!                           reference never declared; accepted and ignored
! The cutoff table holds up to max_cutoff entries (20 in the reference).
module gd_params_mod

  use iso_c_binding, only: c_double

  implicit none

  integer, parameter :: r8 = c_double

  real(r8), parameter :: pi = 3.14159265358979323846_r8
  real(r8), parameter :: omega = 2.0_r8 * pi / 86400.0_r8
  real(r8), parameter :: radius = 6.37122e6_r8
  real(r8), parameter :: g = 9.80616_r8

  integer, parameter :: max_cutoff = 2048

  integer :: num_lon = 0
  integer :: num_lat = 0
  integer :: subcycles = 4
  real(r8) :: time_step_size = 0.0_r8

  logical :: use_diffusion = .false.
  real(r8) :: diffusion_coef = 0.0_r8
  integer :: diffusion_order = 2

  integer :: run_days = 0
  integer :: run_hours = 0
  integer :: run_minutes = 0
  integer :: run_seconds = 0
  integer :: start_time(5) = [0, 0, 0, 0, 0]
  integer :: end_time(5) = [0, 0, 0, 0, 0]
  character(30) :: time_units = 'days'

  character(30) :: test_case = ''
  character(256) :: case_name = ''
  character(256) :: case_desc = ''
  character(256) :: author = ''
  character(30) :: history_periods(1) = ['6 hours']
  character(30) :: restart_period = ''

  character(256) :: namelist_file = ''
  character(256) :: restart_file = ''

  character(30) :: time_scheme = 'predict_correct'
  integer :: time_order = 2
  logical :: qcon_modified = .false.
  character(30) :: split_scheme = ''
  character(30) :: uv_adv_scheme = 'center_diff'
  real(r8) :: uv_adv_upwind_lon_beta = 0.0_r8
  real(r8) :: uv_adv_upwind_lat_beta = 0.5_r8
  integer :: weno_order = 2

  logical :: is_restart_run = .false.

  logical :: use_zonal_tend_filter = .true.
  integer :: zonal_tend_filter_cutoff_wavenumber(max_cutoff) = 0

  ! superset
  real(r8) :: zonal_tend_filter_lat0 = 0.0_r8
  character(30) :: precision = 'fp64'
  integer :: ensemble_members = 1
  integer :: device = 0
  integer :: log_every = 1
  integer :: autotune_steps = 0
  character(256) :: autotune_cache = ''
  logical :: write_output = .true.
  character(256) :: output_dir = ''
  integer :: days = 0
  logical :: use_zonal_reduce = .false.
  logical :: reduce_adv_lon = .false.
  logical :: use_reduce_tend_smooth = .false.
  integer :: zonal_reduce_factors(20) = 0

  namelist /dycore_params/ &
    num_lon, num_lat, subcycles, run_days, run_hours, run_minutes, run_seconds, &
    start_time, end_time, time_units, time_step_size, test_case, case_name, case_desc, &
    author, history_periods, restart_period, restart_file, time_scheme, time_order, &
    qcon_modified, split_scheme, uv_adv_scheme, uv_adv_upwind_lon_beta, &
    uv_adv_upwind_lat_beta, use_zonal_tend_filter, zonal_tend_filter_cutoff_wavenumber, &
    use_diffusion, diffusion_order, diffusion_coef, &
    zonal_tend_filter_lat0, precision, ensemble_members, device, log_every, days, &
    autotune_steps, autotune_cache, write_output, output_dir, &
    use_zonal_reduce, reduce_adv_lon, zonal_reduce_factors, use_reduce_tend_smooth

contains

  subroutine params_read(file_path)

    character(*), intent(in) :: file_path

    integer ios
    character(256) msg

    namelist_file = file_path

    open(10, file=file_path, status='old', action='read', iostat=ios, iomsg=msg)
    if (ios /= 0) then
      write(6, '(A)') '[Error]: Cannot open namelist file ' // trim(file_path) // ': ' // trim(msg)
      stop 1
    end if
    read(10, nml=dycore_params, iostat=ios, iomsg=msg)
    close(10)
    if (ios /= 0) then
      write(6, '(A)') '[Error]: Failed to read &dycore_params: ' // trim(msg)
      stop 1
    end if

    if (days > 0 .and. run_days == 0) run_days = days
    is_restart_run = restart_file /= ''
    if (restart_period == '') restart_period = history_periods(1)

  end subroutine params_read

  ! Number of steps of a run set by run_days/hours/minutes: the loop of
  ! dycore_mod.F90:131 runs while curr_time < end_time (time_mod.F90:138-142),
  ! i.e. ceiling(T / dt); run_seconds is not part of end_time
  ! (time_mod.F90:62-68).
  integer function total_steps() result(n)

    real(r8) t

    t = run_days * 86400.0_r8 + run_hours * 3600.0_r8 + run_minutes * 60.0_r8
    if (time_step_size <= 0.0_r8) then
      n = 0
    else
      n = ceiling(t / time_step_size - 1.0e-9_r8)
    end if

  end function total_steps

end module gd_params_mod
