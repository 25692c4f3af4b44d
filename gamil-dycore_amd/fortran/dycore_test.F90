! dycore_test -- the driver program of the reference (src/dycore_test.F90:1-47):
! read the namelist named on the command line, dycore_init, write the
! initial condition of test_case into the state, dycore_run, dycore_final.
!
!   dycore_test.exe <namelist>             run on the GPU
!   dycore_test.exe <namelist> --check     parse and print the configuration only
!   dycore_test.exe <namelist> --schedule  print the output files the run would
!                                          write (time manager only, no GPU)
!
! Test cases: the reference's steady_geostrophic_flow, rossby_haurwitz_wave
! (&rossby_haurwitz_wave_test_params R, omg, gd0), mountain_zonal_flow,
! jet_zonal_flow and shallow_water_waves, and balanced_sh (&balanced_sh_params): a seeded
! geostrophically balanced spherical-harmonic state, optionally with Gaussian
! mountains; ensemble member m adds the perturbation seed pert_seed + m - 1.
program dycore_test

  use iso_c_binding
  use gd_params_mod
  use gd_capi_mod
  use dycore_mod

  implicit none

  character(256) namelist_file_path, arg
  logical check_only, schedule_only
  integer(c_int) rc
  integer m, ios
  integer(c_int64_t) ld

  real(r8) :: R = 4.0_r8, omg = 7.848e-6_r8, gd0 = 8.0e3_r8 * 9.80616_r8
  logical :: smooth_mountain = .false.
  integer(8) :: seed = 1, mountain_seed = 0, pert_seed = 0
  integer :: nmax = 20, n_mountains = 0, pert_nmax = 0
  real(r8) :: slope = -3.0_r8, rms_height = 150.0_r8, mean_depth = 8000.0_r8, pert_rms = 0.0_r8
  real(r8) :: phase_speed
  type(gd_ic_params) icp

  namelist /rossby_haurwitz_wave_test_params/ R, omg, gd0
  namelist /mountain_zonal_flow_test_params/ smooth_mountain
  namelist /balanced_sh_params/ seed, nmax, slope, rms_height, mean_depth, n_mountains, &
    mountain_seed, pert_seed, pert_nmax, pert_rms

  if (command_argument_count() < 1) then
    write(6, *) 'Usage: ./dycore_test.exe <namelist_file_path> [--check]'
    stop 1
  end if
  call get_command_argument(1, namelist_file_path)
  check_only = .false.
  schedule_only = .false.
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    check_only = trim(arg) == '--check'
    schedule_only = trim(arg) == '--schedule'
  end if

  call params_read(namelist_file_path)

  if (check_only) then
    call dycore_describe()
  else if (schedule_only) then
    call dycore_schedule()
  else
    call run_case()
  end if

contains

  subroutine run_case()

  call dycore_init()

  ld = num_lon + 4
  if (is_restart_run) then
    call dycore_restart()
  else
    select case (test_case)
    case ('steady_geostrophic_flow')
      write(6, *) '[Notice]: Use steady geostrophic flow initial condition.'
      do m = 1, members
        rc = gd_ic_steady_geostrophic(int(num_lon, c_int32_t), int(num_lat, c_int32_t), &
                                      state_u(1, 1, m), state_v(1, 1, m), state_gd(1, 1, m), &
                                      static_ghs(1, 1, m), ld)
        call check(rc)
      end do
    case ('jet_zonal_flow')
      write(6, *) '[Notice]: Use jet zonal flow initial condition.'
      do m = 1, members
        rc = gd_ic_jet_zonal_flow(int(num_lon, c_int32_t), int(num_lat, c_int32_t), 1_c_int32_t, &
                                  state_u(1, 1, m), state_v(1, 1, m), state_gd(1, 1, m), &
                                  static_ghs(1, 1, m), ld)
        call check(rc)
      end do
    case ('shallow_water_waves')
      write(6, *) '[Notice]: Use shallow water waves initial condition.'
      do m = 1, members
        rc = gd_ic_shallow_water_waves(int(num_lon, c_int32_t), int(num_lat, c_int32_t), 0_c_int32_t, &
                                       state_u(1, 1, m), state_v(1, 1, m), state_gd(1, 1, m), &
                                       static_ghs(1, 1, m), ld, phase_speed)
        call check(rc)
      end do
      write(6, '("[Notice]: Phase speed is ", ES27.20)') phase_speed
    case ('rossby_haurwitz_wave')
      write(6, *) '[Notice]: Use Rossby-Haurwitz wave initial condition.'
      call read_group(1)
      do m = 1, members
        rc = gd_ic_rossby_haurwitz(int(num_lon, c_int32_t), int(num_lat, c_int32_t), R, omg, gd0, &
                                   state_u(1, 1, m), state_v(1, 1, m), state_gd(1, 1, m), &
                                   static_ghs(1, 1, m), ld)
        call check(rc)
      end do
    case ('mountain_zonal_flow')
      write(6, *) '[Notice]: Use mountain zonal flow initial condition.'
      call read_group(2)
      if (smooth_mountain) then
        write(6, '(A)') '[Error]: smooth_mountain = .true. is not provided!'
        stop 1
      end if
      do m = 1, members
        rc = gd_ic_mountain_zonal_flow(int(num_lon, c_int32_t), int(num_lat, c_int32_t), &
                                       state_u(1, 1, m), state_v(1, 1, m), state_gd(1, 1, m), &
                                       static_ghs(1, 1, m), ld)
        call check(rc)
      end do
    case ('balanced_sh')
      write(6, *) '[Notice]: Use seeded balanced spherical-harmonic initial condition.'
      call read_group(3)
      do m = 1, members
        icp%seed = seed
        icp%nmax = nmax
        icp%slope = slope
        icp%rms_height = rms_height
        icp%mean_depth = mean_depth
        icp%n_mountains = n_mountains
        icp%mountain_seed = mountain_seed
        icp%pert_seed = pert_seed + m - 1
        icp%pert_nmax = pert_nmax
        icp%pert_rms = pert_rms
        rc = gd_ic_balanced(int(num_lon, c_int32_t), int(num_lat, c_int32_t), icp, &
                            state_u(1, 1, m), state_v(1, 1, m), state_gd(1, 1, m), &
                            static_ghs(1, 1, m), ld)
        call check(rc)
      end do
    case default
      write(6, *) '[Error]: Unknown test case ' // trim(test_case) // '!'
      stop 1
    end select
  end if

  call dycore_run()

  write(6, '("final total_mass   = ", ES27.20)') final_mass
  write(6, '("final total_energy = ", ES27.20)') final_energy

  call dycore_final()

  end subroutine run_case

  subroutine check(code)

    integer(c_int), intent(in) :: code

    if (code /= GD_OK) then
      write(6, '(A)') '[Error]: ' // trim(gd_error_text(c_null_ptr))
      stop 1
    end if

  end subroutine check

  ! A missing test-case group keeps the defaults (the reference requires it).
  subroutine read_group(which)

    integer, intent(in) :: which

    open(11, file=namelist_file_path, status='old', action='read')
    select case (which)
    case (1)
      read(11, nml=rossby_haurwitz_wave_test_params, iostat=ios)
    case (2)
      read(11, nml=mountain_zonal_flow_test_params, iostat=ios)
    case (3)
      read(11, nml=balanced_sh_params, iostat=ios)
    end select
    close(11)

  end subroutine read_group

end program dycore_test
