! gd_capi_mod -- iso_c_binding interface to the C ABI of include/gamil_dycore.h.
! Every argument is a plain C type; 2-D fields are passed as the address of
! element (1,1) of the reference's halo'd arrays x(-1:N+2,-1:M+2) with the
! leading dimension ld = N+4.
module gd_capi_mod

  use iso_c_binding

  implicit none

  integer(c_int), parameter :: GD_OK = 0
  integer(c_int), parameter :: GD_ERR_INVALID_ARG = 1
  integer(c_int), parameter :: GD_ERR_HIP = 2
  integer(c_int), parameter :: GD_ERR_NAN = 3
  integer(c_int), parameter :: GD_ERR_NOT_SUPPORTED = 4

  type, bind(C) :: gd_config
    integer(c_int32_t) :: num_lon
    integer(c_int32_t) :: num_lat
    real(c_double) :: time_step_size
    integer(c_int32_t) :: split_scheme
    integer(c_int32_t) :: subcycles
    integer(c_int32_t) :: qcon_modified
    integer(c_int32_t) :: uv_adv_scheme
    real(c_double) :: uv_adv_upwind_lon_beta
    real(c_double) :: uv_adv_upwind_lat_beta
    integer(c_int32_t) :: use_diffusion
    integer(c_int32_t) :: diffusion_order
    real(c_double) :: diffusion_coef
    integer(c_int32_t) :: use_zonal_tend_filter
    integer(c_int32_t) :: n_cutoff
    type(c_ptr) :: cutoff
    integer(c_int32_t) :: precision
    integer(c_int32_t) :: device
  end type gd_config

  type, bind(C) :: gd_ic_params
    integer(c_int64_t) :: seed
    integer(c_int32_t) :: nmax
    real(c_double) :: slope
    real(c_double) :: rms_height
    real(c_double) :: mean_depth
    integer(c_int32_t) :: n_mountains
    integer(c_int64_t) :: mountain_seed
    integer(c_int64_t) :: pert_seed
    integer(c_int32_t) :: pert_nmax
    real(c_double) :: pert_rms
  end type gd_ic_params

  type, bind(C) :: gd_io_info
    type(c_ptr) :: desc
    type(c_ptr) :: author
    type(c_ptr) :: time_units
    type(c_ptr) :: start_time
    real(c_double) :: time_value
    real(c_double) :: time_step_size
    type(c_ptr) :: time_scheme
    type(c_ptr) :: split_scheme
    integer(c_int32_t) :: subcycles
    type(c_ptr) :: restart_time
    real(c_double) :: elapsed_seconds
  end type gd_io_info

  interface

    integer(c_int) function gd_create(cfg, n_members, h) bind(C, name='gd_create')
      import :: c_int, c_int32_t, c_ptr, gd_config
      type(gd_config), intent(in) :: cfg
      integer(c_int32_t), value :: n_members
      type(c_ptr), intent(out) :: h
    end function gd_create

    integer(c_int) function gd_destroy(h) bind(C, name='gd_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function gd_destroy

    type(c_ptr) function gd_last_error(h) bind(C, name='gd_last_error')
      import :: c_ptr
      type(c_ptr), value :: h
    end function gd_last_error

    integer(c_int) function gd_set_state(h, member, u, v, gd, ghs, ld) bind(C, name='gd_set_state')
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: member
      real(c_double), intent(in) :: u(*), v(*), gd(*), ghs(*)
      integer(c_int64_t), value :: ld
    end function gd_set_state

    integer(c_int) function gd_diag(h, total_mass, total_energy) bind(C, name='gd_diag')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: total_mass(*), total_energy(*)
    end function gd_diag

    integer(c_int) function gd_step(h, nsteps, total_mass, total_energy, beta, nan_step) &
        bind(C, name='gd_step')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nsteps
      real(c_double), intent(out) :: total_mass(*), total_energy(*), beta(*)
      integer(c_int32_t), intent(out) :: nan_step
    end function gd_step

    integer(c_int) function gd_get_state(h, member, u, v, gd, ld) bind(C, name='gd_get_state')
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: member
      real(c_double), intent(out) :: u(*), v(*), gd(*)
      integer(c_int64_t), value :: ld
    end function gd_get_state

    integer(c_int) function gd_get_diag_fields(h, member, div, vor, ld) &
        bind(C, name='gd_get_diag_fields')
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: member
      real(c_double), intent(out) :: div(*), vor(*)
      integer(c_int64_t), value :: ld
    end function gd_get_diag_fields

    integer(c_int) function gd_ic_balanced(num_lon, num_lat, p, u, v, gd, ghs, ld) &
        bind(C, name='gd_ic_balanced')
      import :: c_int, c_int32_t, c_int64_t, c_double, gd_ic_params
      integer(c_int32_t), value :: num_lon, num_lat
      type(gd_ic_params), intent(in) :: p
      real(c_double), intent(out) :: u(*), v(*), gd(*), ghs(*)
      integer(c_int64_t), value :: ld
    end function gd_ic_balanced

    integer(c_int) function gd_ic_rossby_haurwitz(num_lon, num_lat, R, omg, gd0, u, v, gd, ghs, ld) &
        bind(C, name='gd_ic_rossby_haurwitz')
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int32_t), value :: num_lon, num_lat
      real(c_double), value :: R, omg, gd0
      real(c_double), intent(out) :: u(*), v(*), gd(*), ghs(*)
      integer(c_int64_t), value :: ld
    end function gd_ic_rossby_haurwitz

    integer(c_int) function gd_ic_mountain_zonal_flow(num_lon, num_lat, u, v, gd, ghs, ld) &
        bind(C, name='gd_ic_mountain_zonal_flow')
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int32_t), value :: num_lon, num_lat
      real(c_double), intent(out) :: u(*), v(*), gd(*), ghs(*)
      integer(c_int64_t), value :: ld
    end function gd_ic_mountain_zonal_flow

    integer(c_int) function gd_ic_steady_geostrophic(num_lon, num_lat, u, v, gd, ghs, ld) &
        bind(C, name='gd_ic_steady_geostrophic')
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int32_t), value :: num_lon, num_lat
      real(c_double), intent(out) :: u(*), v(*), gd(*), ghs(*)
      integer(c_int64_t), value :: ld
    end function gd_ic_steady_geostrophic

    integer(c_int) function gd_ic_jet_zonal_flow(num_lon, num_lat, perturb, u, v, gd, ghs, ld) &
        bind(C, name='gd_ic_jet_zonal_flow')
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int32_t), value :: num_lon, num_lat, perturb
      real(c_double), intent(out) :: u(*), v(*), gd(*), ghs(*)
      integer(c_int64_t), value :: ld
    end function gd_ic_jet_zonal_flow

    integer(c_int) function gd_ic_shallow_water_waves(num_lon, num_lat, flag, u, v, gd, ghs, ld, &
                                                      phase_speed) bind(C, name='gd_ic_shallow_water_waves')
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int32_t), value :: num_lon, num_lat, flag
      real(c_double), intent(out) :: u(*), v(*), gd(*), ghs(*)
      integer(c_int64_t), value :: ld
      real(c_double), intent(out) :: phase_speed
    end function gd_ic_shallow_water_waves

    integer(c_int) function gd_cutoff_from_lat0(num_lon, num_lat, lat0_deg, cutoff, cap) &
        bind(C, name='gd_cutoff_from_lat0')
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: num_lon, num_lat
      real(c_double), value :: lat0_deg
      integer(c_int32_t), intent(out) :: cutoff(*)
      integer(c_int32_t), value :: cap
    end function gd_cutoff_from_lat0

    integer(c_int) function gd_history_write(path, info, num_lon, num_lat, u, v, gd, ghs, div, vor, &
                                             ld, total_energy, total_mass) bind(C, name='gd_history_write')
      import :: c_int, c_int32_t, c_int64_t, c_double, c_char, gd_io_info
      character(kind=c_char), intent(in) :: path(*)
      type(gd_io_info), intent(in) :: info
      integer(c_int32_t), value :: num_lon, num_lat
      real(c_double), intent(in) :: u(*), v(*), gd(*), ghs(*), div(*), vor(*)
      integer(c_int64_t), value :: ld
      real(c_double), value :: total_energy, total_mass
    end function gd_history_write

    integer(c_int) function gd_restart_write(path, info, num_lon, num_lat, u, v, gd, ghs, ld) &
        bind(C, name='gd_restart_write')
      import :: c_int, c_int32_t, c_int64_t, c_double, c_char, gd_io_info
      character(kind=c_char), intent(in) :: path(*)
      type(gd_io_info), intent(in) :: info
      integer(c_int32_t), value :: num_lon, num_lat
      real(c_double), intent(in) :: u(*), v(*), gd(*), ghs(*)
      integer(c_int64_t), value :: ld
    end function gd_restart_write

    integer(c_int) function gd_restart_read(path, num_lon, num_lat, u, v, gd, ghs, ld, restart_time, &
                                            cap, elapsed_seconds) bind(C, name='gd_restart_read')
      import :: c_int, c_int32_t, c_int64_t, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: num_lon, num_lat
      real(c_double), intent(out) :: u(*), v(*), gd(*), ghs(*)
      integer(c_int64_t), value :: ld
      character(kind=c_char), intent(out) :: restart_time(*)
      integer(c_int32_t), value :: cap
      real(c_double), intent(out) :: elapsed_seconds
    end function gd_restart_read

    type(c_ptr) function gd_io_last_error() bind(C, name='gd_io_last_error')
      import :: c_ptr
    end function gd_io_last_error

    integer(c_int) function gd_autotune(h, steps, cache_path, fused, upd_wpm, ms_per_step) &
                                        bind(C, name='gd_autotune')
      import :: c_int, c_int32_t, c_double, c_ptr, c_char
      type(c_ptr), value :: h
      integer(c_int32_t), value :: steps
      character(kind=c_char), intent(in) :: cache_path(*)
      integer(c_int32_t), intent(out) :: fused, upd_wpm
      real(c_double), intent(out) :: ms_per_step
    end function gd_autotune

  end interface

contains

  ! NUL-terminated copy of a Fortran string (for path arguments).
  function c_string(s) result(res)

    character(*), intent(in) :: s
    character(kind=c_char) :: res(len_trim(s) + 1)

    integer k

    do k = 1, len_trim(s)
      res(k) = s(k:k)
    end do
    res(len_trim(s) + 1) = c_null_char

  end function c_string

  ! Fortran string of a NUL-terminated C string.
  function f_string(p, cap) result(res)

    type(c_ptr), intent(in) :: p
    integer, intent(in) :: cap
    character(512) :: res

    character(kind=c_char), pointer :: s(:)
    integer k

    res = ''
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [cap])
    do k = 1, min(cap, 512)
      if (s(k) == c_null_char) exit
      res(k:k) = s(k)
    end do

  end function f_string

  ! Text of gd_last_error as a Fortran string.
  function gd_error_text(h) result(res)

    type(c_ptr), intent(in) :: h
    character(512) :: res

    res = f_string(gd_last_error(h), 512)

  end function gd_error_text

end module gd_capi_mod
