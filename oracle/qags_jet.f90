! qags_jet.f90 -- test infrastructure only (never part of the product): the
! balance integral of the barotropic-jet test case,
!   int_{-pi/2}^{lat} a u (f + tan(lat) u / a) dlat
! (jet_zonal_flow_test_mod.F90:62-105), evaluated the way the reference does,
! by QUADPACK qags with epsabs 1e-10 and epsrel 1e-3, so that the library's
! own quadrature can be compared with the reference's result.  Linked with
! the reference's lib/quadpack.f90 into oracle/_ref/libqags_ref.so.
module qags_jet_mod

  use iso_c_binding

  implicit none

  real(8), parameter :: pi = atan(1.0d0) * 4.0d0
  real(8), parameter :: omega = 2.0d0 * pi / 86400.0d0
  real(8), parameter :: radius = 6.37122d6
  real(8), parameter :: u_max = 80.0d0
  real(8), parameter :: lat0 = pi / 7.0d0
  real(8), parameter :: lat1 = pi / 2.0d0 - lat0

contains

  real(8) function u_function(lat) result(res)

    real(8), intent(in) :: lat

    real(8) en

    en = exp(-4.0d0 / (lat1 - lat0)**2)
    if (lat <= lat0 .or. lat >= lat1) then
      res = 0.0d0
    else
      res = u_max / en * exp(1 / (lat - lat0) / (lat - lat1))
    end if

  end function u_function

  real(8) function gh_integrand(lat) result(res)

    real(8), intent(in) :: lat

    real(8) u, f

    u = u_function(lat)
    f = 2 * omega * sin(lat)
    res = radius * u * (f + tan(lat) / radius * u)

  end function gh_integrand

  subroutine ref_qags_jet(lat, result, abserr, ier) bind(C, name='ref_qags_jet')

    real(c_double), value :: lat
    real(c_double), intent(out) :: result, abserr
    integer(c_int), intent(out) :: ier

    integer neval, ier8
    real(8) r, e

    call qags(gh_integrand, -0.5d0 * pi, lat, 1.0d-10, 1.0d-3, r, e, neval, ier8)
    result = r
    abserr = e
    ier = ier8

  end subroutine ref_qags_jet

end module qags_jet_mod
