!! bind(c) shim over the reference's own Fortran leaf modules (oracle/_ref).
!! TEST INFRASTRUCTURE ONLY.  Library C: ideal GLM-MHD (9 variables) with the
!! Orszag-Tang setup's setup_config_mod (kappa = 5/3, mu0 = 1): physical
!! fluxes, pressure, fast magnetosonic speed, lambda_max, isvalid, Rusanov and
!! the Powell/GLM non-conservative two-point terms.
module ref_shim_c
use iso_c_binding
use constants_mod
implicit none
contains

subroutine ref_mhd_set_glm(ch) bind(c, name='ref_mhd_set_glm')
  use equations_share_mod, only: glm_ch
  real(c_double), value :: ch
  glm_ch = ch
end subroutine

function ref_mhd_pressure(u) result(p) bind(c, name='ref_mhd_pressure')
  use equations_mod, only: pressure
  real(c_double), intent(in) :: u(9)
  real(c_double) :: p
  p = pressure(u)
end function

subroutine ref_mhd_flux(dir, u, f) bind(c, name='ref_mhd_flux')
  use equations_mod, only: xflux, yflux, zflux
  integer(c_int), value :: dir
  real(c_double), intent(in) :: u(9)
  real(c_double), intent(out) :: f(9)
  select case (dir)
  case (0); f = xflux(u)
  case (1); f = yflux(u)
  case default; f = zflux(u)
  end select
end subroutine

function ref_mhd_lambda_max(dir, u) result(l) bind(c, name='ref_mhd_lambda_max')
  use equations_mod, only: xlambdaMax, ylambdaMax, zlambdaMax
  integer(c_int), value :: dir
  real(c_double), intent(in) :: u(9)
  real(c_double) :: l
  select case (dir)
  case (0); l = xlambdaMax(u)
  case (1); l = ylambdaMax(u)
  case default; l = zlambdaMax(u)
  end select
end function

function ref_mhd_fastspeed(u) result(c) bind(c, name='ref_mhd_fastspeed')
  use equations_mod, only: Fastest_Signal_Speed
  real(c_double), intent(in) :: u(9)
  real(c_double) :: c
  call Fastest_Signal_Speed(u, c)
end function

function ref_mhd_isvalid(u) result(ok) bind(c, name='ref_mhd_isvalid')
  use equations_mod, only: isvalid
  real(c_double), intent(in) :: u(9)
  integer(c_int) :: ok
  ok = merge(1, 0, isvalid(u))
end function

subroutine ref_mhd_rusanov(dir, ul, ur, f) bind(c, name='ref_mhd_rusanov')
  use riemann_mod, only: xriemann, yriemann, zriemann
  integer(c_int), value :: dir
  real(c_double), intent(in) :: ul(9), ur(9)
  real(c_double), intent(out) :: f(9)
  select case (dir)
  case (0); f = xriemann(ul, ur)
  case (1); f = yriemann(ul, ur)
  case default; f = zriemann(ul, ur)
  end select
end subroutine

subroutine ref_mhd_nc(dir, ui, ul, ur, f) bind(c, name='ref_mhd_nc')
  use non_conservative_flux_mod, only: Non_Conservative_two_point_xFlux, &
      Non_Conservative_two_point_yFlux, Non_Conservative_two_point_zFlux
  integer(c_int), value :: dir
  real(c_double), intent(in) :: ui(9), ul(9), ur(9)
  real(c_double), intent(out) :: f(9)
  select case (dir)
  case (0); f = Non_Conservative_two_point_xFlux(ui, ul, ur)
  case (1); f = Non_Conservative_two_point_yFlux(ui, ul, ur)
  case default; f = Non_Conservative_two_point_zFlux(ui, ul, ur)
  end select
end subroutine

end module
