!! bind(c) shim over the reference's own Fortran leaf modules (oracle/_ref).
!! TEST INFRASTRUCTURE ONLY.  Library B: Rusanov, central two-point flux and
!! SSPRK(5,4) constants (module names clash with library A, hence a second
!! shared object).
module ref_shim_b
use iso_c_binding
use constants_mod
implicit none
contains

subroutine ref_rusanov(dir, ul, ur, f) bind(c, name='ref_rusanov')
  use riemann_mod, only: xriemann, yriemann, zriemann
  integer(c_int), value :: dir
  real(c_double), intent(in) :: ul(5), ur(5)
  real(c_double), intent(out) :: f(5)
  select case (dir)
  case (0); f = xriemann(ul, ur)
  case (1); f = yriemann(ul, ur)
  case default; f = zriemann(ul, ur)
  end select
end subroutine

subroutine ref_central(dir, u, v, f) bind(c, name='ref_central')
  use two_point_flux_mod, only: xflux, yflux, zflux
  integer(c_int), value :: dir
  real(c_double), intent(in) :: u(5), v(5)
  real(c_double), intent(out) :: f(5)
  select case (dir)
  case (0); f = xflux(u, v)
  case (1); f = yflux(u, v)
  case default; f = zflux(u, v)
  end select
end subroutine

subroutine ref_ssp_coeffs(k) bind(c, name='ref_ssp_coeffs')
  use timedisc_const_mod
  real(c_double), intent(out) :: k(23)
  k = (/ A0, A1, A2, A3, B0, B1, B2, B3, C0, C1, C2, C3, D0, D1, D2, D3, &
         E0, E1, E2, E3, E4, E5, E6 /)
end subroutine

end module
