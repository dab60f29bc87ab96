!! bind(c) shim over the reference's own Fortran leaf modules (oracle/_ref).
!! TEST INFRASTRUCTURE ONLY: the reference modules are compiled from their
!! original files under the reference tree; this shim only forwards calls.
!! Library A: Euler physics, hybrid HLLC/HLLE, Chandrashekar two-point flux,
!! LSRK 5-4 constants, LGL/Legendre/DiffMatrix utilities.
module ref_shim_a
use iso_c_binding
use constants_mod
implicit none
contains

function ref_pressure(u) result(p) bind(c, name='ref_pressure')
  use equations_mod, only: pressure
  real(c_double), intent(in) :: u(5)
  real(c_double) :: p
  p = pressure(u)
end function

subroutine ref_flux(dir, u, f) bind(c, name='ref_flux')
  use equations_mod, only: xflux, yflux, zflux
  integer(c_int), value :: dir
  real(c_double), intent(in) :: u(5)
  real(c_double), intent(out) :: f(5)
  select case (dir)
  case (0); f = xflux(u)
  case (1); f = yflux(u)
  case default; f = zflux(u)
  end select
end subroutine

function ref_lambda_max(dir, u) result(l) bind(c, name='ref_lambda_max')
  use equations_mod, only: xlambdaMax, ylambdaMax, zlambdaMax
  integer(c_int), value :: dir
  real(c_double), intent(in) :: u(5)
  real(c_double) :: l
  select case (dir)
  case (0); l = xlambdaMax(u)
  case (1); l = ylambdaMax(u)
  case default; l = zlambdaMax(u)
  end select
end function

subroutine ref_cons2evec(u, w) bind(c, name='ref_cons2evec')
  use equations_mod, only: cons2evec
  real(c_double), intent(in) :: u(5)
  real(c_double), intent(out) :: w(5)
  w = cons2evec(u)
end subroutine

function ref_isvalid(u) result(ok) bind(c, name='ref_isvalid')
  use equations_mod, only: isvalid
  real(c_double), intent(in) :: u(5)
  integer(c_int) :: ok
  ok = merge(1, 0, isvalid(u))
end function

subroutine ref_hllc(dir, ul, ur, f) bind(c, name='ref_hllc')
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

subroutine ref_ec(dir, u, v, f) bind(c, name='ref_ec')
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

subroutine ref_lsrk_coeffs(a, b, c) bind(c, name='ref_lsrk_coeffs')
  use timedisc_const_mod, only: a_rk, b_rk, c_rk
  real(c_double), intent(out) :: a(5), b(5), c(5)
  a = a_rk; b = b_rk; c = c_rk
end subroutine

subroutine ref_lgl(n, x, w) bind(c, name='ref_lgl')
  use interpol_mod, only: LegendreGaussLobattoNodesAndWeights
  integer(c_int), value :: n
  real(c_double), intent(out) :: x(0:n), w(0:n)
  call LegendreGaussLobattoNodesAndWeights(n, x, w)
end subroutine

subroutine ref_diff_matrix(n, x, d) bind(c, name='ref_diff_matrix')
  use interpol_mod, only: DiffMatrix
  integer(c_int), value :: n
  real(c_double), intent(in) :: x(0:n)
  real(c_double), intent(out) :: d(0:n, 0:n)
  call DiffMatrix(n, x, d)
end subroutine

subroutine ref_legendre(n, x, l, dl) bind(c, name='ref_legendre')
  use interpol_mod, only: LegendrePolynomialAndDerivative
  integer(c_int), value :: n
  real(c_double), value :: x
  real(c_double), intent(out) :: l, dl
  call LegendrePolynomialAndDerivative(n, x, l, dl)
end subroutine

subroutine ref_vandermonde(m, nodes, vdm, svdm) bind(c, name='ref_vandermonde')
  use interpol_mod, only: mkVanderMondeMatrix
  integer(c_int), value :: m
  real(c_double), intent(in) :: nodes(m)
  real(c_double), intent(out) :: vdm(m, m), svdm(m, m)
  call mkVanderMondeMatrix(m, nodes, vdm, svdm)
end subroutine

function ref_minmod3(a, b, c) result(r) bind(c, name='ref_minmod3')
  use math_mod, only: minmod3
  real(c_double), value :: a, b, c
  real(c_double) :: r
  r = minmod3(a, b, c)
end function

end module
