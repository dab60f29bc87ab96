!! ========================================================================== !!
!! Host-side mesh and initial data for the Fortran driver.
!!
!! mesh_uniform: uniform periodic box of n^3 hexahedra on [0,1]^3 in Morton
!!   order (the leaf order of a uniformly refined single-tree p4est forest,
!!   mesh/p4est/mesh_utils_mod.f90:89-145) flattened into the upload format
!!   of nemo_gpu_upload_mesh: nbr(6,nelem) with faces NOR,SOU,WES,EAS,FRO,BAC
!!   (utils/constants_mod.f90:34-40) and delta(3,nelem) (quad%delta).
!! fill_*: initial data per node through prim2cons, as the reference's setup
!!   hooks do (e.g. setups/euler/3d/sedov/setup_init_mod.f90:7-58), into
!!   state(N_NODES,N_NODES,N_NODES,N_VARS,nelem).
!! ========================================================================== !!
module nemo_setup_mod

use iso_c_binding

implicit none

integer, parameter :: dp = c_double
real(dp), parameter :: PI = 4.0_dp*atan(1.0_dp)

type :: mesh_t
    integer :: n = 0                       !! elements per direction
    integer(c_int64_t) :: nelem = 0
    integer(c_int32_t), allocatable :: nbr(:,:)     !! (6, nelem), 0-based
    real(dp), allocatable :: delta(:,:)             !! (3, nelem)
    real(dp), allocatable :: lower(:,:)             !! (3, nelem)
end type

contains

pure function spread3(v) result(r)
    integer, intent(in) :: v
    integer(c_int64_t)  :: r
    integer :: b
    r = 0
    do b = 0, 20
        if (btest(v, b)) r = ibset(r, 3*b)
    end do
end function

pure function morton(i, j, k) result(r)
    integer, intent(in) :: i, j, k
    integer(c_int64_t)  :: r
    r = ior(ior(spread3(i), ishft(spread3(j), 1)), ishft(spread3(k), 2))
end function

subroutine mesh_uniform(n, mesh)
    integer, intent(in)         :: n
    type(mesh_t), intent(out)   :: mesh
    integer, allocatable :: pos(:), ii(:), jj(:), kk(:)
    integer :: i, j, k, e, lex
    logical :: pow2

    pow2 = iand(n, n-1) == 0
    mesh%n = n
    mesh%nelem = int(n, c_int64_t)**3
    allocate(mesh%nbr(6, mesh%nelem), mesh%delta(3, mesh%nelem), mesh%lower(3, mesh%nelem))
    allocate(pos(0:n**3-1), ii(0:n**3-1), jj(0:n**3-1), kk(0:n**3-1))

    !! element order: Morton key for power-of-two boxes, lexicographic otherwise
    do k = 0, n-1; do j = 0, n-1; do i = 0, n-1
        lex = i + n*(j + n*k)
        if (pow2) then
            e = int(morton(i, j, k))
        else
            e = lex
        end if
        pos(lex) = e
        ii(e) = i; jj(e) = j; kk(e) = k
    end do; end do; end do

    do e = 0, n**3-1
        i = ii(e); j = jj(e); k = kk(e)
        mesh%nbr(1, e+1) = pos(modulo(i-1, n) + n*(j + n*k))
        mesh%nbr(2, e+1) = pos(modulo(i+1, n) + n*(j + n*k))
        mesh%nbr(3, e+1) = pos(i + n*(modulo(j-1, n) + n*k))
        mesh%nbr(4, e+1) = pos(i + n*(modulo(j+1, n) + n*k))
        mesh%nbr(5, e+1) = pos(i + n*(j + n*modulo(k-1, n)))
        mesh%nbr(6, e+1) = pos(i + n*(j + n*modulo(k+1, n)))
        mesh%delta(:, e+1) = 1.0_dp/real(n, dp)
        mesh%lower(:, e+1) = (/ i, j, k /) / real(n, dp)
    end do
end subroutine

!! prim2cons: equations/euler/equations_mod.f90:195-208
pure subroutine prim2cons(kappa, rho, vx, vy, vz, p, u)
    real(dp), intent(in)  :: kappa, rho, vx, vy, vz, p
    real(dp), intent(out) :: u(5)
    u(1) = rho
    u(2) = vx*rho
    u(3) = vy*rho
    u(4) = vz*rho
    u(5) = p/(kappa-1.0_dp) + 0.5_dp*rho*(vx**2 + vy**2 + vz**2)
end subroutine

!! Isentropic vortex (strength 5, background rho = p = 1, u = (1,1,0)) with
!! length scale 0.1 around (0.5,0.5), uniform in z.
subroutine fill_vortex(mesh, nn, xi, kappa, state)
    type(mesh_t), intent(in) :: mesh
    integer, intent(in)      :: nn
    real(dp), intent(in)     :: xi(nn), kappa
    real(dp), intent(out)    :: state(nn, nn, nn, 5, mesh%nelem)
    real(dp), parameter :: eps = 5.0_dp, length = 0.1_dp
    real(dp) :: x, y, dx, dy, r2, du, t, rho
    integer(c_int64_t) :: e
    integer :: i, j, k
    do e = 1, mesh%nelem
        do k = 1, nn; do j = 1, nn; do i = 1, nn
            x = mesh%lower(1, e) + 0.5_dp*(xi(i) + 1.0_dp)*mesh%delta(1, e)
            y = mesh%lower(2, e) + 0.5_dp*(xi(j) + 1.0_dp)*mesh%delta(2, e)
            dx = (x - 0.5_dp)/length
            dy = (y - 0.5_dp)/length
            r2 = dx*dx + dy*dy
            du = eps/(2.0_dp*PI)*exp(0.5_dp*(1.0_dp - r2))
            t = 1.0_dp - (kappa - 1.0_dp)*eps*eps/(8.0_dp*kappa*PI**2)*exp(1.0_dp - r2)
            rho = t**(1.0_dp/(kappa - 1.0_dp))
            call prim2cons(kappa, rho, 1.0_dp - du*dy, 1.0_dp + du*dx, 0.0_dp, rho**kappa, state(i, j, k, :, e))
        end do; end do; end do
    end do
end subroutine

!! Kelvin-Helmholtz: rho = 2 in |y-0.5| < 0.25 and 1 elsewhere, u_x = +-0.5
!! joined by tanh profiles of width 0.025, p = 2.5, plus a deterministic
!! velocity perturbation of amplitude `noise` (a hash of the node position).
subroutine fill_kh(mesh, nn, xi, kappa, noise, state)
    type(mesh_t), intent(in) :: mesh
    integer, intent(in)      :: nn
    real(dp), intent(in)     :: xi(nn), kappa, noise
    real(dp), intent(out)    :: state(nn, nn, nn, 5, mesh%nelem)
    real(dp), parameter :: width = 0.025_dp
    real(dp) :: x, y, z, rho, band, h(3)
    integer(c_int64_t) :: e
    integer :: i, j, k, d
    do e = 1, mesh%nelem
        do k = 1, nn; do j = 1, nn; do i = 1, nn
            x = mesh%lower(1, e) + 0.5_dp*(xi(i) + 1.0_dp)*mesh%delta(1, e)
            y = mesh%lower(2, e) + 0.5_dp*(xi(j) + 1.0_dp)*mesh%delta(2, e)
            z = mesh%lower(3, e) + 0.5_dp*(xi(k) + 1.0_dp)*mesh%delta(3, e)
            rho = merge(2.0_dp, 1.0_dp, abs(y - 0.5_dp) < 0.25_dp)
            band = 0.5_dp*(tanh((y - 0.25_dp)/width) - tanh((y - 0.75_dp)/width))
            do d = 1, 3
                h(d) = sin(12.9898_dp*x + 78.233_dp*y + 37.719_dp*z + 17.0_dp*d)*43758.5453_dp
                h(d) = 2.0_dp*(h(d) - floor(h(d))) - 1.0_dp
            end do
            call prim2cons(kappa, rho, -0.5_dp + band + noise*h(1), noise*h(2), noise*h(3), 2.5_dp, &
                           state(i, j, k, :, e))
        end do; end do; end do
    end do
end subroutine

!! Boundary faces across x (the mesh stays periodic in y and z): the x- face
!! of every element in the first layer and the x+ face of every element in
!! the last one, in element then face order, each carrying
!! nbr = -(b + 1) for its boundary face b (0-based) and the given kind.
subroutine mesh_walls_x(mesh, kind, nbound, bfaces)
    type(mesh_t), intent(inout)      :: mesh
    integer(c_int32_t), intent(in)   :: kind
    integer(c_int64_t), intent(out)  :: nbound
    integer(c_int32_t), allocatable, intent(out) :: bfaces(:,:)
    integer(c_int64_t) :: e
    integer :: i
    allocate(bfaces(3, 2*mesh%n**2))
    nbound = 0
    do e = 1, mesh%nelem
        i = nint(mesh%lower(1, e)*mesh%n)
        if (i == 0) then
            nbound = nbound + 1
            bfaces(:, nbound) = (/ int(e-1, c_int32_t), 0_c_int32_t, kind /)
            mesh%nbr(1, e) = -int(nbound, c_int32_t)
        end if
        if (i == mesh%n - 1) then
            nbound = nbound + 1
            bfaces(:, nbound) = (/ int(e-1, c_int32_t), 1_c_int32_t, kind /)
            mesh%nbr(2, e) = -int(nbound, c_int32_t)
        end if
    end do
end subroutine

!! Sod's shock tube along x: (rho, p) = (1, 1) for x < 0.5 and (0.125, 0.1)
!! beyond, at rest.
subroutine fill_sod(mesh, nn, xi, kappa, state)
    type(mesh_t), intent(in) :: mesh
    integer, intent(in)      :: nn
    real(dp), intent(in)     :: xi(nn), kappa
    real(dp), intent(out)    :: state(nn, nn, nn, 5, mesh%nelem)
    real(dp) :: x
    integer(c_int64_t) :: e
    integer :: i, j, k
    do e = 1, mesh%nelem
        do k = 1, nn; do j = 1, nn; do i = 1, nn
            x = mesh%lower(1, e) + 0.5_dp*(xi(i) + 1.0_dp)*mesh%delta(1, e)
            if (x < 0.5_dp) then
                call prim2cons(kappa, 1.0_dp, 0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, state(i, j, k, :, e))
            else
                call prim2cons(kappa, 0.125_dp, 0.0_dp, 0.0_dp, 0.0_dp, 0.1_dp, state(i, j, k, :, e))
            end if
        end do; end do; end do
    end do
end subroutine

!! prim2cons of GLM-MHD: equations/mhd/equations_mod.f90:294-310 (mu0 = 1;
!! the energy holds B^2/2 and psi^2/2).
pure subroutine prim2cons_mhd(kappa, rho, v, p, b, psi, u)
    real(dp), intent(in)  :: kappa, rho, v(3), p, b(3), psi
    real(dp), intent(out) :: u(9)
    u(1) = rho
    u(2:4) = rho*v
    u(5) = p/(kappa-1.0_dp) + 0.5_dp*rho*sum(v**2) + 0.5_dp*(sum(b**2) + psi**2)
    u(6:8) = b
    u(9) = psi
end subroutine

!! Orszag-Tang-type vortex on the unit cube (GLM-MHD, mu0 = 1):
!! rho = 25/(36 pi), p = 5/(12 pi), v = (-sin 2 pi y, sin 2 pi x, 0.1 sin 2 pi z),
!! B = (-sin 2 pi y, sin 4 pi x, 0)/sqrt(4 pi), psi = 0, plus a deterministic
!! velocity perturbation of amplitude `noise` (a hash of the node position).
subroutine fill_orszag_tang(mesh, nn, xi, kappa, noise, state)
    type(mesh_t), intent(in) :: mesh
    integer, intent(in)      :: nn
    real(dp), intent(in)     :: xi(nn), kappa, noise
    real(dp), intent(out)    :: state(nn, nn, nn, 9, mesh%nelem)
    real(dp) :: x, y, z, v(3), b(3), h(3)
    integer(c_int64_t) :: e
    integer :: i, j, k, d
    do e = 1, mesh%nelem
        do k = 1, nn; do j = 1, nn; do i = 1, nn
            x = mesh%lower(1, e) + 0.5_dp*(xi(i) + 1.0_dp)*mesh%delta(1, e)
            y = mesh%lower(2, e) + 0.5_dp*(xi(j) + 1.0_dp)*mesh%delta(2, e)
            z = mesh%lower(3, e) + 0.5_dp*(xi(k) + 1.0_dp)*mesh%delta(3, e)
            do d = 1, 3
                h(d) = sin(12.9898_dp*x + 78.233_dp*y + 37.719_dp*z + 17.0_dp*d)*43758.5453_dp
                h(d) = 2.0_dp*(h(d) - floor(h(d))) - 1.0_dp
            end do
            v = (/ -sin(2.0_dp*PI*y), sin(2.0_dp*PI*x), 0.1_dp*sin(2.0_dp*PI*z) /) + noise*h
            b = (/ -sin(2.0_dp*PI*y), sin(4.0_dp*PI*x), 0.0_dp /)/sqrt(4.0_dp*PI)
            call prim2cons_mhd(kappa, 25.0_dp/(36.0_dp*PI), v, 5.0_dp/(12.0_dp*PI), b, 0.0_dp, &
                               state(i, j, k, :, e))
        end do; end do; end do
    end do
end subroutine

!! Raw dump: int64 nelem, int64 n_poly, int64 nvars, then the state array.
subroutine write_state(path, n_poly, nelem, state)
    character(len=*), intent(in) :: path
    integer, intent(in)          :: n_poly
    integer(c_int64_t), intent(in) :: nelem
    real(dp), intent(in)         :: state(:,:,:,:,:)
    integer :: unit
    open(newunit=unit, file=path, access='stream', form='unformatted', status='replace')
    write(unit) nelem, int(n_poly, c_int64_t), int(size(state, 4), c_int64_t)
    write(unit) state
    close(unit)
end subroutine

end module
