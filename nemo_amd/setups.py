"""Initial data for the benchmark configurations (host side, FP64).

The reference fills initial data per quad in each setup's hook_init_810
(e.g. setups/euler/3d/sedov/setup_init_mod.f90:7-58) from primitive
variables through prim2cons (equations/euler/equations_mod.f90:195-208).
These generators do the same on the uniform meshes of nemo_amd.mesh and
return the C-ABI state layout u[elem][var][node].

Random perturbations use a counter-based SplitMix64 stream keyed by
(seed, lexicographic element id, node, variable), so the data do not depend
on element order or on how the work is split.
"""
from __future__ import annotations

import numpy as np

from .mesh import Mesh, uniform_periodic
from . import operators

_GOLD = np.uint64(0x9E3779B97F4A7C15)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + _GOLD
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, lex_elem: np.ndarray, M3: int, var: int) -> np.ndarray:
    """Uniform [0,1) numbers, shape [nelem, M3], for one variable slot."""
    node = np.arange(M3, dtype=np.uint64)[None, :]
    key = (lex_elem.astype(np.uint64)[:, None] * np.uint64(M3) + node) * np.uint64(16) + np.uint64(var)
    base = _splitmix64(np.array([seed], dtype=np.uint64))[0]
    r = _splitmix64(base ^ key)
    return (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def prim2cons(rho, vx, vy, vz, p, gamma):
    """prim2cons: equations/euler/equations_mod.f90:195-208."""
    E = p / (gamma - 1.0) + 0.5 * rho * (vx ** 2 + vy ** 2 + vz ** 2)
    return np.stack([rho, vx * rho, vy * rho, vz * rho, E], axis=1)


def isentropic_vortex(mesh: Mesh, N: int, gamma: float = 1.4, strength: float = 5.0,
                      length: float = 0.1, center=(0.5, 0.5)) -> np.ndarray:
    """Isentropic vortex on the unit cube, uniform in z.

    Background rho = 1, p = 1, u = (1, 1, 0).  The classic vortex on
    [-5,5]^2 is scaled into the unit square with length scale `length`
    (radial balance dp/dr = rho v_theta^2 / r is invariant under this
    scaling)."""
    xi, _ = operators.lgl(N)
    X = mesh.node_coords(xi)
    dx = (X[:, 0] - center[0]) / length
    dy = (X[:, 1] - center[1]) / length
    r2 = dx * dx + dy * dy
    eps = strength
    du = eps / (2.0 * np.pi) * np.exp(0.5 * (1.0 - r2))
    T = 1.0 - (gamma - 1.0) * eps * eps / (8.0 * gamma * np.pi ** 2) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho ** gamma
    vx = 1.0 - du * dy
    vy = 1.0 + du * dx
    vz = np.zeros_like(rho)
    return np.ascontiguousarray(prim2cons(rho, vx, vy, vz, p, gamma))


def kelvin_helmholtz(mesh: Mesh, N: int, gamma: float = 1.4, seed: int = 1234,
                     noise: float = 0.01, width: float = 0.025) -> np.ndarray:
    """Kelvin-Helmholtz shear layer on the periodic unit cube.

    rho = 2 for |y-0.5| < 0.25 and 1 elsewhere; u_x = +0.5 inside the band
    and -0.5 outside, joined by tanh profiles of width `width`; p = 2.5; each
    velocity component carries seeded uniform noise in [-noise, noise]."""
    xi, _ = operators.lgl(N)
    M3 = (N + 1) ** 3
    X = mesh.node_coords(xi)
    y = X[:, 1]
    rho = np.where(np.abs(y - 0.5) < 0.25, 2.0, 1.0)
    band = 0.5 * (np.tanh((y - 0.25) / width) - np.tanh((y - 0.75) / width))
    lex = mesh.lex_id()
    vx = -0.5 + band + noise * (2.0 * uniform01(seed, lex, M3, 1) - 1.0)
    vy = noise * (2.0 * uniform01(seed, lex, M3, 2) - 1.0)
    vz = noise * (2.0 * uniform01(seed, lex, M3, 3) - 1.0)
    p = np.full_like(rho, 2.5)
    return np.ascontiguousarray(prim2cons(rho, vx, vy, vz, p, gamma))


def mhd_prim2cons(rho, v, p, B, psi, gamma, mu0=1.0):
    """prim2cons of equations/mhd/equations_mod.f90:294-310; v and B are
    [3, ...] arrays."""
    E = (p / (gamma - 1.0) + 0.5 * rho * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2)
         + 0.5 / mu0 * (B[0] ** 2 + B[1] ** 2 + B[2] ** 2 + psi ** 2))
    return np.stack([rho, rho * v[0], rho * v[1], rho * v[2], E, B[0], B[1], B[2], psi], axis=1)


def orszag_tang(mesh: Mesh, N: int, gamma: float = 5.0 / 3.0, seed: int = 1234, noise: float = 1.0e-3) -> np.ndarray:
    """Three-dimensional Orszag-Tang-type vortex on the periodic unit cube:
    rho = 25/(36 pi), p = 5/(12 pi), v = (-sin 2 pi y, sin 2 pi x, 0.1 sin 2 pi z),
    B = (-sin 2 pi y, sin 4 pi x, 0)/sqrt(4 pi), psi = 0 (mu0 = 1), each velocity
    component with seeded uniform noise in [-noise, noise].  (The reference's
    setups/mhd/3d/orszag-tang/setup_init_mod.f90:44-64 differs in its
    velocity and field.)"""
    xi, _ = operators.lgl(N)
    M3 = (N + 1) ** 3
    X = mesh.node_coords(xi)
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    tp = 2.0 * np.pi
    rho = np.full_like(x, 25.0 / (36.0 * np.pi))
    p = np.full_like(x, 5.0 / (12.0 * np.pi))
    lex = mesh.lex_id()
    v = np.stack([-np.sin(tp * y) + noise * (2.0 * uniform01(seed, lex, M3, 1) - 1.0),
                  np.sin(tp * x) + noise * (2.0 * uniform01(seed, lex, M3, 2) - 1.0),
                  0.1 * np.sin(tp * z) + noise * (2.0 * uniform01(seed, lex, M3, 3) - 1.0)])
    s = 1.0 / np.sqrt(4.0 * np.pi)
    B = np.stack([-np.sin(tp * y) * s, np.sin(2.0 * tp * x) * s, np.zeros_like(x)])
    return np.ascontiguousarray(mhd_prim2cons(rho, v, p, B, np.zeros_like(x), gamma))


def planar_riemann(mesh: Mesh, N: int, rng: np.random.Generator, gamma: float = 1.4) -> np.ndarray:
    """Planar Riemann problem on the periodic unit cube: a plane through the
    centre with a random normal separates two constant states with
    rho in [0.125, 1], p in [0.1, 1] and |u| <= 0.5 (random directions)."""
    xi, _ = operators.lgl(N)
    X = mesh.node_coords(xi)
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    side = ((X - 0.5) * n[None, :, None]).sum(1) < 0.0
    st = []
    for _ in range(2):
        rho = rng.uniform(0.125, 1.0)
        p = rng.uniform(0.1, 1.0)
        d = rng.normal(size=3)
        v = d / np.linalg.norm(d) * rng.uniform(0.0, 0.5)
        st.append((rho, v, p))
    rho = np.where(side, st[0][0], st[1][0])
    vx = np.where(side, st[0][1][0], st[1][1][0])
    vy = np.where(side, st[0][1][1], st[1][1][1])
    vz = np.where(side, st[0][1][2], st[1][1][2])
    p = np.where(side, st[0][2], st[1][2])
    return np.ascontiguousarray(prim2cons(rho, vx, vy, vz, p, gamma))


def ensemble_c5(nruns: int = 256, seed: int = 1234, nmin: int = 8, nmax: int = 32):
    """Config 5: `nruns` independent planar Riemann problems, elements per
    direction uniform in [nmin, nmax], N uniform in {2, 3, 4}.  Returns
    (meshes, n_polys, states)."""
    rng = np.random.default_rng(seed)
    meshes, npolys, states = [], [], []
    for _ in range(nruns):
        n = int(rng.integers(nmin, nmax + 1))
        N = int(rng.integers(2, 5))
        m = uniform_periodic(n)
        meshes.append(m)
        npolys.append(N)
        states.append(planar_riemann(m, N, rng))
    return meshes, npolys, states


def config_c1(n: int = 8, N: int = 3):
    """Config 1: 8^3 periodic unit cube, N=3, isentropic vortex, FP64."""
    mesh = uniform_periodic(n)
    return mesh, isentropic_vortex(mesh, N)


def config_c2(n: int = 64, N: int = 3, seed: int = 1234):
    """Config 2: 64^3 periodic unit cube, N=3, Kelvin-Helmholtz, FP64."""
    mesh = uniform_periodic(n)
    return mesh, kelvin_helmholtz(mesh, N, seed=seed)


def config_c3(n: int = 48, N: int = 3, seed: int = 1234):
    """Config 3: 48^3 periodic unit cube, N=3, GLM-MHD Orszag-Tang-type vortex."""
    mesh = uniform_periodic(n)
    return mesh, orszag_tang(mesh, N, seed=seed)
