"""Non-periodic boxes: the oracle's boundary faces (transmissive and the
Eulerian slip wall), on the CPU.

The slip-wall pressure is checked against the exact solution of the
reflected Riemann problem (root of Toro's pressure function, found
numerically); the assembled operator by free-stream states, by exact
conservation of what the walls cannot change (mass, energy, tangential
momentum) and by the normal-momentum balance of the wall pressures, and
with transmissive sides by the balance of the boundary fluxes f(u)."""
import numpy as np
import pytest
from scipy.optimize import brentq

from nemo_amd import mesh as meshmod
from oracle import oracle as O

GAMMA = 1.4


def _toro_f(p, rho, pk, gamma):
    """Toro's pressure function of one side (rarefaction below pk, shock above)."""
    c = np.sqrt(gamma * pk / rho)
    if p <= pk:
        return 2.0 * c / (gamma - 1.0) * ((p / pk) ** ((gamma - 1.0) / (2.0 * gamma)) - 1.0)
    A = 2.0 / ((gamma + 1.0) * rho)
    B = (gamma - 1.0) / (gamma + 1.0) * pk
    return (p - pk) * np.sqrt(A / (p + B))


def test_slip_wall_pressure_is_the_reflected_riemann_solution():
    """Inner state (rho, v toward the wall, p) against its mirror image
    (rho, -v, p): the star pressure solves 2 f(p*) = 2 v."""
    cfg = O.make_cfg(gamma=GAMMA)
    rng = np.random.default_rng(3)
    for _ in range(200):
        rho, p = 10 ** rng.uniform(-1, 1), 10 ** rng.uniform(-1, 1)
        c = np.sqrt(GAMMA * p / rho)
        v = rng.uniform(-0.8 * 2 * c / (GAMMA - 1.0), 3.0 * c)
        ps = O.slip_wall_pressure(cfg, rho, v, p)
        ref = brentq(lambda q: _toro_f(q, rho, p, GAMMA) - v, 1e-14 * p, 1e3 * p + 10 * rho * v * v,
                     xtol=1e-300, rtol=1e-15, maxiter=500)
        assert abs(ps - ref) <= 1e-12 * ref, (rho, v, p, ps, ref)
        assert (ps > p) == (v > 0)


def _rough(m, N, seed=1, amp=0.1):
    rng = np.random.default_rng(seed)
    M3 = (N + 1) ** 3
    rho = 1.0 + amp * rng.uniform(-1, 1, (m.nelem, M3))
    v = 0.3 * rng.uniform(-1, 1, (3, m.nelem, M3))
    p = 1.0 + amp * rng.uniform(-1, 1, (m.nelem, M3))
    E = p / (GAMMA - 1.0) + 0.5 * rho * (v ** 2).sum(0)
    return np.ascontiguousarray(np.stack([rho, rho * v[0], rho * v[1], rho * v[2], E], axis=1))


def _weights(m, N):
    _, w = O.lgl(N)
    w3 = (w[:, None, None] * w[None, :, None] * w[None, None, :]).reshape(-1)  # (k, j, i) = node order
    return w3[None, :] * (np.prod(m.delta, axis=1) / 8.0)[:, None]


def _face_nodes(N, f):
    """Node indices of face f of an element (i fastest), with their 2-D weights."""
    M = N + 1
    _, w = O.lgl(N)
    d, plus = f // 2, f % 2
    idx, wt = [], []
    for b in range(M):
        for a in range(M):
            ijk = [0, 0, 0]
            t = [x for x in range(3) if x != d]
            ijk[d] = M - 1 if plus else 0
            ijk[t[0]], ijk[t[1]] = a, b
            idx.append(ijk[0] + M * ijk[1] + M * M * ijk[2])
            wt.append(w[a] * w[b])
    return np.array(idx), np.array(wt)


SCHEMES = [(O.VOL_EC, O.SURF_HLLC, O.IND_PP), (O.VOL_WEAK, O.SURF_RUSANOV, O.IND_NONE),
           (O.VOL_CENTRAL, O.SURF_HLLC, O.IND_NONE)]


@pytest.mark.parametrize("N", [2, 3])
@pytest.mark.parametrize("vol,surf,ind", SCHEMES)
def test_free_stream_with_walls_and_transmissive_sides(N, vol, surf, ind):
    M3 = (N + 1) ** 3
    base = meshmod.uniform_periodic(3, 4, 2)
    u0 = np.array([1.2, 1.2 * 0.3, 0.0, 1.2 * -0.2, 2.5 / (GAMMA - 1.0) + 0.5 * 1.2 * (0.09 + 0.04)])
    u = np.ascontiguousarray(np.broadcast_to(u0[None, :, None], (base.nelem, 5, M3)))
    # walls across y (no normal velocity), x and z periodic
    m = meshmod.with_boundaries(base, [None, None, O.BC_SLIP_WALL, O.BC_SLIP_WALL, None, None])
    cfg = O.with_boundaries(O.make_cfg(N=N, volume=vol, surface=surf, indicator=ind, gamma=GAMMA), m.bfaces)
    r = O.rhs(cfg, m.nbr, m.delta, u, np.full(m.nelem, 0.3) if ind != O.IND_NONE else None)
    assert np.abs(r).max() <= 1e-12 * (N + 1) / m.delta.min()
    # every side transmissive: any constant state is steady
    u1 = u.copy()
    u1[:, 2] = 1.2 * 0.4
    u1[:, 4] += 0.5 * 1.2 * 0.16
    m = meshmod.with_boundaries(base, [O.BC_TRANSMISSIVE] * 6)
    cfg = O.with_boundaries(O.make_cfg(N=N, volume=vol, surface=surf, indicator=ind, gamma=GAMMA), m.bfaces)
    r = O.rhs(cfg, m.nbr, m.delta, u1, np.full(m.nelem, 0.3) if ind != O.IND_NONE else None)
    assert np.abs(r).max() <= 1e-12 * (N + 1) / m.delta.min()


@pytest.mark.parametrize("N", [2, 3])
@pytest.mark.parametrize("vol,surf,ind", SCHEMES)
def test_wall_balance(N, vol, surf, ind):
    """Slip walls across y on a rough state: mass, energy and the tangential
    momenta are conserved exactly; the y momentum changes by the wall
    pressures integrated over the two walls."""
    base = meshmod.uniform_periodic(3, 4, 2)
    m = meshmod.with_boundaries(base, [None, None, O.BC_SLIP_WALL, O.BC_SLIP_WALL, None, None])
    u = _rough(m, N)
    cfg = O.with_boundaries(O.make_cfg(N=N, volume=vol, surface=surf, indicator=ind, gamma=GAMMA), m.bfaces)
    r = O.rhs(cfg, m.nbr, m.delta, u, np.random.default_rng(2).uniform(0, 0.5, m.nelem)
              if ind != O.IND_NONE else None)
    W = _weights(m, N)
    tot = np.einsum("en,evn->v", W, r)
    scale = np.einsum("en,evn->v", W, np.abs(r))
    for h in (0, 1, 3, 4):
        assert abs(tot[h]) <= 1e-13 * scale[h], (h, tot[h], scale[h])
    expect = 0.0
    for e, f, _ in m.bfaces:
        idx, wt = _face_nodes(N, int(f))
        dA = m.delta[e, 0] * m.delta[e, 2] / 4.0
        for n, w2 in zip(idx, wt):
            ui = u[e, :, n]
            sgn = 1.0 if f % 2 else -1.0
            ps = O.slip_wall_pressure(cfg, ui[0], sgn * ui[2] / ui[0], O.pressure(cfg, ui))
            expect -= sgn * ps * w2 * dA
    assert abs(tot[2] - expect) <= 1e-12 * scale[2], (tot[2], expect)


@pytest.mark.parametrize("N", [2, 3])
@pytest.mark.parametrize("vol,surf,ind", SCHEMES)
def test_transmissive_balance(N, vol, surf, ind):
    """Every side transmissive: each conserved total changes by the outflow
    of the physical flux of the inner states, -sum over the boundary faces
    of the integral of f(u) . n."""
    base = meshmod.uniform_periodic(3, 2, 2)
    m = meshmod.with_boundaries(base, [O.BC_TRANSMISSIVE] * 6)
    u = _rough(m, N, seed=5)
    cfg = O.with_boundaries(O.make_cfg(N=N, volume=vol, surface=surf, indicator=ind, gamma=GAMMA), m.bfaces)
    r = O.rhs(cfg, m.nbr, m.delta, u, np.random.default_rng(4).uniform(0, 0.5, m.nelem)
              if ind != O.IND_NONE else None)
    W = _weights(m, N)
    tot = np.einsum("en,evn->v", W, r)
    scale = np.einsum("en,evn->v", W, np.abs(r))
    expect = np.zeros(5)
    for e, f, _ in m.bfaces:
        d, sgn = int(f) // 2, (1.0 if f % 2 else -1.0)
        idx, wt = _face_nodes(N, int(f))
        t = [x for x in range(3) if x != d]
        dA = m.delta[e, t[0]] * m.delta[e, t[1]] / 4.0
        for n, w2 in zip(idx, wt):
            expect -= sgn * O.flux(cfg, d, u[e, :, n]) * w2 * dA
    assert np.abs(tot - expect).max() <= 1e-12 * scale.max(), (tot, expect)


def test_mhd_transmissive_free_stream_and_mass():
    """GLM-MHD with transmissive sides: a constant state with psi = 0 is
    steady, and the mass changes by the outflow rho v . n."""
    N = 3
    M3 = (N + 1) ** 3
    base = meshmod.uniform_periodic(3, 2, 2)
    m = meshmod.with_boundaries(base, [O.BC_TRANSMISSIVE] * 6)
    cfg = O.with_boundaries(O.make_cfg(N=N, eq=O.EQ_MHD, volume=O.VOL_EC, surface=O.SURF_HLLD,
                                       indicator=O.IND_PP, timedisc=O.TD_LSRK54, gamma=5.0 / 3.0, mu0=1.0,
                                       glm_ch=0.5), m.bfaces)
    u0 = np.array([1.2, 0.36, -0.24, 0.12, 3.2, 0.4, 0.7, -0.5, 0.0])
    u = np.ascontiguousarray(np.broadcast_to(u0[None, :, None], (m.nelem, 9, M3)))
    r = O.rhs(cfg, m.nbr, m.delta, u, np.full(m.nelem, 0.3))
    assert np.abs(r).max() <= 1e-11 * (N + 1) / m.delta.min()
    rng = np.random.default_rng(6)
    u = u * (1.0 + 0.05 * rng.uniform(-1, 1, u.shape))
    r = O.rhs(cfg, m.nbr, m.delta, u, rng.uniform(0, 0.5, m.nelem))
    W = _weights(m, N)
    tot = (W * r[:, 0]).sum()
    expect = 0.0
    for e, f, _ in m.bfaces:
        d, sgn = int(f) // 2, (1.0 if f % 2 else -1.0)
        idx, wt = _face_nodes(N, int(f))
        t = [x for x in range(3) if x != d]
        dA = m.delta[e, t[0]] * m.delta[e, t[1]] / 4.0
        expect -= sgn * (u[e, 1 + d, idx] * wt).sum() * dA
    assert abs(tot - expect) <= 1e-12 * (W * np.abs(r[:, 0])).sum()


@pytest.mark.parametrize("kind", [O.BC_TRANSMISSIVE, O.BC_SLIP_WALL])
@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("vol,surf,ind", SCHEMES[:2])
def test_boundary_faces_against_periodic_images(kind, d, vol, surf, ind):
    """A state that does not vary along direction d and has no velocity along
    it meets its own values across the d faces: with boundary faces of either
    kind there (the wall's flux is then the normal flux p e_d of the state,
    the transmissive one f(u)) the RHS equals the periodic mesh's at every
    node.  The state varies differently along the two other directions, so a
    boundary flux placed at the wrong face node, side or component fails."""
    N = 3
    base = meshmod.uniform_periodic(3, 3, 3)
    kinds = [None] * 6
    kinds[2 * d] = kinds[2 * d + 1] = kind
    m = meshmod.with_boundaries(base, kinds)
    x, _ = O.lgl(N)
    X = base.node_coords(x)
    t0, t1 = [k for k in range(3) if k != d]
    a, b = X[:, t0], X[:, t1]
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * a) + 0.1 * np.cos(4 * np.pi * b)
    v = np.zeros((3,) + rho.shape)
    v[t0] = 0.3 * np.cos(2 * np.pi * b)
    v[t1] = -0.2 + 0.1 * np.sin(2 * np.pi * a * b)
    p = 1.0 + 0.3 * np.cos(2 * np.pi * (a + 2 * b))
    u = np.ascontiguousarray(np.stack([rho, rho * v[0], rho * v[1], rho * v[2],
                                       p / (GAMMA - 1.0) + 0.5 * rho * (v ** 2).sum(0)], axis=1))
    alpha = np.full(m.nelem, 0.3) if ind != O.IND_NONE else None
    base_cfg = O.make_cfg(N=N, volume=vol, surface=surf, indicator=ind, gamma=GAMMA)
    r_bc = O.rhs(O.with_boundaries(base_cfg, m.bfaces), m.nbr, m.delta, u, alpha)
    r_per = O.rhs(base_cfg, base.nbr, base.delta, u, alpha)
    assert np.abs(r_bc - r_per).max() <= 1e-12 * np.abs(r_per).max()
