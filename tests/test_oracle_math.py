"""Pins the oracle's composite operators where the reference has no
implementation (3-D LGL weak form, flux differencing, FV subcells,
blending, indicator, repair) by what the mathematics fixes.  Each check is
chosen so that a plausible mistake (dropped term, wrong sign or index,
transposed operand) fails it."""
import numpy as np
import pytest

from nemo_amd import mesh as meshmod, setups
from oracle import oracle as O

GAMMA = 1.4


def _cfg(**kw):
    base = dict(N=3, volume=O.VOL_WEAK, surface=O.SURF_RUSANOV, indicator=O.IND_NONE, timedisc=O.TD_LSRK54)
    base.update(kw)
    return O.make_cfg(**base)


def _quad_weights(N, m):
    _, w = O.lgl(N)
    W = np.einsum("i,j,k->kji", w, w, w).reshape(-1)  # node = i + M j + M^2 k
    vol = np.prod(m.delta, axis=1) / 8.0
    return W[None, :] * vol[:, None]


def _rough(m, N, seed=1, amp=0.2):
    rng = np.random.default_rng(seed)
    M3 = (N + 1) ** 3
    rho = 1.0 + amp * rng.uniform(-1, 1, (m.nelem, M3))
    v = amp * rng.uniform(-1, 1, (3, m.nelem, M3))
    p = 1.0 + amp * rng.uniform(-1, 1, (m.nelem, M3))
    return np.ascontiguousarray(setups.prim2cons(rho, v[0], v[1], v[2], p, GAMMA))


# ---- LGL operators -----------------------------------------------------------

def test_lgl_closed_form_n3():
    x, w = O.lgl(3)
    assert np.allclose(x, [-1, -1 / np.sqrt(5), 1 / np.sqrt(5), 1], rtol=0, atol=4e-16)
    assert np.allclose(w, [1 / 6, 5 / 6, 5 / 6, 1 / 6], rtol=0, atol=4e-16)


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_diff_matrix_identities(N):
    x, w = O.lgl(N)
    D = O.diff_matrix(N, x)
    assert np.allclose(D @ np.ones(N + 1), 0, atol=1e-13)
    for k in range(1, N + 1):  # exact on polynomials of degree <= N
        assert np.allclose(D @ x ** k, k * x ** (k - 1), atol=1e-12)
    # SBP: W D + D^T W = B = diag(-1, 0, ..., 0, 1)
    W = np.diag(w)
    B = np.zeros((N + 1, N + 1))
    B[0, 0], B[-1, -1] = -1, 1
    assert np.allclose(W @ D + D.T @ W, B, atol=1e-13)


@pytest.mark.parametrize("N", [2, 3, 4])
def test_vandermonde_orthonormal(N):
    x, w = O.lgl(N)
    V, Vi = O.vandermonde(N, x)
    G = V.T @ np.diag(w) @ V  # LGL quadrature exact below degree 2N
    off = np.ones_like(G, dtype=bool)
    off[N, N] = False
    assert np.allclose(G[off], np.eye(N + 1)[off], atol=1e-13)
    assert np.allclose(Vi @ V, np.eye(N + 1), atol=1e-13)


# ---- pointwise -----------------------------------------------------------------

def test_riemann_consistency_and_upwinding():
    cfg = _cfg()
    rng = np.random.default_rng(0)
    for _ in range(50):
        u = _rough(meshmod.uniform_periodic(1), 0, seed=rng.integers(1 << 30))[0, :, 0]
        for d in range(3):
            f = O.flux(cfg, d, u)
            for k in (O.SURF_RUSANOV, O.SURF_HLLC):
                assert np.allclose(O.riemann(cfg, k, d, u, u), f, rtol=1e-14, atol=1e-14)
            for k in (O.VOL_EC, O.VOL_CENTRAL):
                assert np.allclose(O.two_point(cfg, k, d, u, u), f, rtol=1e-12, atol=1e-13)
    # supersonic to the right: the HLLC hybrid returns the left flux
    uL = np.array([1.0, 5.0, 0.1, 0.2, 1 / 0.4 + 0.5 * (25 + 0.05)])
    uR = np.array([0.5, 2.0, 0.0, 0.1, 0.5 / 0.4 + 0.5 * (16 + 0.02) * 0.5 / 0.25 * 0.25])
    assert np.array_equal(O.riemann(cfg, O.SURF_HLLC, 0, uL, uR), O.flux(cfg, 0, uL))


def test_ec_flux_symmetry_and_tadmor_condition():
    """f#(u,v) = f#(v,u) and (w_v - w_u).f#(u,v) = psi_v - psi_u, psi = rho v_n
    for the entropy S = -rho s/(gamma-1) of equations_mod.f90:212-286."""
    cfg = _cfg()
    rng = np.random.default_rng(3)
    for _ in range(200):
        a = _rough(meshmod.uniform_periodic(1), 0, seed=rng.integers(1 << 30), amp=0.6)[0, :, 0]
        b = _rough(meshmod.uniform_periodic(1), 0, seed=rng.integers(1 << 30), amp=0.6)[0, :, 0]
        for d in range(3):
            f = O.two_point(cfg, O.VOL_EC, d, a, b)
            assert np.allclose(f, O.two_point(cfg, O.VOL_EC, d, b, a), rtol=1e-13, atol=1e-14)
            lhs = (O.cons2evec(cfg, b) - O.cons2evec(cfg, a)) @ f
            rhs = b[1 + d] - a[1 + d]
            assert abs(lhs - rhs) <= 1e-11 * (1 + abs(rhs))


def test_logmean_branches():
    for a, b in [(1.0, 2.0), (1.0, 1.00005), (0.3, 0.30001), (5.0, 4.9), (1e-3, 2e-3)]:
        exact = (a - b) / (np.log(a) - np.log(b)) if a != b else a
        assert abs(O.logmean(a, b) - exact) <= 1e-12 * exact


# ---- element operators ---------------------------------------------------------

@pytest.mark.parametrize("vol", [O.VOL_WEAK, O.VOL_CENTRAL, O.VOL_EC])
@pytest.mark.parametrize("surf", [O.SURF_RUSANOV, O.SURF_HLLC])
@pytest.mark.parametrize("N", [2, 3, 4])
def test_free_stream_preservation(vol, surf, N):
    m = meshmod.uniform_periodic(3)
    M3 = (N + 1) ** 3
    u = np.zeros((m.nelem, 5, M3))
    u[:, 0], u[:, 1], u[:, 2], u[:, 3] = 1.3, 1.3 * 0.7, -1.3 * 0.2, 1.3 * 0.4
    u[:, 4] = 2.0 / 0.4 + 0.5 * 1.3 * (0.49 + 0.04 + 0.16)
    alpha = np.linspace(0, 1, m.nelem)
    r = O.rhs(_cfg(N=N, volume=vol, surface=surf, indicator=O.IND_PP), m.nbr, m.delta, u, alpha)
    assert np.abs(r).max() < 1e-11


@pytest.mark.parametrize("N", [2, 3, 4])
def test_weak_form_equals_central_flux_differencing(N):
    """On LGL nodes the SBP property makes the weak form and split-form flux
    differencing with the central two-point flux the same operator."""
    m = meshmod.uniform_periodic(3)
    u = setups.isentropic_vortex(m, N, length=0.3)
    a = O.rhs(_cfg(N=N, volume=O.VOL_WEAK), m.nbr, m.delta, u)
    b = O.rhs(_cfg(N=N, volume=O.VOL_CENTRAL), m.nbr, m.delta, u)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()


@pytest.mark.parametrize("vol", [O.VOL_WEAK, O.VOL_EC])
@pytest.mark.parametrize("surf", [O.SURF_RUSANOV, O.SURF_HLLC])
def test_conservation_for_any_alpha(vol, surf):
    """sum over elements and nodes of w J rhs = 0 (periodic) for any blending,
    the invariant of setups/euler/3d/conservation-test (random blend
    factors, setup_modify_limiter_mod.f90:31-40)."""
    N = 3
    m = meshmod.uniform_periodic(4)
    u = _rough(m, N, seed=11)
    rng = np.random.default_rng(5)
    alpha = rng.uniform(0, 1, m.nelem)
    r = O.rhs(_cfg(N=N, volume=vol, surface=surf, indicator=O.IND_PP), m.nbr, m.delta, u, alpha)
    Wq = _quad_weights(N, m)
    for v in range(5):
        tot = (Wq * r[:, v]).sum()
        scale = (Wq * np.abs(r[:, v])).sum()
        assert abs(tot) <= 1e-13 * scale


def test_entropy_conservation_of_flux_differencing():
    """With the EC flux at the faces as well, the semi-discrete entropy
    change sum w J (dS/du . rhs) vanishes; with Rusanov faces it is <= 0."""
    N = 3
    m = meshmod.uniform_periodic(3)
    u = _rough(m, N, seed=2, amp=0.3)
    Wq = _quad_weights(N, m)
    cfg_ec = _cfg(volume=O.VOL_EC, surface=O.SURF_EC_TEST)
    cfg_rus = _cfg(volume=O.VOL_EC, surface=O.SURF_RUSANOV)
    w = np.stack([[O.cons2evec(cfg_ec, u[e, :, n]) for n in range(64)] for e in range(m.nelem)])  # [e,n,v]
    for cfg, sign in ((cfg_ec, 0), (cfg_rus, -1)):
        r = O.rhs(cfg, m.nbr, m.delta, u)
        dS = (Wq * np.einsum("env,evn->en", w, r)).sum()
        scale = (Wq * np.abs(np.einsum("env,evn->en", w, r))).sum()
        if sign == 0:
            assert abs(dS) <= 1e-12 * scale
        else:
            assert dS < -1e-6 * scale


def test_fv_subcell_operator_against_1d_construction():
    """alpha = 1 with data varying only in x: the blended operator is the
    first-order FV update on the LGL sub-grid, built here directly from the
    pointwise Riemann solver: rhs_i = -(2/dx)(F_{i+1/2} - F_{i-1/2})/w_i."""
    N, M = 3, 4
    m = meshmod.uniform_periodic(3)
    cfg = _cfg(N=N, volume=O.VOL_EC, surface=O.SURF_HLLC, indicator=O.IND_PP)
    x, w = O.lgl(N)
    X = m.node_coords(x)[:, 0]
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * X)
    vx = 0.2 * np.cos(2 * np.pi * X)
    p = 1.0 + 0.2 * np.sin(4 * np.pi * X)
    u = np.ascontiguousarray(setups.prim2cons(rho, vx, 0 * vx, 0 * vx, p, GAMMA))
    r = O.rhs(cfg, m.nbr, m.delta, u, np.ones(m.nelem))
    for e in range(m.nelem):
        em, ep = m.nbr[e, 0], m.nbr[e, 1]
        line = [u[e, :, i] for i in range(M)]
        F = [O.riemann(cfg, O.SURF_HLLC, 0, u[em, :, M - 1], line[0])]
        F += [O.riemann(cfg, O.SURF_HLLC, 0, line[i], line[i + 1]) for i in range(M - 1)]
        F += [O.riemann(cfg, O.SURF_HLLC, 0, line[M - 1], u[ep, :, 0])]
        jac = 2.0 / m.delta[e, 0]
        for i in range(M):
            expect = -jac * (F[i + 1] - F[i]) / w[i]
            for jk in range(M * M):  # every x-line of the element is the same
                assert np.allclose(r[e, :, i + M * jk], expect, rtol=1e-12, atol=1e-12)


def test_alpha_zero_is_pure_dg_and_blend_is_affine():
    N = 3
    m = meshmod.uniform_periodic(3)
    u = _rough(m, N, seed=4)
    c_b = _cfg(volume=O.VOL_EC, surface=O.SURF_HLLC, indicator=O.IND_PP)
    c_dg = _cfg(volume=O.VOL_EC, surface=O.SURF_HLLC)
    r0 = O.rhs(c_b, m.nbr, m.delta, u, np.zeros(m.nelem))
    assert np.array_equal(r0, O.rhs(c_dg, m.nbr, m.delta, u))
    r1 = O.rhs(c_b, m.nbr, m.delta, u, np.ones(m.nelem))
    rh = O.rhs(c_b, m.nbr, m.delta, u, np.full(m.nelem, 0.3))
    assert np.abs(rh - (0.7 * r0 + 0.3 * r1)).max() <= 1e-13 * np.abs(r1).max()


def test_indicator_closed_forms():
    N, M = 3, 4
    m = meshmod.uniform_periodic(1)
    x, _ = O.lgl(N)
    cfg = _cfg(indicator=O.IND_PP, alpha_max=0.5, alpha_min=0.001)
    # constant state: no high-mode energy -> alpha = 0
    u = np.zeros((1, 5, M ** 3))
    u[0, 0], u[0, 4] = 1.0, 2.5
    a, e = O.indicator(cfg, u)
    assert a[0] == 0.0 and abs(e[0]) < 1e-28
    # rho = 1, p = p0 + b Lhat_N(x): rho*p has modes (0,0,0) = 2^1.5 p0 and
    # (N,0,0) = 2 b, so energy = (2b)^2 / ((2^1.5 p0)^2 + (2b)^2)
    T = 0.5 * 10 ** (-1.8 * M ** 0.25)
    s = np.log((1 - 1e-4) / 1e-4)
    for b in (0.02, 0.06, 0.1, 0.5):
        p0 = 1.0
        LN = np.array([O.legendre(N, xi)[0] for xi in x])
        X = np.tile(LN, M * M)  # node = i + M j + M^2 k, varies with i only
        p = p0 + b * X
        u = np.ascontiguousarray(setups.prim2cons(np.ones((1, M ** 3)), 0 * p[None], 0 * p[None], 0 * p[None],
                                                  p[None], GAMMA))
        a, e = O.indicator(cfg, u)
        E = (2 * b) ** 2 / ((2 ** 1.5 * p0) ** 2 + (2 * b) ** 2)
        assert abs(e[0] - E) <= 1e-12
        al = 1 / (1 + np.exp(-s / T * (E - T)))
        al = 0.0 if al < 0.001 else (1.0 if al > 0.999 else al)
        assert abs(a[0] - min(al, 0.5)) <= 1e-9
    # energy in the second shell only (p = p0 + b Lhat_{N-1}(x) + c Lhat_N(y)
    # with c small): Hennemann's second ratio second / (total - top) decides,
    # E = max(top/total, second/(total - top))
    for b, c in ((0.05, 0.0), (0.08, 0.01), (0.2, 0.002)):
        LN1 = np.array([O.legendre(N - 1, xi)[0] for xi in x])
        LN = np.array([O.legendre(N, xi)[0] for xi in x])
        i = np.arange(M ** 3) % M
        j = (np.arange(M ** 3) // M) % M
        p = 1.0 + b * LN1[i] + c * LN[j]
        u = np.ascontiguousarray(setups.prim2cons(np.ones((1, M ** 3)), 0 * p[None], 0 * p[None], 0 * p[None],
                                                  p[None], GAMMA))
        a, e = O.indicator(cfg, u)
        e0, e1, e2 = (2 ** 1.5) ** 2, (2 * b) ** 2, (2 * c) ** 2
        E = max(e2 / (e0 + e1 + e2), e1 / (e0 + e1))
        assert abs(e[0] - E) <= 1e-12, (e[0], E)


@pytest.mark.parametrize("timedisc", [O.TD_LSRK54, O.TD_SSPRK54])
def test_rk_order_conditions(timedisc):
    """Butcher form of the reference's stage recursions satisfies all eight
    4th-order conditions."""
    a_rk, b_rk, c_rk = O.lsrk54_coeffs()
    k = O.ssprk54_coeffs()
    S = 5
    A = np.zeros((S + 1, S))  # row s: state after s stages = u0 + sum_j A[s,j] dt F(stage j input)
    if timedisc == O.TD_LSRK54:
        # reg_s = a_s reg_{s-1} + F(u_{s-1});  u_s = u_{s-1} + c_s reg_s
        reg = np.zeros(S)
        for s in range(S):
            reg = a_rk[s] * reg
            reg[s] += 1.0
            A[s + 1] = A[s] + c_rk[s] * reg
    else:
        A3, B1, B2, B3, C1, C2, C3 = k[3], k[5], k[6], k[7], k[9], k[10], k[11]
        D1, D2, D3, E1, E2, E3, E4, E5, E6 = k[13], k[14], k[15], k[17], k[18], k[19], k[20], k[21], k[22]
        e = np.eye(S)
        A[1] = A3 * e[0]
        A[2] = B2 * A[1] + B3 * e[1]
        A[3] = C2 * A[2] + C3 * e[2]
        A[4] = D2 * A[3] + D3 * e[3]
        A[5] = E2 * A[4] + E3 * e[4] + E4 * A[2] + E5 * A[3] + E6 * e[3]
        # consistency of the convex combinations
        for row in ((B1 + B2), (C1 + C2), (D1 + D2), (E1 + E2 + E4 + E5)):
            assert abs(row - 1.0) < 1e-13
    a = A[:S, :]
    b = A[S]
    c = a.sum(1)
    # the SSP constants are given to 14 digits (ssp/4-5/timedisc_const_mod.f90:5-32)
    # and meet the conditions to ~1e-10
    tol = 1e-12 if timedisc == O.TD_LSRK54 else 5e-10
    assert abs(b.sum() - 1) < tol
    assert abs(b @ c - 1 / 2) < tol
    assert abs(b @ c ** 2 - 1 / 3) < tol
    assert abs(b @ a @ c - 1 / 6) < tol
    assert abs(b @ c ** 3 - 1 / 4) < tol
    assert abs((b * c) @ a @ c - 1 / 8) < tol
    assert abs(b @ a @ c ** 2 - 1 / 12) < tol
    assert abs(b @ a @ a @ c - 1 / 24) < tol
    if timedisc == O.TD_LSRK54:
        assert np.allclose(c, b_rk, atol=1e-12)  # b_rk are the stage times


def test_repair_invariants():
    N = 3
    cfg = _cfg(repair=1)
    m = meshmod.uniform_periodic(1)
    u = _rough(m, N, seed=9)
    ok, out = O.repair(cfg, u[0].copy())
    assert ok and np.array_equal(out, u[0])  # valid element untouched
    bad = u[0].copy()
    nodes = [3, 17, 40]
    ke = 0.5 * (bad[1, nodes] ** 2 + bad[2, nodes] ** 2 + bad[3, nodes] ** 2) / bad[0, nodes]
    bad[4, nodes] = 0.5 * ke  # total energy below kinetic energy: p < 0
    assert not all(O.isvalid(cfg, bad[:, n]) for n in range(64))
    ok, out = O.repair(cfg, bad.copy())
    assert ok
    assert all(O.isvalid(cfg, out[:, n]) for n in range(64))
    Wq = _quad_weights(N, m)[0]
    for v in range(5):
        assert abs((Wq * out[v]).sum() - (Wq * bad[v]).sum()) <= 1e-14 * (Wq * np.abs(bad[v])).sum()


def test_cfl_closed_form():
    N = 3
    m = meshmod.uniform_periodic(2, 2, 2, extent=(1.0, 2.0, 0.5))
    M3 = (N + 1) ** 3
    rho, v, p = 1.2, np.array([0.3, -0.5, 0.1]), 0.8
    u = np.zeros((m.nelem, 5, M3))
    u[:, 0], u[:, 1:4] = rho, (rho * v)[None, :, None]
    u[:, 4] = p / 0.4 + 0.5 * rho * v @ v
    cfg = _cfg(cfl=0.5)
    dt, lm, vm = O.calc_timestep(cfg, m.delta, u)
    c = np.sqrt(GAMMA * p / rho)
    lam = np.abs(v) + c
    expect = 0.5 * 0.5 / 7 * np.min(m.delta[0] / lam)
    assert abs(dt - expect) <= 1e-14 * expect
    assert abs(lm - lam.max()) <= 1e-14 and abs(vm - np.abs(v).max()) <= 1e-15


def test_isentropic_vortex_convergence():
    """Weak-form DGSEM (the vortex scheme's operator) converges at about order N+1 on
    the advected isentropic vortex (exact solution: the initial vortex
    translated by u_inf t)."""
    N = 3
    T = 0.04
    errs = []
    for n in (16, 32, 64):
        m = meshmod.uniform_periodic(n, n, 1)
        u = setups.isentropic_vortex(m, N)
        cfg = _cfg(N=N, volume=O.VOL_WEAK, surface=O.SURF_RUSANOV, timedisc=O.TD_SSPRK54, cfl=0.5)
        r0, r1 = np.zeros_like(u), np.zeros_like(u)
        t = 0.0
        while t < T - 1e-14:
            dt, _, _ = O.calc_timestep(cfg, m.delta, u)
            dt = min(dt, T - t)
            assert O.step(cfg, m.nbr, m.delta, u, r0, r1, dt) == 0
            t += dt
        ex = setups.isentropic_vortex(m, N, center=(0.5 + T, 0.5 + T))
        errs.append(np.abs(u[:, 0] - ex[:, 0]).max())
    eoc = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert errs[-1] < 5e-5
    assert eoc[-1] > 3.0, (errs, eoc)


# ---- temporal order of the oracle's stage recursions -------------------------

def _smooth_periodic(m, N):
    """Smooth, non-uniform periodic state (all five variables vary)."""
    xi, _ = O.lgl(N)
    X = m.node_coords(xi)
    s = np.sin(2 * np.pi * (X[:, 0] + X[:, 1] + X[:, 2]))
    c = np.cos(2 * np.pi * (X[:, 0] - X[:, 1]))
    return np.ascontiguousarray(setups.prim2cons(1.0 + 0.2 * s, 0.3 + 0.1 * c, 0.2 - 0.1 * s, 0.1 + 0.05 * c,
                                                 1.0 + 0.1 * c, GAMMA))


def _temporal_eoc(step, cfg, m, u0, T, counts=(2, 4, 8, 16, 32)):
    """Observed orders log2(d_k / d_{k+1}) of the differences d_k between
    solutions at T with T/n_k and T/n_{k+1} steps of `step`."""
    sols = []
    for n in counts:
        u = u0.copy()
        r0, r1 = np.zeros_like(u), np.zeros_like(u)
        for _ in range(n):
            if step(cfg, m.nbr, m.delta, u, r0, r1, T / n) != 0:
                return None
        sols.append(u)
    d = [np.abs(sols[i] - sols[i + 1]).max() for i in range(len(sols) - 1)]
    if not np.all(np.isfinite(d)) or min(d) <= 0:
        return None
    return [float(np.log2(d[i] / d[i + 1])) for i in range(len(d) - 1)]


_ORDER_CASES = [(O.TD_LSRK54, O.VOL_WEAK, O.SURF_RUSANOV), (O.TD_LSRK54, O.VOL_EC, O.SURF_HLLC),
                (O.TD_SSPRK54, O.VOL_WEAK, O.SURF_RUSANOV), (O.TD_SSPRK54, O.VOL_EC, O.SURF_HLLC)]


@pytest.mark.parametrize("timedisc,volume,surface", _ORDER_CASES)
def test_oracle_step_temporal_order(timedisc, volume, surface):
    """O.step itself (not a restatement of its recursion) converges at
    fourth order in time on a fixed mesh: the semi-discrete ODE is the same
    for every dt, so successive differences of the solutions at a fixed time
    shrink by 2^4 per halving of dt (Richardson).  Pins the oracle's LSRK
    5-4 probe and register recursion (low-storage/4-5/timedisc_probe_mod.f90:
    31-37, timedisc_hooks_mod.f90:39-47) and the SSPRK(5,4) combinations
    (ssp/4-5/timedisc_probe_mod.f90:24-46, timedisc_hooks_mod.f90:40-67)."""
    N = 3
    m = meshmod.uniform_periodic(4, 4, 4)
    u0 = _smooth_periodic(m, N)
    cfg = _cfg(N=N, volume=volume, surface=surface, timedisc=timedisc, cfl=0.5)
    dt0, _, _ = O.calc_timestep(cfg, m.delta, u0)
    eoc = _temporal_eoc(O.step, cfg, m, u0, 8 * dt0)
    assert eoc is not None and min(eoc[1:]) >= 3.8, eoc


# Mistakes planted in a private copy of the oracle's stage recursion; each
# must make the order test above fail.
_PLANTED = [
    ("lsrk: a_s mis-indexed", O.TD_LSRK54,
     "la[s - 1] * q0[n] + fe[n];", "la[s % 5] * q0[n] + fe[n];"),
    ("lsrk: c_s mis-indexed", O.TD_LSRK54,
     "st[n] = ue[n] + lc[s - 1] * dt * reg;", "st[n] = ue[n] + lc[s % 5] * dt * reg;"),
    ("lsrk: old register used", O.TD_LSRK54,
     "st[n] = ue[n] + lc[s - 1] * dt * reg;", "st[n] = ue[n] + lc[s - 1] * dt * q0[n];"),
    ("lsrk: a and c swapped", O.TD_LSRK54,
     "la[s - 1] * q0[n] + fe[n];", "lc[s - 1] * q0[n] + fe[n];"),
    ("ssp: B1/B2 swapped", O.TD_SSPRK54,
     "case 2: st[n] = B1 * q0[n] + B2 * ue[n]", "case 2: st[n] = B2 * q0[n] + B1 * ue[n]"),
    ("ssp: E5 register term dropped", O.TD_SSPRK54,
     "case 3: q1[n] = q1[n] + E5 * st[n];", "case 3: q1[n] = q1[n];"),
]


@pytest.mark.parametrize("name,timedisc,old,new", _PLANTED, ids=[p[0] for p in _PLANTED])
def test_planted_rk_mistakes_fail_the_order_test(tmp_path, name, timedisc, old, new):
    import ctypes
    import os
    import shutil
    import subprocess
    src = open(os.path.join(os.path.dirname(O.__file__), "nemo_oracle.c")).read()
    assert src.count(old) == 1, name
    d = tmp_path / "mut"
    d.mkdir()
    (d / "nemo_oracle.c").write_text(src.replace(old, new))
    shutil.copy(os.path.join(os.path.dirname(O.__file__), "nemo_oracle.h"), d / "nemo_oracle.h")
    so = d / "libmut.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-fopenmp", "-std=c11", "-shared", "-o", str(so),
                    str(d / "nemo_oracle.c"), "-lm"], check=True)
    L = ctypes.CDLL(str(so))
    P, PI = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)
    L.orc_step.argtypes = [ctypes.POINTER(O.Cfg), ctypes.c_int64, PI, P, P, P, P, ctypes.c_double]

    def mut_step(cfg, nbr, delta, u, r0, r1, dt):
        nb = np.ascontiguousarray(nbr, dtype=np.int32)
        de = np.ascontiguousarray(delta, dtype=np.float64)
        return L.orc_step(ctypes.byref(cfg), u.shape[0], nb.ctypes.data_as(PI), de.ctypes.data_as(P),
                          u.ctypes.data_as(P), r0.ctypes.data_as(P), r1.ctypes.data_as(P), dt)

    N = 3
    m = meshmod.uniform_periodic(4, 4, 4)
    u0 = _smooth_periodic(m, N)
    cfg = _cfg(N=N, volume=O.VOL_WEAK, surface=O.SURF_RUSANOV, timedisc=timedisc, cfl=0.5)
    dt0, _, _ = O.calc_timestep(cfg, m.delta, u0)
    eoc = _temporal_eoc(mut_step, cfg, m, u0, 8 * dt0)
    assert eoc is None or min(eoc[1:]) < 3.8, (name, eoc)
