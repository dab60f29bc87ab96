"""Adaptive mesh (nemo_amd.amr) and the oracle's non-conforming faces and
state transfer, on the CPU.

The forest is checked against geometry (face planes, half-face positions,
2:1 balance, symmetric conforming neighbours) and against the uniform mesh
module; the oracle's 2:1 operators against Gauss-quadrature L2 projections;
its mortar fluxes against closed forms: free-stream preservation,
conservation, and an exactly representable state whose RHS is known
(-grad p at rest), which fails for a wrong half, orientation, slot or
transposed projection."""
import numpy as np
import pytest

from nemo_amd import amr, indicator_mlp, mesh as meshmod
from oracle import oracle as O


def _forest(seed=0, lmin=2, lmax=4):
    """Two rounds of refinement around a point, balanced."""
    f = amr.Forest.uniform(lmin, maxlevel=lmax)
    c0 = np.array([0.45, 0.4, 0.5])
    for r in (0.3, 0.15):
        cen = f.lower + 0.5 * f.delta
        f = f.refine(np.linalg.norm(cen - c0, axis=1) < r).balance()
    return f


def test_uniform_forest_matches_uniform_mesh():
    for level in (1, 2, 3):
        f = amr.Forest.uniform(level)
        m = meshmod.uniform_periodic(1 << level)
        nbr, mort = f.connectivity()
        assert mort.shape == (0, 6)
        assert np.array_equal(nbr, m.nbr)
        assert np.allclose(f.delta, m.delta) and np.allclose(f.lower, m.lower)


def test_forest_geometry_and_balance():
    f = _forest()
    nbr, mort = f.connectivity()
    lo, d = f.lower, f.delta
    assert len(mort) > 0 and len(set(f.level.tolist())) == 3
    # conforming neighbours: symmetric, same size, touching (periodic)
    for e in range(f.nelem):
        for face in range(6):
            n = nbr[e, face]
            if n < 0:
                continue
            ax, side = face // 2, face % 2
            assert nbr[n, face ^ 1] == e
            assert f.level[n] == f.level[e]
            gap = (lo[n, ax] - (lo[e, ax] + d[e, ax])) if side else (lo[e, ax] - (lo[n, ax] + d[n, ax]))
            assert abs(gap - round(gap)) < 1e-12
    # mortars: the four halves tile the big face in the documented order
    seen = set()
    for m, (big, face, *sm) in enumerate(mort):
        ax, side = face // 2, face % 2
        t1, t2 = amr.TANGENTIAL[ax]
        assert nbr[big, face] == -(5 * m + 1)
        plane = lo[big, ax] + (d[big, ax] if side else 0.0)
        for k, s in enumerate(sm):
            assert nbr[s, face ^ 1] == -(5 * m + 2 + k)
            assert f.level[s] == f.level[big] + 1
            sp = lo[s, ax] + (0.0 if side else d[s, ax])
            assert abs((sp - plane) - round(sp - plane)) < 1e-12
            assert abs(lo[s, t1] - (lo[big, t1] + (k & 1) * 0.5 * d[big, t1])) < 1e-12
            assert abs(lo[s, t2] - (lo[big, t2] + (k >> 1) * 0.5 * d[big, t2])) < 1e-12
            seen.add((int(s), int(face ^ 1)))
        seen.add((int(big), int(face)))
    assert len(seen) == 5 * len(mort) == int((nbr < 0).sum())


def test_balance_enforced():
    f = amr.Forest.uniform(2, maxlevel=5)
    # refine one leaf three times without balancing -> 2:1 violated
    for _ in range(3):
        flags = np.zeros(f.nelem, dtype=bool)
        flags[np.argmax(f.level)] = True
        f = f.refine(flags)
    with pytest.raises(ValueError):
        f.connectivity()
    b = f.balance()
    nbr, _ = b.connectivity()
    assert b.nelem > f.nelem


def test_adapt_units_and_round_trip():
    f = _forest(lmax=5)
    cen = f.lower + 0.5 * f.delta
    ref = np.linalg.norm(cen - 0.45, axis=1) < 0.1
    g, units = f.adapt(ref, np.zeros(f.nelem, bool))
    assert units is not None
    kinds = np.bincount(units[:, 0], minlength=3)
    assert kinds[amr.UNIT_REFINE] >= ref.sum() and kinds[amr.UNIT_COARSEN] == 0
    # every old leaf is consumed once, every new leaf produced once
    no = np.where(units[:, 0] == amr.UNIT_COARSEN, 8, 1).sum()
    nn = np.where(units[:, 0] == amr.UNIT_REFINE, 8, 1).sum()
    assert no == f.nelem and nn == g.nelem
    # coarsening everywhere merges families but never goes below minlevel
    assert g.family_starts().sum() > 0
    back, units3 = g.adapt(np.zeros(g.nelem, bool), np.ones(g.nelem, bool))
    assert back.nelem < g.nelem and (units3[:, 0] == amr.UNIT_COARSEN).any()
    assert back.level.min() >= g.minlevel
    back.connectivity()  # still 2:1 balanced


def test_adapt_pruning_gives_the_same_forest():
    """Dropping the coarsenings that balance would undo (adapt's pruning)
    changes neither the new forest nor the replacement units, on random
    refine/coarsen flags over random balanced forests; a steady no-op pass
    is recognised as such."""
    rng = np.random.default_rng(5)
    checked = changed = 0
    for trial in range(24):
        f = amr.Forest.uniform(2 + trial % 2, maxlevel=5 + trial % 2)
        for lv in range(3):
            cen = f.lower + 0.5 * f.delta
            f, _ = f.adapt(np.linalg.norm(cen - rng.random(3), axis=1) < 0.3 - 0.07 * lv + 0.05 * rng.random(),
                           np.zeros(f.nelem, bool))
        for k in range(4):
            refine = rng.random(f.nelem) < rng.random() * 0.3 * (k % 2)
            coarsen = rng.random(f.nelem) < rng.random()
            a, ua = f.adapt(refine, coarsen, prune=False)
            b, ub = f.adapt(refine, coarsen, prune=True)
            assert (ua is None) == (ub is None)
            assert np.array_equal(a.level, b.level) and np.array_equal(a.anchor, b.anchor)
            if ua is not None:
                assert np.array_equal(ua, ub)
                changed += 1
            checked += 1
            f = a
    assert checked == 96 and changed > 20
    # steady state: finest elements flagged at maxlevel, the rest flagged for
    # coarsening next to finer ones -- nothing changes, with or without pruning
    f = amr.Forest.uniform(3, minlevel=3, maxlevel=6)
    for lv in range(3):
        cen = f.lower + 0.5 * f.delta
        f, _ = f.adapt(np.linalg.norm(cen - 0.5, axis=1) < 0.2 - 0.03 * lv, np.zeros(f.nelem, bool))

    def flags(g):
        cen = g.lower + 0.5 * g.delta
        alpha = np.where((g.level == 6) & (np.linalg.norm(cen - 0.5, axis=1) < 0.1), 0.3, 0.0)
        from nemo_amd import amr_run
        return amr_run.adapt_flags(g, alpha)

    for _ in range(4):
        g, u = f.adapt(*flags(f))
        if u is None:
            break
        f = g
    r, c = flags(f)
    assert c.sum() > 0
    assert f.adapt(r, c, prune=True)[1] is None and f.adapt(r, c, prune=False)[1] is None


def test_face_neighbour_flags_by_geometry():
    """amr_run.face_neighbours_any marks a leaf when it or any leaf sharing
    a face with it is flagged; checked against touching boxes found by
    brute force (periodic), on a forest with many 2:1 faces."""
    from nemo_amd import amr_run
    rng = np.random.default_rng(4)
    f = amr.Forest.uniform(2, minlevel=2, maxlevel=4)
    for _ in range(2):
        f = f.refine(rng.random(f.nelem) < 0.25).balance()
    lo = f.anchor.astype(np.int64)
    h = (amr.ROOT >> f.level).astype(np.int64)
    n = f.nelem
    touch = np.zeros((n, n), dtype=bool)
    for d in range(3):
        for s in (-1, 1):
            # leaves whose face plane coincides and whose face rectangles overlap
            plane_i = (lo[:, d] + (h if s > 0 else 0)) % amr.ROOT
            plane_j = (lo[:, d] + (0 if s > 0 else h)) % amr.ROOT
            same = plane_i[:, None] == plane_j[None, :]
            ov = np.ones((n, n), dtype=bool)
            for t in range(3):
                if t != d:
                    ov &= (lo[:, None, t] < lo[None, :, t] + h[None, :]) & (lo[None, :, t] < lo[:, None, t] + h[:, None])
            touch |= same & ov
    np.fill_diagonal(touch, False)
    for _ in range(20):
        flag = rng.random(n) < rng.random() * 0.3
        want = flag | (touch & flag[None, :]).any(axis=1)
        assert np.array_equal(amr_run.face_neighbours_any(f, flag), want)


def _touching(f):
    """[n, n] leaves sharing a face (periodic), by brute force on the boxes."""
    lo = f.anchor.astype(np.int64)
    h = (amr.ROOT >> f.level).astype(np.int64)
    n = f.nelem
    touch = np.zeros((n, n), dtype=bool)
    for d in range(3):
        for s in (-1, 1):
            plane_i = (lo[:, d] + (h if s > 0 else 0)) % amr.ROOT
            plane_j = (lo[:, d] + (0 if s > 0 else h)) % amr.ROOT
            ov = plane_i[:, None] == plane_j[None, :]
            for t in range(3):
                if t != d:
                    ov &= (lo[:, None, t] < lo[None, :, t] + h[None, :]) & (lo[None, :, t] < lo[:, None, t] + h[:, None])
            touch |= ov
    np.fill_diagonal(touch, False)
    return touch


def test_balance_matches_brute_force_rounds():
    """Forest.balance (rounds that look only from new leaves and from leaves
    whose check failed) equals refining, round by round until none is left,
    every leaf that shares a face with a leaf two or more levels finer,
    found by brute force on the boxes; inputs refined several times without
    balancing (jumps of up to four levels)."""
    rng = np.random.default_rng(12)
    for trial in range(12):
        f = amr.Forest.uniform(1, minlevel=1, maxlevel=5)
        for r in range(3):
            flag = (rng.random(f.nelem) < 0.2) & (f.level < 5)
            if r:
                flag |= (f.level == f.level.max()) & (rng.random(f.nelem) < 0.5) & (f.level < 5)
            f = f.refine(flag)
        g = f
        while True:
            t = _touching(g)
            need = (t & (g.level[None, :] > g.level[:, None] + 1)).any(axis=1)
            if not need.any():
                break
            g = g.refine(need)
        b = f.balance()
        assert np.array_equal(b.level, g.level) and np.array_equal(b.anchor, g.anchor), trial


def test_morton_code_moves():
    """Moves on the Morton code (amr._key_move, amr._key_floor) equal the
    Morton codes of the moved or rounded anchors, periodic wrap included
    (anchors on the box's first and last cells of each level)."""
    rng = np.random.default_rng(11)
    n = 4000
    level = rng.integers(0, amr.GRID_BITS + 1, n)
    h = amr._size(level)
    cells = np.left_shift(np.int64(1), level.astype(np.int64))
    anchor = rng.integers(0, cells[:, None], (n, 3)) * h[:, None]
    anchor[:8] = 0
    anchor[8:16] = ((cells[8:16] - 1) * h[8:16])[:, None]
    key = amr._mkeys(anchor)
    for d in range(3):
        for s in (-1, 1):
            moved = anchor.copy()
            moved[:, d] = (moved[:, d] + s * h) % amr.ROOT
            assert np.array_equal(amr._key_move(key, d, amr._dilated(h, d), s), amr._mkeys(moved))
    big = np.minimum(2 * h, amr.ROOT)
    assert np.array_equal(amr._key_floor(key, big), amr._mkeys(anchor & ~(big - 1)[:, None]))


class _OracleSolver:
    """The calls amr_run.AdaptiveRun makes of a solver, answered by the
    oracle (indicator, CPU steps with the forest's mortars, state transfer)."""
    M3 = 64

    def __init__(self, cfg, forest, u):
        self.cfg, self.forest, self.u = cfg, forest, np.ascontiguousarray(u)

    def indicator(self):
        return O.indicator(self.cfg, self.u)[0]

    def run(self, k, record_dt=False):
        rc, h = O.run(O.with_mortars(self.cfg, self.forest.mortars), self.forest.nbr, self.forest.delta, self.u, k)
        assert rc == 0
        return h

    def remesh(self, new, units):
        u, nfb = O.remap(3, self.cfg.gamma, self.u, units, new.nelem)
        self.u, self.forest = np.ascontiguousarray(u), new
        return nfb


def test_adaptive_run_steady_shortcuts_are_exact():
    """AdaptiveRun skips a pass whose indicator masks, or whose refine and
    coarsen flags (as Forest.adapt reads them), equal those of the last pass
    that left the same forest unchanged.  A Sedov blast on levels 2-4
    (oracle steps, adaptation every 8 of 64 steps) gives the same meshes,
    replacement units, time steps and final state with the shortcuts as
    with every pass run in full, and the flag test decides passes whose
    masks moved."""
    from nemo_amd import amr_run
    cfg = O.make_cfg(N=3, volume=O.VOL_EC, surface=O.SURF_HLLC, indicator=O.IND_PP, timedisc=O.TD_LSRK54,
                     repair=1, cfl=0.5)
    ic = lambda f: indicator_mlp.sedov(f, 3, seed=1234, radius=0.08, p0=1e-5)
    f = amr.Forest.uniform(2, minlevel=2, maxlevel=4)
    for _ in range(8):
        a = O.indicator(cfg, np.ascontiguousarray(ic(f)))[0]
        new, units = f.adapt(amr_run.face_neighbours_any(f, a > 0), np.zeros(f.nelem, bool))
        if units is None:
            break
        f = new
    runs = {}
    for sc in (True, False):
        s = _OracleSolver(cfg, f, ic(f))
        run = amr_run.AdaptiveRun(s, f, every=8, steady_shortcuts=sc)
        run.advance(64, record=True)
        runs[sc] = (run, s.u)
    (a, ua), (b, ub) = runs[True], runs[False]
    assert a.flag_shortcuts > 0
    assert b.flag_shortcuts == b.mask_shortcuts == 0
    assert a.remeshes == b.remeshes >= 1
    assert len(a.history) == len(b.history) == 8
    for x, y in zip(a.history, b.history):
        assert np.array_equal(x.forest.level, y.forest.level) and np.array_equal(x.forest.anchor, y.forest.anchor)
        assert (x.units is None) == (y.units is None)
        if x.units is not None:
            assert np.array_equal(x.units, y.units)
        assert np.array_equal(x.dts, y.dts)
    assert np.array_equal(ua, ub)


def test_adapt_refine_coarsen_inverse():
    f = _forest(lmax=5)
    flags = np.zeros(f.nelem, dtype=bool)
    flags[f.level == f.level.max()] = True
    g, units = f.adapt(flags, np.zeros(f.nelem, bool))
    # coarsen exactly the families the refinement (and its balancing) made
    c = np.zeros(g.nelem, bool)
    for kind, io, jn in units:
        if kind == amr.UNIT_REFINE:
            c[jn:jn + 8] = True
    assert c.sum() > 8 * flags.sum()  # balancing refined neighbours too
    h, units2 = g.adapt(np.zeros(g.nelem, bool), c)
    assert np.array_equal(h.level, f.level) and np.array_equal(h.anchor, f.anchor)
    assert np.all(units2[:, 0] != amr.UNIT_REFINE)


@pytest.mark.parametrize("N", [2, 3, 4])
def test_mortar_ops_are_l2_projection(N):
    M = N + 1
    P, R = O.mortar_ops(N)
    x, w = O.lgl(N)

    def lag(xs, p):
        v = np.ones_like(xs)
        for q in range(M):
            if q != p:
                v = v * (xs - x[q]) / (x[p] - x[q])
        return v

    g, gw = np.polynomial.legendre.leggauss(M + 2)
    mass = np.array([[np.sum(gw * lag(g, p) * lag(g, q)) for q in range(M)] for p in range(M)])
    for c in (0, 1):
        # interpolation: P applied to samples of a degree-N polynomial
        coef = np.arange(1, M + 1) / 7.0
        poly = lambda t: np.polyval(coef, t)
        xs = 0.5 * (x + (2 * c - 1))
        assert np.abs(P[c] @ poly(x) - poly(xs)).max() < 1e-14
        xi = 0.5 * (g + (2 * c - 1))
        B = np.array([[0.5 * np.sum(gw * lag(xi, p) * lag(g, i)) for i in range(M)] for p in range(M)])
        assert np.abs(np.linalg.solve(mass, B) - R[c]).max() < 1e-13
        assert np.abs(w @ R[c] - 0.5 * w).max() < 1e-15
    assert np.abs(R[0] @ P[0] + R[1] @ P[1] - np.eye(M)).max() < 1e-13


def _state_at_rest(f, N):
    """rho, p polynomials of degree <= 2 at rest; exact RHS = (0, -grad p, 0)."""
    M = N + 1
    x, _ = O.lgl(N)
    X = f.node_coords(x)
    xx, yy, zz = X[:, 0] - 0.5, X[:, 1] - 0.5, X[:, 2] - 0.5
    p = 1.0 + 0.3 * xx ** 2 + 0.2 * yy * zz + 0.1 * xx
    u = np.zeros((f.nelem, 5, M ** 3))
    u[:, 0] = 1.0 + 0.1 * yy + 0.05 * xx * zz
    u[:, 4] = p / 0.4
    ex = np.zeros_like(u)
    ex[:, 1] = -(0.6 * xx + 0.1)
    ex[:, 2] = -0.2 * zz
    ex[:, 3] = -0.2 * yy
    return u, ex


@pytest.mark.parametrize("N", [2, 3, 4])
@pytest.mark.parametrize("vol,surf", [(O.VOL_CENTRAL, O.SURF_RUSANOV), (O.VOL_WEAK, O.SURF_HLLC)])
def test_oracle_mortar_exact_state(N, vol, surf):
    f = _forest()
    cfg = O.with_mortars(O.make_cfg(N=N, volume=vol, surface=surf), f.mortars)
    u, ex = _state_at_rest(f, N)
    r = O.rhs(cfg, f.nbr, f.delta, u)
    inner = np.all((f.lower > 1e-12) & (f.lower + f.delta < 1 - 1e-12), axis=1)  # periodic wrap breaks the polynomial
    on_mortar = np.zeros(f.nelem, bool)
    on_mortar[f.mortars[:, 0]] = True
    on_mortar[f.mortars[:, 2:].ravel()] = True
    assert (inner & on_mortar).sum() > 50
    scale = np.abs(ex).max() * N / f.delta.min()
    assert np.abs(r - ex)[inner].max() <= 1e-13 * scale


@pytest.mark.parametrize("N", [2, 3])
@pytest.mark.parametrize("vol,surf", [(O.VOL_CENTRAL, O.SURF_RUSANOV), (O.VOL_CENTRAL, O.SURF_HLLD)])
def test_oracle_mhd_mortar_exact_state(N, vol, surf):
    """GLM-MHD at rest in a uniform field (B constant, psi = 0) with the
    pressure and density of the Euler case: the field adds a constant to the
    momentum flux and nothing else, so the exact RHS is still (0, -grad p, 0,
    0, 0); the 9-variable mortar path must reproduce it with the same halves
    and orientation."""
    f = _forest()
    cfg = O.with_mortars(O.make_cfg(N=N, eq=O.EQ_MHD, gamma=1.4, mu0=1.0, volume=vol, surface=surf,
                                    timedisc=O.TD_LSRK54, glm_ch=0.7), f.mortars)
    ue, ex5 = _state_at_rest(f, N)
    B = np.array([0.3, -0.5, 0.4])
    u = np.zeros((f.nelem, 9, (N + 1) ** 3))
    u[:, :5] = ue
    u[:, 4] += 0.5 * (B @ B)
    u[:, 5:8] = B[None, :, None]
    ex = np.zeros_like(u)
    ex[:, :5] = ex5
    r = O.rhs(cfg, f.nbr, f.delta, u)
    inner = np.all((f.lower > 1e-12) & (f.lower + f.delta < 1 - 1e-12), axis=1)
    scale = np.abs(ex).max() * N / f.delta.min()
    assert np.abs(r - ex)[inner].max() <= 1e-13 * scale


@pytest.mark.parametrize("N", [2, 3, 4])
def test_oracle_mortar_free_stream_and_conservation(N):
    f = _forest()
    M = N + 1
    cfg = O.with_mortars(O.make_cfg(N=N, volume=O.VOL_EC, surface=O.SURF_HLLC, indicator=O.IND_PP), f.mortars)
    u = np.zeros((f.nelem, 5, M ** 3))
    u[:, 0], u[:, 1], u[:, 2], u[:, 3], u[:, 4] = 1.3, 0.3, -0.2, 0.1, 2.5
    r = O.rhs(cfg, f.nbr, f.delta, u)
    assert np.abs(r).max() < 1e-12 * M / f.delta.min()
    u = indicator_mlp.sedov(f, N, radius=0.05, p0=1e-2)
    a, _ = O.indicator(cfg, u)
    assert (a > 0).any()
    r = O.rhs(cfg, f.nbr, f.delta, u)
    _, w = O.lgl(N)
    w3 = (w[:, None, None] * w[None, :, None] * w[None, None, :]).reshape(-1)
    tot = np.einsum("evn,n,e->v", r, w3, np.prod(f.delta, axis=1) / 8.0)
    assert np.abs(tot).max() <= 1e-13 * np.abs(r).max()


def _poly_state(f, N, shift=0.0):
    M = N + 1
    x, _ = O.lgl(N)
    X = f.node_coords(x)
    u = np.zeros((f.nelem, 5, M ** 3))
    xx, yy, zz = X[:, 0], X[:, 1], X[:, 2]
    u[:, 0] = 2.0 + 0.3 * xx - 0.2 * yy * zz + shift
    u[:, 1] = 0.1 * xx * yy
    u[:, 2] = -0.2 * zz
    u[:, 3] = 0.05
    u[:, 4] = 3.0 + 0.1 * xx * xx * zz
    return u


@pytest.mark.parametrize("N", [2, 3, 4])
def test_oracle_remap_exact_and_conservative(N):
    f = _forest(lmax=5)
    flags = np.zeros(f.nelem, bool)
    flags[::3] = True
    g, units = f.adapt(flags, np.zeros(f.nelem, bool))
    # refinement interpolates polynomials of degree <= N exactly
    u = _poly_state(f, N)
    ug, nfb = O.remap(N, 1.4, u, units, g.nelem)
    assert nfb == 0
    assert np.abs(ug - _poly_state(g, N)).max() < 1e-13
    # coarsening back is the inverse of refinement (L2 projection)
    c = np.zeros(g.nelem, bool)
    for kind, io, jn in units:
        if kind == amr.UNIT_REFINE:
            c[jn:jn + 8] = True
    h, units2 = g.adapt(np.zeros(g.nelem, bool), c)
    assert np.array_equal(h.anchor, f.anchor)
    rng = np.random.default_rng(1)
    v = u * (1.0 + 0.01 * rng.random(u.shape))
    vg, _ = O.remap(N, 1.4, v, units, g.nelem)
    vh, _ = O.remap(N, 1.4, vg, units2, h.nelem)
    assert np.abs(vh - v).max() < 1e-13
    # a coarsening of non-polynomial data conserves every variable
    wg = vg * (1.0 + 0.01 * rng.random(vg.shape))
    wh, _ = O.remap(N, 1.4, wg, units2, h.nelem)
    _, w = O.lgl(N)
    w3 = (w[:, None, None] * w[None, :, None] * w[None, None, :]).reshape(-1)
    assert np.allclose(np.einsum("evn,n,e->v", wh, w3, np.prod(h.delta, 1)),
                       np.einsum("evn,n,e->v", wg, w3, np.prod(g.delta, 1)), rtol=1e-14, atol=0)


def test_oracle_remap_positivity_fallback():
    N = 3
    f = amr.Forest.uniform(1, maxlevel=2)
    M = N + 1
    x, w = O.lgl(N)
    u = np.zeros((f.nelem, 5, M ** 3))
    u[:, 0] = 1.0
    # nodal pressure (1e-3, 1e-8, 1e-8, 1e-3) along x: positive at the nodes,
    # its cubic interpolant is -2.5e-4 at the element centre, a child node
    i = np.arange(M ** 3) % M
    u[:, 4] = np.where((i == 0) | (i == M - 1), 1e-3, 1e-8)[None, :] / 0.4
    assert all(O.isvalid(O.make_cfg(), u[0, :, n]) for n in range(M ** 3))
    flags = np.ones(f.nelem, bool)
    g, units = f.adapt(flags, np.zeros(f.nelem, bool))
    ug, nfb = O.remap(N, 1.4, u, units, g.nelem)
    assert nfb > 0
    w3 = (w[:, None, None] * w[None, :, None] * w[None, None, :]).reshape(-1)
    assert np.allclose(np.einsum("evn,n,e->v", ug, w3, np.prod(g.delta, 1)),
                       np.einsum("evn,n,e->v", u, w3, np.prod(f.delta, 1)), rtol=1e-14, atol=1e-18)
    for e in range(g.nelem):
        rho = ug[e, 0]
        p = 0.4 * (ug[e, 4] - 0.5 * (ug[e, 1:4] ** 2).sum(0) / rho)
        assert (rho > 0).all() and (p > 0).all()


def test_oracle_remap_mhd_exact_and_magnetic_fallback():
    """GLM-MHD states (9 variables): refinement is exact for polynomials of
    degree <= N in every variable, and the fallback tests the gas pressure
    with the magnetic (and psi) energy taken out: a unit whose interpolated
    gas pressure turns negative under a strong uniform field falls back,
    although its total energy, read as an Euler state, stays positive."""
    N = 3
    M = N + 1
    f = amr.Forest.uniform(1, maxlevel=2)
    x, w = O.lgl(N)
    X = f.node_coords(x)
    xx, yy, zz = X[:, 0], X[:, 1], X[:, 2]
    u = np.zeros((f.nelem, 9, M ** 3))
    u[:, :5] = _poly_state(f, N)
    u[:, 4] += 5.0
    u[:, 5], u[:, 6], u[:, 7], u[:, 8] = 0.5 + 0.1 * xx * yy, -0.3 * zz ** 3, 0.2, 0.01 * yy
    g, units = f.adapt(np.ones(f.nelem, bool), np.zeros(f.nelem, bool))
    ug, nfb = O.remap(N, 5.0 / 3.0, u, units, g.nelem, mu0=1.0)
    assert nfb == 0
    Xg = g.node_coords(x)
    assert np.abs(ug[:, :5] - (_poly_state(g, N) + np.array([0, 0, 0, 0, 5.0])[None, :, None])).max() < 1e-13
    assert np.abs(ug[:, 5] - (0.5 + 0.1 * Xg[:, 0] * Xg[:, 1])).max() < 1e-13
    assert np.abs(ug[:, 6] + 0.3 * Xg[:, 2] ** 3).max() < 1e-13
    # gas pressure (1e-3, 1e-8, 1e-8, 1e-3) along x under B = (1, 0, 0): the
    # interpolant is negative at a child node, the total energy is not
    i = np.arange(M ** 3) % M
    u = np.zeros((f.nelem, 9, M ** 3))
    u[:, 0] = 1.0
    u[:, 5] = 1.0
    u[:, 4] = np.where((i == 0) | (i == M - 1), 1e-3, 1e-8)[None, :] * 1.5 + 0.5
    ug, nfb = O.remap(N, 5.0 / 3.0, u, units, g.nelem, mu0=1.0)
    assert nfb == f.nelem
    _, nfb_euler = O.remap(N, 5.0 / 3.0, u[:, :5], units, g.nelem)
    assert nfb_euler == 0
    w3 = (w[:, None, None] * w[None, :, None] * w[None, None, :]).reshape(-1)
    assert np.allclose(np.einsum("evn,n,e->v", ug, w3, np.prod(g.delta, 1)),
                       np.einsum("evn,n,e->v", u, w3, np.prod(f.delta, 1)), rtol=1e-14, atol=1e-18)


# ---- the same faces on a non-conforming and a conforming mesh -----------------------

def _slab_meshes(level=2):
    """(coarse, mixed, fine): uniform forests at `level` and `level + 1`, and
    the coarse one with the slab of elements at x-index 1 refined, so that
    every non-conforming face has its normal along x."""
    coarse = amr.Forest.uniform(level, maxlevel=level + 1)
    h = 1.0 / (1 << level)
    cen = coarse.lower + 0.5 * coarse.delta
    mixed = coarse.refine(np.abs(cen[:, 0] - 1.5 * h) < 0.25 * h).balance()
    fine = amr.Forest.uniform(level + 1, maxlevel=level + 1)
    return coarse, mixed, fine, h


def _slab_state(f, N, h, eq, seed=3):
    """A state that varies along x only, as a polynomial of degree N within
    each coarse cell [i h, (i+1) h) and discontinuous between cells: every
    element of every mesh holds it exactly, so the faces between elements of
    one coarse cell are continuous and the faces between cells carry the same
    Riemann problem on every mesh."""
    rng = np.random.default_rng(seed)
    M = N + 1
    ncell = int(round(1.0 / h))
    nv = 9 if eq == O.EQ_MHD else 5
    # per cell and variable: polynomial coefficients in the local coordinate
    coef = 0.1 * rng.uniform(-1, 1, (ncell, nv, N + 1))
    base = np.array([1.2, 0.3, -0.2, 0.1, 6.0, 0.5, -0.4, 0.6, 0.1][:nv])
    x, _ = O.lgl(N)
    X = f.node_coords(x)
    cell = np.floor((f.lower[:, 0] + 0.5 * f.delta[:, 0]) / h).astype(int)
    xi = (X[:, 0] - cell[:, None] * h) / h
    u = np.empty((f.nelem, nv, M ** 3))
    for v in range(nv):
        u[:, v] = base[v] + sum(coef[cell, v, k][:, None] * xi ** k for k in range(N + 1))
    return u


def _by_position(f):
    return {tuple(np.round(lo * 1024).astype(int)) + (int(lv),): e
            for e, (lo, lv) in enumerate(zip(f.lower, f.level))}


def _slab_cfgs(N):
    yield O.make_cfg(N=N, volume=O.VOL_EC, surface=O.SURF_HLLC, indicator=O.IND_PP)
    yield O.make_cfg(N=N, volume=O.VOL_WEAK, surface=O.SURF_RUSANOV, indicator=O.IND_PP)
    mhd = dict(N=N, eq=O.EQ_MHD, gamma=5.0 / 3.0, mu0=1.0, indicator=O.IND_PP, timedisc=O.TD_LSRK54,
               glm_ch=0.8, glm_alpha=0.5)
    yield O.make_cfg(volume=O.VOL_EC, surface=O.SURF_HLLD, **mhd)
    yield O.make_cfg(volume=O.VOL_CENTRAL, surface=O.SURF_RUSANOV, **mhd)


@pytest.mark.parametrize("N", [2, 3, 4])
def test_oracle_mortar_matches_conforming_faces(N):
    """With a state constant across x-normal faces (tangentially) the mortar
    sequence has a closed form: interpolating the big face is exact, every
    half face carries the conforming face's Riemann problem, and the
    projection of a tangentially constant flux is that flux.  So on the slab
    mesh the RHS of every big element equals its RHS on the coarse uniform
    mesh and that of every small element its RHS on the fine uniform mesh, to
    rounding, for any blending factor and, for GLM-MHD, with the
    non-conservative term of each side (the small side's own node, the big
    side's interpolated one).  A wrong side state in that term, a term added
    twice or dropped, or a transposed projection fails it."""
    coarse, mixed, fine, h = _slab_meshes()
    assert len(mixed.mortars) and np.all(mixed.mortars[:, 1] // 2 == 0)
    pc, pf = _by_position(coarse), _by_position(fine)
    on_mortar = np.zeros(mixed.nelem, bool)
    on_mortar[mixed.mortars[:, 0]] = True
    on_mortar[mixed.mortars[:, 2:].ravel()] = True
    for cfg in _slab_cfgs(N):
        eq = cfg.eq
        um, uc, uf = (_slab_state(f, N, h, eq) for f in (mixed, coarse, fine))
        am = np.full(mixed.nelem, 0.3)
        rm = O.rhs(O.with_mortars(cfg, mixed.mortars), mixed.nbr, mixed.delta, um, am)
        rc = O.rhs(cfg, coarse.nbr, coarse.delta, uc, np.full(coarse.nelem, 0.3))
        rf = O.rhs(cfg, fine.nbr, fine.delta, uf, np.full(fine.nelem, 0.3))
        scale = max(np.abs(rc).max(), np.abs(rf).max())
        assert scale > 1.0
        worst = 0.0
        for e, key in enumerate(_by_position(mixed)):
            ref = rc[pc[key]] if key in pc else rf[pf[key]]
            assert np.array_equal(um[e], (uc[pc[key]] if key in pc else uf[pf[key]]))
            worst = max(worst, np.abs(rm[e] - ref).max())
        assert worst <= 1e-12 * scale, (eq, cfg.volume, cfg.surface, worst / scale)
        # the mortar faces carry a real jump: a conforming pairing would differ
        assert np.abs(rm[on_mortar]).max() > 1e-3 * scale


@pytest.mark.parametrize("N", [2, 3])
def test_oracle_mhd_mortar_free_stream_and_mass(N):
    """GLM-MHD on the refined-around-a-point forest: a constant state with
    psi = 0 is steady, and the mass of a rough state is conserved across the
    non-conforming faces (the Powell/GLM terms have no mass component)."""
    f = _forest()
    M3 = (N + 1) ** 3
    cfg = O.with_mortars(O.make_cfg(N=N, eq=O.EQ_MHD, gamma=5.0 / 3.0, mu0=1.0, volume=O.VOL_EC,
                                    surface=O.SURF_HLLD, indicator=O.IND_PP, timedisc=O.TD_LSRK54,
                                    glm_ch=0.8, glm_alpha=0.0), f.mortars)
    u0 = np.array([1.2, 0.36, -0.24, 0.12, 3.2, 0.4, 0.7, -0.5, 0.0])
    u = np.ascontiguousarray(np.broadcast_to(u0[None, :, None], (f.nelem, 9, M3)))
    r = O.rhs(cfg, f.nbr, f.delta, u, np.full(f.nelem, 0.4))
    assert np.abs(r).max() < 1e-11 * (N + 1) / f.delta.min()
    rng = np.random.default_rng(5)
    u = u * (1.0 + 0.1 * rng.uniform(-1, 1, u.shape))
    r = O.rhs(cfg, f.nbr, f.delta, u, rng.uniform(0, 0.5, f.nelem))
    _, w = O.lgl(N)
    w3 = (w[:, None, None] * w[None, :, None] * w[None, None, :]).reshape(-1)
    W = w3[None, :] * (np.prod(f.delta, axis=1) / 8.0)[:, None]
    assert abs((W * r[:, 0]).sum()) <= 1e-13 * (W * np.abs(r[:, 0])).sum()


# Plausible mistakes in the mortar sequence, planted in a copy of the
# oracle: each must make the slab comparison above fail.  (Its state is
# constant along the faces, so a wrong half or orientation goes unseen
# there; the exactly representable state at rest above catches those.)
_PLANTED_MORTAR = [
    ("big side takes the small side's term", "phys_nc(c, d, ub, uL, uR, nb);", "phys_nc(c, d, us, uL, uR, nb);"),
    ("small side's term dropped", "flux[((5 * m + 1 + k) * V + h) * M2 + ta + M * tb] = f[h] + ns[h];",
     "flux[((5 * m + 1 + k) * V + h) * M2 + ta + M * tb] = f[h];"),
    ("big side's term after the projection", "fk[k][h][ta + M * tb] = f[h] + nb[h];", "fk[k][h][ta + M * tb] = f[h];"),
    ("term added again at the element", "if (em >= 0) {\n                        phys_nc(c, d, ul[0], unm",
     "if (1) {\n                        if (em < 0) for (int h = 0; h < V; ++h) unm[h] = ul[0][h];\n"
     "                        phys_nc(c, d, ul[0], unm"),
    ("projection transposed", "r += R[(ka * M + p) * M + ta] * fk[k][h][ta + M * tb];",
     "r += R[(ka * M + ta) * M + p] * fk[k][h][ta + M * tb];"),
]


@pytest.mark.parametrize("name,old,new", _PLANTED_MORTAR, ids=[p[0] for p in _PLANTED_MORTAR])
def test_planted_mortar_mistakes_fail_the_slab_comparison(tmp_path, name, old, new):
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
    L.orc_rhs.argtypes = [ctypes.POINTER(O.Cfg), ctypes.c_int64, PI, P, P, P, P]

    def rhs(cfg, f, u, alpha):
        nb = np.ascontiguousarray(f.nbr, dtype=np.int32)
        de = np.ascontiguousarray(f.delta, dtype=np.float64)
        u = np.ascontiguousarray(u)
        a = np.ascontiguousarray(alpha)
        out = np.zeros_like(u)
        L.orc_rhs(ctypes.byref(cfg), u.shape[0], nb.ctypes.data_as(PI), de.ctypes.data_as(P), u.ctypes.data_as(P),
                  a.ctypes.data_as(P), out.ctypes.data_as(P))
        return out

    N = 3
    coarse, mixed, fine, h = _slab_meshes()
    pc, pf = _by_position(coarse), _by_position(fine)
    cfg = list(_slab_cfgs(N))[2]  # GLM-MHD, EC volume, HLLD faces
    um, uc, uf = (_slab_state(f, N, h, cfg.eq) for f in (mixed, coarse, fine))
    rm = rhs(O.with_mortars(cfg, mixed.mortars), mixed, um, np.full(mixed.nelem, 0.3))
    rc = O.rhs(cfg, coarse.nbr, coarse.delta, uc, np.full(coarse.nelem, 0.3))
    rf = O.rhs(cfg, fine.nbr, fine.delta, uf, np.full(fine.nelem, 0.3))
    scale = max(np.abs(rc).max(), np.abs(rf).max())
    worst = max(np.abs(rm[e] - (rc[pc[k]] if k in pc else rf[pf[k]])).max()
                for e, k in enumerate(_by_position(mixed)))
    assert worst > 1e-6 * scale, (name, worst / scale)
