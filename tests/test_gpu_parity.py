"""Parity of the HIP path (through the C-ABI) with the FP64 CPU oracle.

Bars (from the task statement of the plain FP64 path):
  * RHS per call:        max|rhs_gpu - rhs_orc| / max|rhs_orc| <= 1e-12
  * after 100 RK steps:  max|u_gpu - u_orc|     / max|u_orc|   <= 1e-10
Both runs start from the same host-generated initial data on the same mesh.
"""
import numpy as np
import pytest

from nemo_amd import _lib, mesh as meshmod, setups, solver
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-12
RUN_TOL = 1e-10


def _cfgs(N, vol, surf, ind, td, repair=0, cfl=0.5):
    g = solver.make_config(N=N, volume=vol, surface=surf, indicator=ind, timedisc=td, repair=repair, cfl=cfl)
    o = O.make_cfg(N=N, volume=vol, surface=surf, indicator=ind, timedisc=td, repair=repair, cfl=cfl)
    return g, o


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


COMBOS = [
    # (N, volume, surface, indicator, timedisc, setup)
    (3, _lib.VOL_WEAK, _lib.SURF_RUSANOV, _lib.IND_NONE, _lib.TD_SSPRK54, "vortex"),   # isentropic vortex scheme
    (3, _lib.VOL_EC, _lib.SURF_HLLC, _lib.IND_PP, _lib.TD_LSRK54, "kh"),               # Kelvin-Helmholtz scheme
    (3, _lib.VOL_CENTRAL, _lib.SURF_RUSANOV, _lib.IND_NONE, _lib.TD_LSRK54, "vortex"),
    (3, _lib.VOL_EC, _lib.SURF_RUSANOV, _lib.IND_NONE, _lib.TD_SSPRK54, "kh"),
    (3, _lib.VOL_WEAK, _lib.SURF_HLLC, _lib.IND_PP, _lib.TD_LSRK54, "kh"),
    (2, _lib.VOL_EC, _lib.SURF_HLLC, _lib.IND_PP, _lib.TD_LSRK54, "kh"),
    (4, _lib.VOL_EC, _lib.SURF_HLLC, _lib.IND_PP, _lib.TD_LSRK54, "kh"),
    (2, _lib.VOL_WEAK, _lib.SURF_RUSANOV, _lib.IND_NONE, _lib.TD_SSPRK54, "vortex"),
    (4, _lib.VOL_WEAK, _lib.SURF_RUSANOV, _lib.IND_NONE, _lib.TD_SSPRK54, "vortex"),
    # smooth data with many log-mean pairs just past the series switch
    (3, _lib.VOL_EC, _lib.SURF_HLLC, _lib.IND_PP, _lib.TD_LSRK54, "vortex"),
    (4, _lib.VOL_EC, _lib.SURF_RUSANOV, _lib.IND_NONE, _lib.TD_LSRK54, "vortex"),
]


def _setup(name, n, N):
    m = meshmod.uniform_periodic(n)
    if name == "vortex":
        return m, setups.isentropic_vortex(m, N)
    return m, setups.kelvin_helmholtz(m, N, seed=1234)


@pytest.mark.parametrize("combo", COMBOS, ids=lambda c: f"N{c[0]}-vol{c[1]}-surf{c[2]}-ind{c[3]}-td{c[4]}-{c[5]}")
def test_rhs_parity(combo):
    N, vol, surf, ind, td, name = combo
    m, u0 = _setup(name, 8, N)
    g, o = _cfgs(N, vol, surf, ind, td)
    print(f"rhs parity {combo}", flush=True)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        r = s.kernel_calc_flux()
        a = s.indicator()
    aref, _ = O.indicator(o, u0) if ind == _lib.IND_PP else (np.zeros(m.nelem), None)
    rref = O.rhs(o, m.nbr, m.delta, u0, aref if ind == _lib.IND_PP else None)
    assert np.array_equal(a > 0, aref > 0)
    assert np.abs(a - aref).max() <= 1e-12
    if ind == _lib.IND_PP:
        assert (aref > 0).any(), "blending never active: the FV path is not exercised"
    err = _rel(r, rref)
    print(f"  rel Linf {err:.3e}")
    assert err <= RHS_TOL


def test_timestep_parity():
    m, u0 = _setup("kh", 8, 3)
    g, o = _cfgs(3, _lib.VOL_EC, _lib.SURF_HLLC, _lib.IND_PP, _lib.TD_LSRK54)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        dt, lm, vm = s.hook_calc_timestep()
    dto, lmo, vmo = O.calc_timestep(o, m.delta, u0)
    assert abs(dt - dto) <= 1e-15 * dto
    assert abs(lm - lmo) <= 1e-15 * lmo
    assert abs(vm - vmo) <= 1e-15 * vmo


@pytest.mark.parametrize("which", ["c1", "c2", "ec-ssp"])
def test_run_100_steps(which):
    """Configs 1 and 2, and the two-point volume form under SSPRK(5,4) (the
    FP64 kernel's unfused path: with LSRK its last volume pass applies the
    update itself)."""
    N, vol, surf, ind, td, name = {"c1": COMBOS[0], "c2": COMBOS[1], "ec-ssp": COMBOS[3]}[which]
    m, u0 = _setup(name, 8, N)
    g, o = _cfgs(N, vol, surf, ind, td, repair=1)
    print(f"100-step parity {which}", flush=True)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        hist = s.run(100, record_dt=True)
        u = s.get_state()
    uref = u0.copy()
    rc, href = O.run(o, m.nbr, m.delta, uref, 100)
    assert rc == 0
    assert np.abs(hist - href).max() <= 1e-12 * href.max()
    err = _rel(u, uref)
    print(f"  rel Linf after 100 steps {err:.3e}")
    assert err <= RUN_TOL


def test_stage_api_matches_run():
    """nemo_gpu_step with a host dt equals nemo_gpu_run's device-side dt path."""
    N, vol, surf, ind, td, name = COMBOS[1]
    m, u0 = _setup(name, 4, N)
    g, _ = _cfgs(N, vol, surf, ind, td, repair=1)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        h = s.run(3, record_dt=True)
        ua = s.get_state()
        s.set_state(u0)
        for dt in h:
            s.step(float(dt))
        s.synchronize()
        ub = s.get_state()
    assert np.array_equal(ua, ub)


def _rough_state(m, N, seed=3):
    """Random primitive state with pressure near zero (repair territory)."""
    rng = np.random.default_rng(seed)
    M3 = (N + 1) ** 3
    rho = rng.uniform(0.5, 2.0, (m.nelem, M3))
    v = rng.uniform(-1.0, 1.0, (3, m.nelem, M3))
    p = rng.uniform(1e-3, 2e-3, (m.nelem, M3))
    return np.ascontiguousarray(setups.prim2cons(rho, v[0], v[1], v[2], p, 1.4))


def _pressure(u, gamma=1.4):
    return (gamma - 1.0) * (u[:, 4] - 0.5 * (u[:, 1] ** 2 + u[:, 2] ** 2 + u[:, 3] ** 2) / u[:, 0])


def test_repair_parity():
    """Stage 1 with 4x the CFL step leaves nodes with p <= 0; the device
    repair (tile squeeze) must reproduce the oracle's repaired state."""
    N = 3
    m = meshmod.uniform_periodic(4)
    u0 = _rough_state(m, N)
    g, o = _cfgs(N, _lib.VOL_EC, _lib.SURF_HLLC, _lib.IND_PP, _lib.TD_LSRK54, repair=1)
    dt = 4.0 * O.calc_timestep(o, m.delta, u0)[0]
    # the unrepaired stage really is invalid somewhere
    o_norep = O.make_cfg(N=N, volume=O.VOL_EC, surface=O.SURF_HLLC, indicator=O.IND_PP, timedisc=O.TD_LSRK54)
    ua = u0.copy()
    O.stage(o_norep, m.nbr, m.delta, ua, np.zeros_like(ua), np.zeros_like(ua), 1, dt)
    assert ((ua[:, 0] <= 0) | (_pressure(ua) <= 0)).sum() > 0
    uref = u0.copy()
    assert O.stage(o, m.nbr, m.delta, uref, np.zeros_like(uref), np.zeros_like(uref), 1, dt) == 0
    print("repair parity", flush=True)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        s._call("nemo_gpu_stage", s._h, 1, dt)
        s.synchronize()
        u = s.get_state()
    assert ((u[:, 0] > 0) & (_pressure(u) > 0)).all()
    err = _rel(u, uref)
    print(f"  rel Linf {err:.3e}")
    assert err <= 1e-10


@pytest.mark.parametrize("N", [2, 3, 4])
def test_partial_last_block(N):
    """125 elements (5^3) leave a partial last workgroup tile for every
    degree (7, 4 and 2 elements per workgroup): RHS, indicator, CFL and a
    20-step run still match the oracle."""
    m, u0 = _setup("kh", 5, N)
    g, o = _cfgs(N, _lib.VOL_EC, _lib.SURF_HLLC, _lib.IND_PP, _lib.TD_LSRK54, repair=1)
    print(f"partial block N={N}", flush=True)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        r = s.kernel_calc_flux()
        a = s.indicator()
        dt, _, _ = s.hook_calc_timestep()
        h = s.run(20, record_dt=True)
        u = s.get_state()
    aref, _ = O.indicator(o, u0)
    assert np.abs(a - aref).max() <= 1e-12
    assert _rel(r, O.rhs(o, m.nbr, m.delta, u0, aref)) <= RHS_TOL
    assert abs(dt - O.calc_timestep(o, m.delta, u0)[0]) <= 1e-15 * dt
    uref = u0.copy()
    rc, href = O.run(o, m.nbr, m.delta, uref, 20)
    assert rc == 0
    assert np.abs(h - href).max() <= 1e-12 * href.max()
    assert _rel(u, uref) <= RUN_TOL


# elements per workgroup of the stage kernel by degree (tiles of 7 x 9,
# 4 x 16 and 2 x 25 lanes)
TILE_BE = {2: 7, 3: 4, 4: 2}


def _near1_edge_state(m, N, F=1.085):
    """Blocks of the stage kernel's tile (N = 3: four consecutive Morton
    elements, a 2x2x1 patch; N = 2: seven; N = 4: two) whose density and
    pressure vary as f^g and f^-g around the block's first node, g = sin(2 pi
    u) in the block's x coordinate u, with f set per block so that the
    extreme node sits 0.1 % inside or 0.1 % outside the stage kernel's near-1
    bound (every node within a factor F of the first node's density and
    pressure): both variants of the direction passes run, the near-1 one at
    beta ratios up to F^4.  Returns the state and the host's restatement of
    the block test (a last partial block's padding slots hold element 0's
    first node, density and pressure 1, as its own first node)."""
    from nemo_amd import operators
    BE = TILE_BE[N]
    xi, _ = operators.lgl(N)
    X = m.node_coords(xi)                                  # [nelem, 3, M^3]
    nb = -(-m.nelem // BE)
    rho = np.empty((m.nelem, X.shape[2]))
    p = np.empty_like(rho)
    for b in range(nb):
        els = np.arange(b * BE, min((b + 1) * BE, m.nelem))
        x = X[els, 0, :]
        u = (x - X[els[0], 0, 0]) / max(np.ptp(x), 1e-300)
        g = np.sin(2 * np.pi * u)                          # 0 at the block's first node
        gmax = max(np.abs(g).max(), 1e-3)
        lnf = np.log(F) / gmax * (0.999 if b % 2 == 0 else 1.001)
        rho[els] = np.exp(lnf * g)
        p[els] = np.exp(-lnf * g)
    vx, vy, vz = 0.3 + 0 * rho, -0.2 + 0 * rho, 0.1 + 0 * rho
    U = setups.prim2cons(rho, vx, vy, vz, p, 1.4)           # [nelem, 5, M^3]
    # the kernel's test, restated: pressures from the conserved variables
    pp = 0.4 * (U[:, 4] - 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2 + U[:, 3] ** 2) / U[:, 0])
    ok = np.empty(nb, dtype=bool)
    for b in range(nb):
        els = np.arange(b * BE, min((b + 1) * BE, m.nelem))
        r, q = U[els, 0].ravel(), pp[els].ravel()
        rr, pr = r[0], q[0]
        ok[b] = ((r <= F * rr) & (rr <= F * r) & (q <= F * pr) & (pr <= F * q)).all()
    return np.ascontiguousarray(U), ok


@pytest.mark.parametrize("N", [2, 3, 4])
def test_near1_log_means_at_their_bound(N):
    """The FP64 two-point kernel's two variants of the log-means (with and
    without the log's range reduction, chosen per block) against the oracle
    on blocks built to sit just inside and just outside the choice's bound:
    RHS to 1e-12 and 20 LSRK steps to 1e-10 (at N = 2, 3 the near-1 variant
    fuses the update into its x pass and the general one leaves it to the
    separate update; at N = 4 both leave it to the separate update)."""
    m = meshmod.uniform_periodic(8)
    u0, ok = _near1_edge_state(m, N)
    assert ok.any() and not ok.all()
    g, o = _cfgs(N, _lib.VOL_EC, _lib.SURF_HLLC, _lib.IND_PP, _lib.TD_LSRK54, repair=1)
    print(f"near-1 bound, N = {N}: {int(ok.sum())} of {ok.size} blocks inside", flush=True)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        r = s.kernel_calc_flux()
        a = s.indicator()
        hist = s.run(20, record_dt=True)
        u = s.get_state()
    aref, _ = O.indicator(o, u0)
    rref = O.rhs(o, m.nbr, m.delta, u0, aref)
    err = _rel(r, rref)
    print(f"  rhs rel Linf {err:.3e}")
    assert np.abs(a - aref).max() <= 1e-12
    assert err <= RHS_TOL
    uref = u0.copy()
    rc, href = O.run(o, m.nbr, m.delta, uref, 20)
    assert rc == 0
    assert np.abs(hist - href).max() <= 1e-12 * href.max()
    err = _rel(u, uref)
    print(f"  rel Linf after 20 steps {err:.3e}")
    assert err <= RUN_TOL
