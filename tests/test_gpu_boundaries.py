"""Non-periodic boxes on the GPU (through the C-ABI) against the FP64
oracle: transmissive sides and Eulerian slip walls, read by the stage
kernels from the face slots a boundary pass fills before each stage.

Bars as for periodic meshes: RHS per call <= 1e-12 relative L-inf (per
variable for GLM-MHD), 20 RK steps <= 1e-10, FP32 operator RHS <= 5e-5."""
import numpy as np
import pytest

from nemo_amd import _lib, mesh as meshmod, setups, solver
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-12
RUN_TOL = 1e-10
GAMMA = 1.4

# walls across y, transmissive across x, periodic in z
KINDS = [_lib.BC_TRANSMISSIVE, _lib.BC_TRANSMISSIVE, _lib.BC_SLIP_WALL, _lib.BC_SLIP_WALL, None, None]


def _mesh():
    return meshmod.with_boundaries(meshmod.uniform_periodic(6, 5, 4), KINDS)


def _state(m, N):
    """Smooth flow with velocity toward and away from the walls and a narrow
    pressure bump (drives the indicator)."""
    x, _ = O.lgl(N)
    X = m.node_coords(x)
    s = lambda t: np.sin(2 * np.pi * t)
    c = lambda t: np.cos(2 * np.pi * t)
    rho = 1.0 + 0.2 * s(X[:, 0]) * c(X[:, 2])
    vx, vy, vz = 0.3 + 0.1 * c(X[:, 1]), 0.25 * s(X[:, 0] + X[:, 2]), 0.1 * c(X[:, 0])
    r2 = ((X - np.array([0.3, 0.45, 0.5])[None, :, None]) ** 2).sum(1)
    p = 1.0 + 0.2 * c(X[:, 1] + X[:, 2]) + 3.0 * np.exp(-r2 / (2 * 0.05 ** 2))
    u = np.stack([rho, rho * vx, rho * vy, rho * vz, p / (GAMMA - 1) + 0.5 * rho * (vx ** 2 + vy ** 2 + vz ** 2)], 1)
    return np.ascontiguousarray(u)


def _cfgs(m, N, vol, surf, ind, td=_lib.TD_LSRK54, repair=0, precision=0):
    g = solver.make_config(N=N, volume=vol, surface=surf, indicator=ind, timedisc=td, repair=repair,
                           precision=precision, gamma=GAMMA)
    o = O.with_boundaries(O.make_cfg(N=N, volume=vol, surface=surf, indicator=ind, timedisc=td, repair=repair,
                                     gamma=GAMMA), m.bfaces)
    return g, o


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


COMBOS = [(3, _lib.VOL_EC, _lib.SURF_HLLC, _lib.IND_PP), (3, _lib.VOL_WEAK, _lib.SURF_RUSANOV, _lib.IND_NONE),
          (3, _lib.VOL_CENTRAL, _lib.SURF_HLLC, _lib.IND_NONE), (2, _lib.VOL_EC, _lib.SURF_HLLC, _lib.IND_PP),
          (4, _lib.VOL_EC, _lib.SURF_RUSANOV, _lib.IND_PP)]


@pytest.mark.parametrize("combo", COMBOS, ids=lambda c: f"N{c[0]}-vol{c[1]}-surf{c[2]}-ind{c[3]}")
def test_boundary_rhs_parity(combo):
    N, vol, surf, ind = combo
    m = _mesh()
    u0 = _state(m, N)
    g, o = _cfgs(m, N, vol, surf, ind)
    print(f"boundary rhs parity {combo}: {m.nelem} elements, {len(m.bfaces)} boundary faces", flush=True)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        r = s.kernel_calc_flux()
    aref = O.indicator(o, u0)[0] if ind == _lib.IND_PP else None
    if ind == _lib.IND_PP:
        assert (aref > 0).any()
    err = _rel(r, O.rhs(o, m.nbr, m.delta, u0, aref))
    print(f"  rel Linf {err:.3e}")
    assert err <= RHS_TOL


@pytest.mark.parametrize("td", [_lib.TD_LSRK54, _lib.TD_SSPRK54])
def test_boundary_run_parity(td):
    """20 steps with walls and transmissive sides (EC + HLLC, blending,
    repair) against the oracle's run, dt history to 1e-12."""
    N = 3
    m = _mesh()
    u0 = _state(m, N)
    g, o = _cfgs(m, N, _lib.VOL_EC, _lib.SURF_HLLC, _lib.IND_PP, td=td, repair=1)
    print(f"boundary run parity timedisc={td}", flush=True)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        h = s.run(20, record_dt=True)
        u = s.get_state()
    uref = u0.copy()
    rc, href = O.run(o, m.nbr, m.delta, uref, 20)
    assert rc == 0
    assert np.abs(h - href).max() <= 1e-12 * href.max()
    err = _rel(u, uref)
    print(f"  rel Linf after 20 steps {err:.3e}")
    assert err <= RUN_TOL


def test_boundary_fp32_operator():
    N = 3
    m = _mesh()
    u0 = _state(m, N)
    g, o = _cfgs(m, N, _lib.VOL_EC, _lib.SURF_HLLC, _lib.IND_PP, precision=1)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        r = s.kernel_calc_flux()
    err = _rel(r, O.rhs(o, m.nbr, m.delta, u0, O.indicator(o, u0)[0]))
    print(f"boundary fp32 rhs rel Linf {err:.3e}")
    assert err <= 5e-5


def test_mhd_transmissive_parity():
    """GLM-MHD with transmissive sides in x and y (periodic z): RHS per
    variable <= 1e-12 with the GLM speed and damping set, 20 steps <= 1e-10."""
    N = 3
    base = meshmod.uniform_periodic(5, 4, 4)
    m = meshmod.with_boundaries(base, [_lib.BC_TRANSMISSIVE] * 4 + [None, None])
    u0 = setups.orszag_tang(m, N, gamma=5.0 / 3.0, seed=3)
    mk = lambda repair=0, glm=(0.0, 0.0): (
        solver.make_config(N=N, equations=_lib.EQ_GLM_MHD, volume=_lib.VOL_EC, surface=_lib.SURF_HLLD,
                           indicator=_lib.IND_PP, timedisc=_lib.TD_LSRK54, repair=repair, gamma=5.0 / 3.0, mu0=1.0),
        O.with_boundaries(O.make_cfg(N=N, eq=O.EQ_MHD, volume=O.VOL_EC, surface=O.SURF_HLLD, indicator=O.IND_PP,
                                     timedisc=O.TD_LSRK54, repair=repair, gamma=5.0 / 3.0, mu0=1.0,
                                     glm_ch=glm[0], glm_alpha=glm[1]), m.bfaces))
    # per variable, relative to its own largest value floored at 1e-3 of the
    # largest of all (psi and its RHS are rounding noise on this divergence-
    # free field)
    per_var = lambda a, b: max(float(np.abs(a[:, h] - b[:, h]).max() / max(np.abs(b[:, h]).max(),
                                                                           1e-3 * np.abs(b).max()))
                               for h in range(a.shape[1]))
    glm = (0.8, 4.0)
    g, o = mk(glm=glm)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        s.set_glm(*glm)
        r = s.kernel_calc_flux()
    err = per_var(r, O.rhs(o, m.nbr, m.delta, u0, O.indicator(o, u0)[0]))
    print(f"mhd transmissive rhs rel Linf (worst variable) {err:.3e}", flush=True)
    assert err <= RHS_TOL
    g, o = mk(repair=1)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        h = s.run(20, record_dt=True)
        u = s.get_state()
    uref = u0.copy()
    rc, href = O.run(o, m.nbr, m.delta, uref, 20)
    assert rc == 0 and np.abs(h - href).max() <= 1e-12 * href.max()
    err = max(float(np.abs(u[:, h] - uref[:, h]).max() / max(np.abs(uref[:, h]).max(), 1e-3 * np.abs(uref).max()))
              for h in range(9))
    print(f"  20 steps rel Linf (worst variable) {err:.3e}")
    assert err <= RUN_TOL


def test_boundary_checks_reject_bad_tables():
    m = _mesh()
    g, _ = _cfgs(m, 3, _lib.VOL_EC, _lib.SURF_HLLC, _lib.IND_PP)
    cases = []
    bad = m.bfaces.copy()
    bad[0, 2] = 7
    cases.append(("unknown boundary kind", bad))
    bad = m.bfaces.copy()
    bad[[0, 1]] = bad[[1, 0]]
    cases.append(("does not match nbr", bad))
    with solver.Solver(g) as s:
        for word, bf in cases:
            class _M:
                nbr, delta, nelem, bfaces = m.nbr, m.delta, m.nelem, bf
            with pytest.raises(_lib.NemoGpuError) as exc:
                s.upload_mesh(_M)
            assert word in str(exc.value), (word, str(exc.value))
    gm = solver.make_config(N=3, equations=_lib.EQ_GLM_MHD, volume=_lib.VOL_EC, surface=_lib.SURF_HLLD,
                            gamma=5.0 / 3.0, mu0=1.0)
    with solver.Solver(gm) as s:
        with pytest.raises(_lib.NemoGpuError) as exc:
            s.upload_mesh(m)
        assert "slip wall" in str(exc.value)
