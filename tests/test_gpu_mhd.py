"""GLM-MHD on the GPU (through the C-ABI) against the FP64 CPU oracle.

Bars as for Euler: RHS per call max|d| / max|orc| <= 1e-12 (per variable),
100 RK steps max|d| / max|orc| <= 1e-10.  The FP32 operator (FP64 RK
accumulation, GLM update, repair and CFL) is compared with the FP64 GPU path
per variable in L1 and Linf.
"""
import numpy as np
import pytest

from nemo_amd import _lib, mesh as meshmod, setups, solver
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GAMMA = 5.0 / 3.0
RHS_TOL = 1e-12
RUN_TOL = 1e-10


def _cfgs(vol, surf, ind, td=_lib.TD_LSRK54, repair=0, precision=0, glm=(0.0, 0.0), N=3):
    g = solver.make_config(N=N, equations=_lib.EQ_GLM_MHD, volume=vol, surface=surf, indicator=ind, timedisc=td,
                           repair=repair, gamma=GAMMA, precision=precision, mu0=1.0)
    o = O.make_cfg(N=N, eq=O.EQ_MHD, volume=vol, surface=surf, indicator=ind, timedisc=td, repair=repair,
                   gamma=GAMMA, mu0=1.0, glm_ch=glm[0], glm_alpha=glm[1])
    return g, o


def _state(m, rough, N=3):
    """Orszag-Tang data; `rough` adds density and pressure jumps (and a
    non-zero psi) so that the indicator blends and every term is active."""
    u = setups.orszag_tang(m, N, gamma=GAMMA, seed=7)
    if not rough:
        return u
    xi, _ = O.lgl(N)
    X = m.node_coords(xi)
    rho = u[:, 0] * (1.0 + 0.5 * (X[:, 0] > 0.5))
    v = u[:, 1:4] / u[:, 0:1]
    B = u[:, 5:8]
    p = 5.0 / (12.0 * np.pi) * (1.0 + 0.3 * (X[:, 1] > 0.5))
    psi = 0.02 * np.sin(2 * np.pi * X[:, 2])
    return np.ascontiguousarray(setups.mhd_prim2cons(rho, v.transpose(1, 0, 2), p, B.transpose(1, 0, 2), psi, GAMMA))


def _rel_per_var(a, b):
    return [float(np.abs(a[:, h] - b[:, h]).max() / max(np.abs(b[:, h]).max(), 1e-300)) for h in range(a.shape[1])]


COMBOS = [(v, s, i) for v in (_lib.VOL_EC, _lib.VOL_CENTRAL) for s in (_lib.SURF_HLLD, _lib.SURF_RUSANOV)
          for i in (_lib.IND_NONE, _lib.IND_PP)]


@pytest.mark.parametrize("combo", COMBOS, ids=lambda c: f"vol{c[0]}-surf{c[1]}-ind{c[2]}")
def test_mhd_rhs_parity(combo):
    vol, surf, ind = combo
    m = meshmod.uniform_periodic(6)
    u0 = _state(m, rough=True)
    glm = (0.8, 4.0)
    g, o = _cfgs(vol, surf, ind, glm=glm)
    print(f"mhd rhs parity {combo}", flush=True)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        s.set_glm(*glm)
        r = s.kernel_calc_flux()
        a = s.indicator()
    aref = O.indicator(o, u0)[0] if ind == _lib.IND_PP else None
    if ind == _lib.IND_PP:
        assert np.abs(a - aref).max() <= 1e-12
        assert (aref > 0).any() and (aref == 0).any()
    rref = O.rhs(o, m.nbr, m.delta, u0, aref)
    err = _rel_per_var(r, rref)
    print("  rel Linf per var", " ".join(f"{e:.2e}" for e in err))
    assert max(err) <= RHS_TOL


@pytest.mark.parametrize("ind", [_lib.IND_NONE, _lib.IND_PP])
def test_mhd_fp32_rhs_against_fp64_oracle(ind):
    """FP32 operator's RHS on the rough state (density and pressure jumps,
    non-zero psi, GLM speed and damping set) against the FP64 oracle, per
    variable: the compact FP32 layout's FP64 psi for the damping term and
    its indicator on FP32-rounded rho are both on this path.  Bound as for
    the Euler FP32 RHS: 5e-5 of the variable's max."""
    m = meshmod.uniform_periodic(6)
    u0 = _state(m, rough=True)
    glm = (0.8, 4.0)
    g, o = _cfgs(_lib.VOL_EC, _lib.SURF_HLLD, ind, precision=1, glm=glm)
    print(f"mhd fp32 rhs ind={ind}", flush=True)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        s.set_glm(*glm)
        r = s.kernel_calc_flux()
        a = s.indicator()
    aref = O.indicator(o, u0)[0] if ind == _lib.IND_PP else None
    if ind == _lib.IND_PP:
        assert (aref > 0).any()
        assert np.abs(a - aref).max() <= 1e-4
    rref = O.rhs(o, m.nbr, m.delta, u0, aref)
    err = _rel_per_var(r, rref)
    print("  rel Linf per var", " ".join(f"{e:.2e}" for e in err))
    assert max(err) <= 5e-5
    # the damping term -alpha psi is in the psi equation: dropping it moves
    # that variable far beyond the bound
    damp = 4.0 * np.abs(u0[:, 8]).max() / np.abs(rref[:, 8]).max()
    assert damp > 1e-2


def test_mhd_timestep_and_glm():
    m = meshmod.uniform_periodic(6)
    u0 = _state(m, rough=True)
    g, o = _cfgs(_lib.VOL_EC, _lib.SURF_HLLD, _lib.IND_PP)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        dt, lm, vm = s.hook_calc_timestep()
        ch, alpha = s.get_glm()
    dto, lmo, vmo = O.calc_timestep(o, m.delta, u0)
    assert abs(dt - dto) <= 1e-15 * dto
    assert lm == lmo and vm == vmo
    assert ch == 2.0 * abs(lmo - vmo)
    assert alpha == ch / 0.18


@pytest.mark.parametrize("surf", [_lib.SURF_HLLD, _lib.SURF_RUSANOV])
def test_mhd_run_parity_100_steps(surf):
    m = meshmod.uniform_periodic(4)
    u0 = _state(m, rough=True)
    g, o = _cfgs(_lib.VOL_EC, surf, _lib.IND_PP, repair=1)
    print(f"mhd 100-step parity surf={surf}", flush=True)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        dts = s.run(100, record_dt=True)
        ug = s.get_state()
    uo = u0.copy()
    rc, dto = O.run(o, m.nbr, m.delta, uo, 100)
    assert rc == 0
    assert np.abs(dts - dto).max() <= 1e-12 * dto.max()
    err = _rel_per_var(ug, uo)
    print("  rel Linf per var", " ".join(f"{e:.2e}" for e in err))
    assert max(err) <= RUN_TOL


def test_mhd_fp32_operator_error_per_variable():
    """FP32 operator vs FP64 on the same dt sequence (16^3 Orszag-Tang,
    EC + HLLD, 100 steps): L1 error per variable <= 1e-6 of the variable's
    scale (momentum and field components share one scale)."""
    m = meshmod.uniform_periodic(16)
    u0 = setups.orszag_tang(m, 3, gamma=GAMMA, seed=1234)
    g64, _ = _cfgs(_lib.VOL_EC, _lib.SURF_HLLD, _lib.IND_PP, repair=1)
    g32, _ = _cfgs(_lib.VOL_EC, _lib.SURF_HLLD, _lib.IND_PP, repair=1, precision=1)
    with solver.Solver(g64) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        dts = s.run(100, record_dt=True)
        u64 = s.get_state()
    with solver.Solver(g32) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        for dt in dts:
            s.hook_calc_timestep()  # c_h and alpha of the step, as in the run
            s.step(float(dt))
        s.synchronize()
        u32 = s.get_state()
    _, w = O.lgl(3)
    W = np.einsum("i,j,k->kji", w, w, w).reshape(-1)[None, :]
    # error scale per variable: its own L1 norm for rho and E; the mean over
    # the three components for momentum and field (B_z and psi start at
    # zero, so their own norms are no scale); the field's for psi
    l1 = lambda x: float((W * np.abs(x)).sum())
    group = {0: (0,), 1: (1, 2, 3), 2: (1, 2, 3), 3: (1, 2, 3), 4: (4,), 5: (5, 6, 7), 6: (5, 6, 7), 7: (5, 6, 7),
             8: (5, 6, 7)}
    for h in range(9):
        e1 = l1(u32[:, h] - u64[:, h]) / np.mean([l1(u64[:, k]) for k in group[h]])
        einf = np.abs(u32[:, h] - u64[:, h]).max() / max(np.abs(u64[:, k]).max() for k in group[h])
        print(f"  var {h}: L1 {e1:.2e}  Linf {einf:.2e}")
        assert e1 <= 1e-6, h


@pytest.mark.parametrize("N", [2, 4])
@pytest.mark.parametrize("combo", [(_lib.VOL_EC, _lib.SURF_HLLD, _lib.IND_PP),
                                   (_lib.VOL_CENTRAL, _lib.SURF_RUSANOV, _lib.IND_NONE)],
                         ids=lambda c: f"vol{c[0]}-surf{c[1]}-ind{c[2]}")
def test_mhd_other_degrees_rhs_and_run(N, combo):
    """GLM-MHD at n_poly = 2 and 4: RHS per variable <= 1e-12 against the
    oracle on the rough state (GLM speed and damping set), 20 LSRK steps with
    repair <= 1e-10 per variable, and the FP32 operator's RHS within 5e-5."""
    vol, surf, ind = combo
    m = meshmod.uniform_periodic(4)
    u0 = _state(m, rough=True, N=N)
    glm = (0.8, 4.0)
    g, o = _cfgs(vol, surf, ind, glm=glm, N=N)
    print(f"mhd N={N} {combo}", flush=True)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        s.set_glm(*glm)
        r = s.kernel_calc_flux()
    aref = O.indicator(o, u0)[0] if ind == _lib.IND_PP else None
    if ind == _lib.IND_PP:
        assert (aref > 0).any()
    err = _rel_per_var(r, O.rhs(o, m.nbr, m.delta, u0, aref))
    print("  rhs rel Linf per var", " ".join(f"{e:.2e}" for e in err))
    assert max(err) <= RHS_TOL
    g, o = _cfgs(vol, surf, ind, repair=1, N=N)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        dts = s.run(20, record_dt=True)
        ug = s.get_state()
    uo = u0.copy()
    rc, dto = O.run(o, m.nbr, m.delta, uo, 20)
    assert rc == 0
    assert np.abs(dts - dto).max() <= 1e-12 * dto.max()
    err = _rel_per_var(ug, uo)
    print("  20 steps rel Linf per var", " ".join(f"{e:.2e}" for e in err))
    assert max(err) <= RUN_TOL
    g, o = _cfgs(vol, surf, ind, precision=1, glm=glm, N=N)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        s.set_glm(*glm)
        r32 = s.kernel_calc_flux()
    err = _rel_per_var(r32, O.rhs(o, m.nbr, m.delta, u0, aref))
    print("  fp32 rhs rel Linf per var", " ".join(f"{e:.2e}" for e in err))
    assert max(err) <= 5e-5
