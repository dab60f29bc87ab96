"""FP32 operator (FP64 state, RK accumulation, CFL and repair) against the
FP64 oracle.  Tolerances: the RHS to FP32 round-off amplified by the
operator's cancellations (5e-5 relative L-inf), and the accuracy bar of
the chosen form, L1 density error <= 1e-6 relative to the FP64 oracle,
after 100 steps on the same dt sequence."""
import numpy as np
import pytest

from nemo_amd import _lib, mesh as meshmod, setups, solver
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _cfgs(N=3):
    g = solver.make_config(N=N, volume=_lib.VOL_EC, surface=_lib.SURF_HLLC, indicator=_lib.IND_PP,
                           timedisc=_lib.TD_LSRK54, repair=1, precision=1)
    o = O.make_cfg(N=N, volume=O.VOL_EC, surface=O.SURF_HLLC, indicator=O.IND_PP, timedisc=O.TD_LSRK54, repair=1)
    return g, o


def _l1_rel(a, b, w):
    return float((w * np.abs(a - b)).sum() / (w * np.abs(b)).sum())


@pytest.mark.parametrize("N", [2, 3, 4])
def test_fp32_rhs_against_fp64_oracle(N):
    m = meshmod.uniform_periodic(8)
    u0 = setups.kelvin_helmholtz(m, N, seed=1234)
    g, o = _cfgs(N)
    print(f"fp32 rhs N={N}", flush=True)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        r = s.kernel_calc_flux()
        a = s.indicator()
    aref, _ = O.indicator(o, u0)
    rref = O.rhs(o, m.nbr, m.delta, u0, aref)
    err = float(np.abs(r - rref).max() / np.abs(rref).max())
    print(f"  rel Linf {err:.3e}, alpha max diff {np.abs(a - aref).max():.2e}")
    assert err <= 5e-5
    assert np.abs(a - aref).max() <= 1e-4


def test_fp32_100_steps_l1_density_error():
    N = 3
    m = meshmod.uniform_periodic(8)
    u0 = setups.kelvin_helmholtz(m, N, seed=1234)
    g, o = _cfgs(N)
    uref = u0.copy()
    rc, dts = O.run(o, m.nbr, m.delta, uref, 100)
    assert rc == 0
    print("fp32 100 steps", flush=True)
    with solver.Solver(g) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        for dt in dts:
            s.step(float(dt))
        s.synchronize()
        u = s.get_state()
    _, w = O.lgl(N)
    W = np.einsum("i,j,k->kji", w, w, w).reshape(-1)[None, :]
    err = _l1_rel(u[:, 0], uref[:, 0], W)
    print(f"  L1 density error {err:.3e}")
    assert err <= 1e-6
