"""The learned indicator on the GPU (through the C-ABI): training kernels
against the FP64 reference network, inference fused in the stage kernel,
the Persson-Peraire judge and fallback, and what the trained network gains.
"""
import os

import numpy as np
import pytest

from nemo_amd import _lib, indicator_mlp, mesh as meshmod, setups, solver
from oracle import mlp_ref, oracle as O

pytestmark = pytest.mark.gpu


def _cfg(ind=_lib.IND_PP, precision=0, **kw):
    base = dict(indicator=ind, precision=precision)
    base.update(indicator_mlp.SCHEME)
    base.update(kw)
    return solver.make_config(**base)


def _features_np(u):
    """Features of every element from the oracle's Vandermonde matrix."""
    x, _ = O.lgl(3)
    _, Vi = O.vandermonde(3, x)
    M = 4
    out = np.zeros((u.shape[0], 4))
    rho = u[:, 0]
    mom = u[:, 1:4]
    p = 0.4 * (u[:, 4] - 0.5 * (mom ** 2).sum(1) / rho)
    for c, q in enumerate((rho, p)):
        Q = q.reshape(-1, M, M, M)  # [e][k][j][i]
        m = np.einsum("ai,bj,ck,ekji->ecba", Vi, Vi, Vi, Q)
        e2 = m ** 2
        tot = e2.sum((1, 2, 3))
        idx = np.arange(M)
        top = np.maximum.outer(np.maximum.outer(idx, idx), idx)  # max mode index, [c][b][a]
        s1 = (e2 * (top == M - 1)).sum((1, 2, 3))
        s2 = (e2 * (top >= M - 2)).sum((1, 2, 3))
        out[:, 2 * c] = mlp_ref.feature(s1 / tot)
        out[:, 2 * c + 1] = mlp_ref.feature(s2 / tot)
    return out


def _clamp(y, amin=0.001, amax=0.5):
    a = np.where(y < amin, 0.0, y)
    a = np.where(a > 1 - amin, 1.0, a)
    return np.minimum(a, amax)


def test_mlp_gradient_against_reference():
    rng = np.random.default_rng(11)
    p = indicator_mlp.init_params(3)
    X = rng.normal(-1.0, 0.7, (333, 4)).astype(np.float32)
    t = rng.uniform(0, 0.5, 333).astype(np.float32)
    with solver.Solver(_cfg()) as s:
        s.mlp_set_params(p)
        loss, g = s.mlp_grad(X, t)
        y = s.mlp_forward(X)
    lref, gref = mlp_ref.loss_grad(p.astype(np.float64), X.astype(np.float64), t.astype(np.float64))
    yref, _ = mlp_ref.forward(p.astype(np.float64), X.astype(np.float64))
    print(f"loss {loss:.6e} ref {lref:.6e}; grad rel {np.abs(g - gref).max() / np.abs(gref).max():.2e}")
    assert np.abs(y - yref).max() <= 2e-6
    assert abs(loss - lref) <= 1e-5 * lref
    assert np.abs(g - gref).max() <= 1e-4 * np.abs(gref).max()


def test_training_reduces_loss():
    m = meshmod.uniform_periodic(12)
    feat, alpha = indicator_mlp.collect([(m, indicator_mlp.sedov(m, 3, seed=3))], steps=40, every=4, precision=0,
                                        log=lambda *a: None)
    assert (alpha > 0).mean() > 0.01
    params, hist = indicator_mlp.train(feat, alpha, epochs=12, batch=512, lr=2e-3, seed=1)
    print("loss per epoch", hist)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.5 * hist[0]
    # the trained network reproduces the reference forward pass
    y_ref, _ = mlp_ref.forward(params.astype(np.float64), feat[:500].astype(np.float64))
    with solver.Solver(_cfg()) as s:
        s.mlp_set_params(params)
        y = s.mlp_forward(feat[:500])
    assert np.abs(y - y_ref).max() <= 2e-6


@pytest.mark.parametrize("precision", [0, 1])
def test_features_and_fused_inference(precision):
    m = meshmod.uniform_periodic(8)
    u0 = indicator_mlp.sedov(m, 3, seed=5, radius=0.05)
    params = indicator_mlp.init_params(9)
    # the features the stage kernel computes
    with solver.Solver(_cfg(precision=precision)) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        f, a_pp = s.indicator_features()
    fref = _features_np(u0)
    if precision == 0:
        assert np.abs(f - fref).max() <= 1e-6
    else:
        # FP32 modal coefficients carry ~1e-7 |q| of rounding, so a shell
        # energy fraction f is known to ~2e-7/sqrt(f): tight for f > 1e-4
        # (feature > -0.92), only rough near the 1e-12 floor
        hi = fref > -0.92
        assert np.abs(f - fref)[hi].max() <= 1e-5
        assert np.abs(f - fref).max() <= 5e-3
    aref, _ = O.indicator(O.make_cfg(N=3, indicator=O.IND_PP), u0)
    assert np.abs(a_pp - aref).max() <= (1e-12 if precision == 0 else 1e-4)
    # inference fused in the stage kernel, judge off
    with solver.Solver(_cfg(ind=_lib.IND_MLP, precision=precision)) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        s.mlp_set_params(params)
        s.mlp_set_policy(0.0, 0.05)
        a = s.indicator()
        stats = s.mlp_stats()
    y, _ = mlp_ref.forward(params.astype(np.float64), f.astype(np.float64))
    expect = _clamp(y)
    ok = np.abs(y - 0.001) > 1e-5  # away from the alpha_min cut
    assert np.abs(a - expect)[ok].max() <= 2e-6
    assert stats == (0, 0, 0)


def test_judge_and_fallback():
    m = meshmod.uniform_periodic(8)
    u0 = indicator_mlp.sedov(m, 3, seed=6, radius=0.05)
    # one element with a non-positive pressure node
    bad = 37
    u0[bad, 4, 5] = 0.5 * (u0[bad, 1:4, 5] ** 2).sum() / u0[bad, 0, 5] - 1e-3
    params = indicator_mlp.init_params(4)
    with solver.Solver(_cfg(ind=_lib.IND_MLP)) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        s.mlp_set_params(params)
        s.mlp_set_policy(1.0, 0.0)   # judge every element, no tolerance
        a = s.indicator()
        # the judge's counters belong to the stage sequence: a first stage
        # (dt = 0 leaves the state as it is) refreshes and judges the factor
        s.stage(1, 0.0)
        s.synchronize()
        judged, overridden, fallback = s.mlp_stats()
        f, a_pp = s.indicator_features()
    y, _ = mlp_ref.forward(params.astype(np.float64), f.astype(np.float64))
    a_mlp = _clamp(y)
    expect = np.where(a_pp > a_mlp, a_pp, a_mlp)
    expect[bad] = a_pp[bad]
    ok = np.abs(y - 0.001) > 1e-5
    assert np.abs(a - expect)[ok].max() <= 2e-6
    assert judged == m.nelem
    assert abs(overridden - int((a_pp > a_mlp).sum())) <= 2
    assert fallback == 1


def test_sampled_judge_fraction():
    m = meshmod.uniform_periodic(16)
    u0 = setups.kelvin_helmholtz(m, 3, seed=2)
    with solver.Solver(_cfg(ind=_lib.IND_MLP, precision=1)) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        s.mlp_set_params(indicator_mlp.init_params(1))
        s.mlp_set_policy(0.05, 0.05)
        s.run(10)
        judged, _, _ = s.mlp_stats()
    frac = judged / (10 * m.nelem)
    print("judged fraction", frac)
    assert 0.04 < frac < 0.06


@pytest.mark.skipif(not os.path.exists(indicator_mlp.DEFAULT_PARAMS), reason="no trained parameters")
def test_trained_network_gain_on_sedov():
    """The trained network on Sedov blasts it was not trained on (new
    centres, 16^3 and 24^3), FP32 operator, 80 steps.  Against the method
    without it (pure DG, no indicator) it keeps the post-shock density
    undershoot of the plain indicator (min rho ~ 5e-2 instead of ~1e-3) and
    the plain indicator's solution (L1 density difference < 1e-4, against
    ~3e-3 for pure DG), while the Persson-Peraire indicator only runs on the
    sampled 5% of elements per step.  At the final state fewer than 0.5% of
    the elements have a prediction more than the judge's tolerance (0.05)
    below the plain indicator (the judge corrects those only where it
    samples them)."""
    params = indicator_mlp.load_params()
    for n, seed in ((16, 987), (24, 4321)):
        m = meshmod.uniform_periodic(n)
        u0 = indicator_mlp.sedov(m, 3, seed=seed)
        res, alphas = {}, {}
        for name, ind in (("pp", _lib.IND_PP), ("mlp", _lib.IND_MLP), ("none", _lib.IND_NONE)):
            with solver.Solver(_cfg(ind=ind, precision=1)) as s:
                s.upload_mesh(m)
                s.set_state(u0)
                if ind == _lib.IND_MLP:
                    s.mlp_set_params(params)
                try:
                    s.run(80)
                    res[name] = s.get_state()
                    alphas[name] = s.indicator()
                    if ind == _lib.IND_MLP:
                        print(f"  {n}^3 judge stats", s.mlp_stats())
                except _lib.NemoGpuError as e:
                    res[name] = e
            print(f"  {n}^3 {name}: {'ok' if isinstance(res[name], np.ndarray) else res[name]}")
        assert isinstance(res["pp"], np.ndarray) and isinstance(res["mlp"], np.ndarray)
        d = np.abs(res["mlp"][:, 0] - res["pp"][:, 0]).sum() / np.abs(res["pp"][:, 0]).sum()
        a_pp, a_mlp = alphas["pp"], alphas["mlp"]
        under = np.mean(a_pp - a_mlp > 0.05)
        print(f"  {n}^3 L1 density mlp vs pp {d:.3e}; blended pp {np.mean(a_pp > 0):.4f} mlp {np.mean(a_mlp > 0):.4f};"
              f" max underestimate {np.max(a_pp - a_mlp):.3f}, fraction under by > 0.05: {under:.5f}")
        assert d < 1e-4
        assert under < 0.005
        rmin = {k: float(res[k][:, 0].min()) for k in ("pp", "mlp")}
        assert rmin["mlp"] >= 0.9 * rmin["pp"]
        if isinstance(res["none"], np.ndarray):
            dn = np.abs(res["none"][:, 0] - res["pp"][:, 0]).sum() / np.abs(res["pp"][:, 0]).sum()
            rn = float(res["none"][:, 0].min())
            print(f"  {n}^3 L1 density none vs pp {dn:.3e}, min rho none {rn:.3e}"
                  f" pp {rmin['pp']:.3e} mlp {rmin['mlp']:.3e}")
            assert d < 0.05 * dn
            assert rmin["mlp"] > 5.0 * rn


@pytest.mark.parametrize("precision", [0, 1])
def test_diagnostics_between_stages_leave_the_step_untouched(precision):
    """nemo_gpu_rhs, nemo_gpu_indicator and nemo_gpu_indicator_features called
    between the stages of a step evaluate a fresh factor, but neither the
    factor that stages 2..5 use nor the judge's counters change: the step
    ends bit for bit where an uninterrupted step ends."""
    m = meshmod.uniform_periodic(8)
    u0 = indicator_mlp.sedov(m, 3, seed=7, radius=0.05)
    params = indicator_mlp.init_params(5)
    out = {}
    for interrupt in (False, True):
        with solver.Solver(_cfg(ind=_lib.IND_MLP, precision=precision)) as s:
            s.upload_mesh(m)
            s.set_state(u0)
            s.mlp_set_params(params)
            s.mlp_set_policy(1.0, 0.0)   # every element judged: the counters move each step
            dt, _, _ = s.hook_calc_timestep()
            s.stage(1, dt)
            if interrupt:
                s.kernel_calc_flux()
                a_mid = s.indicator()
                s.indicator_features()
            for st in range(2, 6):
                s.stage(st, dt)
            s.synchronize()
            out[interrupt] = (s.get_state(), s.mlp_stats())
    assert np.array_equal(out[False][0], out[True][0])
    assert out[False][1] == out[True][1] and out[False][1][0] == m.nelem
    # blending is active, so a factor written back would change stages 2..5
    assert np.any(a_mid > 0)
