"""CPU checks of the learned indicator's reference network (oracle/mlp_ref.py),
the checker of the GPU training kernels: its gradient against central finite
differences, and its layout against the C-ABI."""
import re
import os

import numpy as np

from nemo_amd import _lib, indicator_mlp
from oracle import mlp_ref


def test_layout_matches_header():
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "nemo_gpu.h")).read()
    n = int(re.search(r"#define NEMO_MLP_NPARAMS (\d+)", hdr).group(1))
    assert n == mlp_ref.NP == _lib.MLP_NPARAMS == indicator_mlp.NPARAMS
    p = np.arange(mlp_ref.NP, dtype=np.float64)
    assert np.array_equal(mlp_ref.pack(*mlp_ref.unpack(p)), p)


def test_gradient_against_finite_differences():
    rng = np.random.default_rng(3)
    p = indicator_mlp.init_params(5).astype(np.float64)
    X = rng.normal(-1.0, 0.8, (17, 4))
    t = rng.uniform(0, 0.5, 17)
    loss, g = mlp_ref.loss_grad(p, X, t)
    # every block of the layout, including the output bias
    probe = [0, 3, 255, 256, 300, 320, 321, 4415, 4416, 4480, 6000, 8575, 8576, 8640, 8703, 8704]
    for k in probe:
        h = 1e-6 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mlp_ref.loss_grad(pp, X, t)[0] - mlp_ref.loss_grad(pm, X, t)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 * max(1.0, abs(fd)) + 1e-9, (k, fd, g[k])


def test_gelu_closed_forms():
    x = np.linspace(-6, 6, 101)
    assert np.allclose(mlp_ref.gelu(np.array([0.0])), 0.0)
    assert np.allclose(mlp_ref.gelu(x[x > 5]), x[x > 5], rtol=1e-6)   # -> x for large x
    assert np.allclose(mlp_ref.gelu(x[x < -5]), 0.0, atol=1e-6)        # -> 0 for large -x
    h = 1e-6
    fd = (mlp_ref.gelu(x + h) - mlp_ref.gelu(x - h)) / (2 * h)
    assert np.allclose(mlp_ref.gelu_grad(x), fd, atol=1e-8)


def test_sedov_setup_energy():
    """The deposited energy integrates to E (quadrature over the LGL nodes)."""
    from nemo_amd import mesh as meshmod, operators
    m = meshmod.uniform_periodic(16)
    u = indicator_mlp.sedov(m, 3, jitter=0.0, radius=0.05, p0=0.0)
    _, w = operators.lgl(3)
    W = np.einsum("i,j,k->kji", w, w, w).reshape(-1)[None, :] * (np.prod(m.delta, axis=1) / 8.0)[:, None]
    E = (W * u[:, 4]).sum()
    assert abs(E - 1.0) < 1e-6
