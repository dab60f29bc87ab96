"""Host analysis layer (nemo_amd.analyze): quadrature volumes and error norms
against closed forms."""
import numpy as np

from nemo_amd import amr, analyze, mesh as meshmod
from oracle import oracle as O


def test_quadrature_volumes_integrate_polynomials_exactly():
    N = 3
    for m in (meshmod.uniform_periodic(4), amr.Forest.uniform(1).refine(np.array([1, 0, 0, 0, 0, 0, 0, 1], bool))):
        vol = analyze.quadrature_volumes(m, N)
        x, _ = O.lgl(N)
        X = m.node_coords(x)
        assert abs(vol.sum() - 1.0) < 1e-14
        # x^2 y z^3 on the unit cube: 1/3 * 1/2 * 1/4 (degree <= 2N-1 per direction: exact)
        f = X[:, 0] ** 2 * X[:, 1] * X[:, 2] ** 3
        assert abs((vol * f).sum() - 1.0 / 24.0) < 1e-15


def test_error_norms_closed_form():
    N = 2
    m = meshmod.uniform_periodic(2)
    M3 = (N + 1) ** 3
    u = np.zeros((m.nelem, 5, M3))
    ue = u.copy()
    x, _ = O.lgl(N)
    X = m.node_coords(x)
    u[:, 0] = X[:, 0]            # error x: L1 = 1/2, L2 = 1/sqrt(3), Linf = 1
    u[:, 4] = -2.0               # constant error 2
    e = analyze.error_norms(m, N, u, ue)
    assert abs(e[0]["l1"] - 0.5) < 1e-15 and abs(e[0]["l2"] - 1 / np.sqrt(3)) < 1e-15 and e[0]["linf"] == 1.0
    assert abs(e[4]["l1"] - 2) < 1e-15 and abs(e[4]["l2"] - 2) < 1e-15 and e[4]["linf"] == 2.0
    assert e[1]["linf"] == 0.0
