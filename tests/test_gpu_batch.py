"""Batched ensemble (nemo_gpu_batch_*): independent runs of different sizes
and polynomial degrees advanced together must give, run by run, bit for bit
what a single-run solver gives (same kernels, same time steps), with or
without graph replay, and the autotuner must leave the states untouched."""
import numpy as np
import pytest

from nemo_amd import _lib, mesh as meshmod, setups, solver

pytestmark = pytest.mark.gpu

KW = dict(volume=_lib.VOL_EC, surface=_lib.SURF_HLLC, indicator=_lib.IND_PP, timedisc=_lib.TD_LSRK54, repair=1,
          cfl=0.5, alpha_max=0.5, alpha_min=0.001)


def _runs(seed=5):
    rng = np.random.default_rng(seed)
    spec = [(5, 2), (6, 3), (3, 4), (4, 3), (7, 2), (4, 4)]
    meshes, npolys, states = [], [], []
    for n, N in spec:
        m = meshmod.uniform_periodic(n)
        meshes.append(m)
        npolys.append(N)
        states.append(setups.planar_riemann(m, N, rng))
    return meshes, npolys, states


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_matches_single_runs(precision):
    meshes, npolys, states = _runs()
    steps = 7
    cfg = solver.make_config(N=3, precision=precision, **KW)
    print(f"batch vs single precision={precision}", flush=True)
    with solver.Batch(cfg, meshes, npolys) as b:
        for r, u in enumerate(states):
            b.upload(r, u)
        dts = b.run(steps, use_graph=True, record_dt=True)
        ub = [b.download(r) for r in range(len(states))]
    for r, (m, N, u0) in enumerate(zip(meshes, npolys, states)):
        with solver.Solver(solver.make_config(N=N, precision=precision, **KW)) as s:
            s.upload_mesh(m)
            s.set_state(u0)
            d = s.run(steps, record_dt=True)
            us = s.get_state()
        assert np.array_equal(d, dts[:, r]), r
        assert np.array_equal(us, ub[r]), (r, np.abs(us - ub[r]).max())


def test_graph_and_eager_and_geometry_agree():
    meshes, npolys, states = _runs(seed=9)
    cfg = solver.make_config(N=3, precision=1, **KW)
    out = []
    for use_graph, grid, xcd in ((False, 0, False), (True, 0, False), (True, 2, True), (False, 3, False)):
        with solver.Batch(cfg, meshes, npolys) as b:
            for r, u in enumerate(states):
                b.upload(r, u)
            b.set_launch(grid, xcd)
            b.run(5, use_graph=use_graph)
            out.append([b.download(r) for r in range(len(states))])
    for o in out[1:]:
        for a, c in zip(out[0], o):
            assert np.array_equal(a, c)


def test_autotune_restores_state():
    meshes, npolys, states = _runs(seed=3)
    cfg = solver.make_config(N=3, precision=1, **KW)
    with solver.Batch(cfg, meshes, npolys) as b:
        for r, u in enumerate(states):
            b.upload(r, u)
        grid, xcd, ms = b.autotune(steps=2)
        print("autotuned", grid, xcd, ms)
        after = [b.download(r) for r in range(len(states))]
        b.run(3)
        ua = [b.download(r) for r in range(len(states))]
    for u0, u1 in zip(states, after):
        assert np.array_equal(u0, u1)
    with solver.Batch(cfg, meshes, npolys) as b:
        for r, u in enumerate(states):
            b.upload(r, u)
        b.run(3)
        ub = [b.download(r) for r in range(len(states))]
    for a, c in zip(ua, ub):
        assert np.array_equal(a, c)
