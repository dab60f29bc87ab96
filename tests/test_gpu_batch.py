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
    geoms = ((False, 0, False, True, 0), (True, 0, False, True, 0), (True, 2, True, True, 1), (False, 3, False, True, 0),
             (True, 0, False, False, 0), (True, 2, True, False, 0), (False, 3, False, False, 0))
    for use_graph, grid, xcd, one, order in geoms:
        with solver.Batch(cfg, meshes, npolys) as b:
            for r, u in enumerate(states):
                b.upload(r, u)
            b.set_launch(grid, xcd, one, order)
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
        tuned = b.autotune(steps=2)
        table = b.tune_table()
        print("autotuned", tuned)
        # every candidate was timed and the kept geometry is the fastest
        assert len(table) == 9
        assert {t["one_launch"] for t in table} == {True, False}
        best = min(table, key=lambda t: t["ms_per_step"])
        assert (best["one_launch"], best["order"], best["grid_per_cu"], best["xcd_map"]) == \
            (tuned["one_launch"], tuned["order"], tuned["grid_per_cu"], tuned["xcd_map"])
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


def _oracle_cfg(N):
    from oracle import oracle as O
    return O.make_cfg(N=N, volume=O.VOL_EC, surface=O.SURF_HLLC, indicator=O.IND_PP, timedisc=O.TD_LSRK54,
                      repair=1, cfl=0.5, alpha_max=0.5, alpha_min=0.001)


def _uneven_riemann_runs(seed):
    """Five planar Riemann problems of uneven size and degree (3^3..6^3
    elements, N in {2, 3, 4}): the ensemble workload at test scale."""
    rng = np.random.default_rng(seed)
    spec = [(3, 2), (5, 3), (4, 4), (6, 2), (4, 3)]
    meshes, npolys, states = [], [], []
    for n, N in spec:
        m = meshmod.uniform_periodic(n)
        meshes.append(m)
        npolys.append(N)
        states.append(setups.planar_riemann(m, N, rng))
    return meshes, npolys, states


def test_batch_fp64_against_oracle():
    """The batched path itself (ghost-padded blocks, per-run time steps from
    the per-run CFL reduction, graph replay of the stage sequence) against
    the FP64 oracle run by run: 20 LSRK 5-4 steps with blending and repair,
    state <= 1e-10 and dt history <= 1e-12 relative."""
    from oracle import oracle as O
    meshes, npolys, states = _uneven_riemann_runs(seed=21)
    steps = 20
    cfg = solver.make_config(N=3, precision=0, **KW)
    print("batch fp64 vs oracle", flush=True)
    with solver.Batch(cfg, meshes, npolys) as b:
        for r, u in enumerate(states):
            b.upload(r, u)
        dts = b.run(steps, use_graph=True, record_dt=True)
        ub = [b.download(r) for r in range(len(states))]
    blended = 0
    for r, (m, N, u0) in enumerate(zip(meshes, npolys, states)):
        uo = u0.copy()
        rc, dto = O.run(_oracle_cfg(N), m.nbr, m.delta, uo, steps)
        assert rc == 0
        ddt = np.abs(dts[:, r] - dto).max() / dto.max()
        err = np.abs(ub[r] - uo).max() / np.abs(uo).max()
        a, _ = O.indicator(_oracle_cfg(N), uo)
        blended += int((a > 0).sum())
        print(f"  run {r}: {m.nelem} el N={N}: state {err:.2e} dt {ddt:.2e}")
        assert ddt <= 1e-12, r
        assert err <= 1e-10, r
    assert blended > 0  # the FV subcells took part


def test_batch_fp32_against_oracle():
    """FP32 operator in the batch (FP64 state, RK, CFL and repair): L1
    density error <= 1e-6 relative to the FP64 oracle after 20 steps, run
    by run (each on its own CFL time steps)."""
    from oracle import oracle as O
    meshes, npolys, states = _uneven_riemann_runs(seed=22)
    steps = 20
    cfg = solver.make_config(N=3, precision=1, **KW)
    print("batch fp32 vs oracle", flush=True)
    with solver.Batch(cfg, meshes, npolys) as b:
        for r, u in enumerate(states):
            b.upload(r, u)
        b.run(steps, use_graph=True)
        ub = [b.download(r) for r in range(len(states))]
    for r, (m, N, u0) in enumerate(zip(meshes, npolys, states)):
        uo = u0.copy()
        rc, _ = O.run(_oracle_cfg(N), m.nbr, m.delta, uo, steps)
        assert rc == 0
        _, w = O.lgl(N)
        W = np.einsum("i,j,k->kji", w, w, w).reshape(-1)[None, :]
        l1 = float((W * np.abs(ub[r][:, 0] - uo[:, 0])).sum() / (W * np.abs(uo[:, 0])).sum())
        print(f"  run {r}: {m.nelem} el N={N}: L1 density {l1:.2e}")
        assert l1 <= 1e-6, r
