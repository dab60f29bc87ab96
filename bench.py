#!/usr/bin/env python3
"""Throughput of the MI355X spatial operator + RK stage update.

Workload of the headline throughput: 3-D Euler on the
periodic unit cube, 64^3 elements, N=3 (4^3 LGL nodes, 16.8M DOF per
variable), Kelvin-Helmholtz data (seed 1234), HLLC surface flux,
entropy-conservative (Chandrashekar) flux differencing, DG/FV subcell
blending driven by the Persson-Peraire indicator on rho*p with alpha in
[0, 0.5], low-storage RK 5-4, CFL 0.5, positivity repair on.  A "step" is
one RK time step: the CFL reduction plus five fused stage launches, all
device resident.

Metric: DOF-updates per second per RK stage,
    value = n_elem * (N+1)^3 * n_stages * steps / wall.

Prints one JSON line; `value` is the plain FP64 path.  Extra objects:
  chosen        the chosen form (FP32 operator, FP64 RK accumulation, CFL and
                repair) on the same workload: its DOF-updates/s and its L1
                density error against the plain FP64 run on the same dt
                sequence (the metric's "at fixed L1 err", bar 1e-6).
  roofline      achieved HBM bandwidth of the stage kernel (algorithmic bytes
                of the LSRK 2N update: stage 1 moves u in, u out, du out; stages
                2-5 also read du: 3 or 4 x 8 B x 5 variables per DOF) over its
                average launch time, measured with HIP events on the solver's
                stream, against the 8 TB/s HBM3E peak.
  precision_mhd the GLM-MHD precision arm (48^3 Orszag-Tang-type vortex,
                EC + HLLD): FP32 operator vs FP64, per-variable L1/Linf after
                100 steps, and both stage times.
  ensemble      the batched ensemble: 256 independent Riemann-problem runs
                of uneven size and N through persistent, graph-captured stage
                launches with an autotuned launch geometry.
  amr_sedov     Sedov blast on an adaptive forest (levels 3-6) refined by the
                indicator every 8 steps: FP64 + Persson-Peraire and the chosen
                form (FP32 + learned indicator), each adapting on its own
                factors, and the chosen form replayed on the FP64 run's meshes
                and time steps for its L1 density error.
  cpu_baseline  the FP64 C oracle on the host cores (OpenMP), timed on a
                bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
NSTAGES = 5
NVARS = 5


def workload(n: int, N: int, seed: int):
    from nemo_amd import mesh as meshmod, setups
    m = meshmod.uniform_periodic(n)
    u0 = setups.kelvin_helmholtz(m, N, seed=seed)
    return m, u0


def _solver(cfg_kw, N, precision, mlp_params):
    from nemo_amd import solver
    cfg = solver.make_config(N=N, precision=precision, **cfg_kw)
    s = solver.Solver(cfg)
    if mlp_params is not None:
        s.mlp_set_params(mlp_params)
    return s


def gpu_run(m, u0, N, steps, warmup, cfg_kw, precision=0, mlp_params=None):
    """Warm up, then time `steps` device-resident steps from u0.  Returns the
    wall time, stage-launch timing, the final state and the dt sequence."""
    with _solver(cfg_kw, N, precision, mlp_params) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        if warmup:
            s.run(warmup)
        s.set_state(u0)
        if mlp_params is not None:
            s.mlp_stats()  # reset the judge counters
        s.synchronize()
        s.set_timing(True)
        t0 = time.perf_counter()
        dts = s.run(steps, record_dt=True)
        s.synchronize()
        t1 = time.perf_counter()
        launches, kernel_ms = s.read_timing()
        s.set_timing(False)
        u = s.get_state()
        status = s.runtime.stat.ok
        stats = s.mlp_stats() if mlp_params is not None else None
    return t1 - t0, launches, kernel_ms, status, u, dts, stats


def gpu_replay(m, u0, N, dts, cfg_kw, precision, mlp_params=None):
    """Advance u0 with a prescribed dt sequence (same times as another run)."""
    with _solver(cfg_kw, N, precision, mlp_params) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        for dt in dts:
            s.step(float(dt))
        s.synchronize()
        return s.get_state()


def l1_rel(a, b, N):
    from nemo_amd import operators
    _, w = operators.lgl(N)
    W = np.einsum("i,j,k->kji", w, w, w).reshape(-1)[None, :]
    return float((W * np.abs(a - b)).sum() / (W * np.abs(b)).sum())


def mhd_precision(n: int, steps: int, seed: int):
    """GLM-MHD Orszag-Tang-type vortex, n^3 elements, N=3, EC volume flux +
    HLLD, PP blending, LSRK 5-4: the FP32 operator (FP64 RK accumulation,
    GLM update, repair, CFL) against the FP64 path on the same dt sequence,
    per-variable L1 (quadrature-weighted) and Linf errors after `steps` steps.
    Momentum and field components are measured against the mean scale of
    their vector (B_z and psi start at zero); psi against the field's."""
    from nemo_amd import _lib, mesh as meshmod, operators, setups, solver
    m = meshmod.uniform_periodic(n)
    u0 = setups.orszag_tang(m, 3, seed=seed)
    kw = dict(N=3, equations=_lib.EQ_GLM_MHD, volume=_lib.VOL_EC, surface=_lib.SURF_HLLD, indicator=_lib.IND_PP,
              timedisc=_lib.TD_LSRK54, repair=1, gamma=5.0 / 3.0, mu0=1.0, cfl=0.5)
    out = {}
    with solver.Solver(solver.make_config(precision=0, **kw)) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        s.set_timing(True)
        dts = s.run(steps, record_dt=True)
        n64, ms64 = s.read_timing()
        u64 = s.get_state()
    with solver.Solver(solver.make_config(precision=1, **kw)) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        s.set_timing(True)
        for dt in dts:
            s.hook_calc_timestep()  # c_h and alpha of the step
            s.step(float(dt))
        s.synchronize()
        n32, ms32 = s.read_timing()
        u32 = s.get_state()
    _, w = operators.lgl(3)
    W = np.einsum("i,j,k->kji", w, w, w).reshape(-1)[None, :]
    l1 = lambda x: float((W * np.abs(x)).sum())
    group = {0: (0,), 1: (1, 2, 3), 2: (1, 2, 3), 3: (1, 2, 3), 4: (4,), 5: (5, 6, 7), 6: (5, 6, 7), 7: (5, 6, 7),
             8: (5, 6, 7)}
    names = ["rho", "mx", "my", "mz", "E", "Bx", "By", "Bz", "psi"]
    for h in range(9):
        d = u32[:, h] - u64[:, h]
        out[names[h]] = {"l1": l1(d) / np.mean([l1(u64[:, k]) for k in group[h]]),
                         "linf": float(np.abs(d).max() / max(np.abs(u64[:, k]).max() for k in group[h]))}
    return {"workload": f"glm-mhd orszag-tang {n}^3 N=3 ec-hlld-pp-blend lsrk54, {steps} steps, seed {seed}",
            "reference": "FP64 GPU path (pinned to the FP64 oracle: <= 1e-10 after 100 steps on 4^3)",
            "errors_fp32_vs_fp64": out,
            "fp64_stage_us": round(ms64 / max(n64, 1) * 1e3, 1), "fp32_stage_us": round(ms32 / max(n32, 1) * 1e3, 1),
            "dof_per_var": int(m.nelem * 64)}


def ensemble(nruns: int, steps: int, seed: int):
    """Batched ensemble: `nruns` independent planar Riemann problems, elements
    per direction uniform in [8, 32], N uniform in {2, 3, 4}, EC + HLLC with
    PP blending, LSRK 5-4, FP32 operator with FP64 stage accumulation; one
    launch per RK stage over all runs and degrees, graph-captured
    steps, autotuned launch geometry."""
    from nemo_amd import _lib, mesh as meshmod, setups, solver
    rng = np.random.default_rng(seed)
    spec = [(int(rng.integers(8, 33)), int(rng.integers(2, 5))) for _ in range(nruns)]
    meshes = [meshmod.uniform_periodic(n) for n, _ in spec]
    npolys = [N for _, N in spec]
    cfg = solver.make_config(N=3, precision=1, volume=_lib.VOL_EC, surface=_lib.SURF_HLLC, indicator=_lib.IND_PP,
                             timedisc=_lib.TD_LSRK54, repair=1, cfl=0.5, alpha_max=0.5, alpha_min=0.001)
    dof = sum(m.nelem * (N + 1) ** 3 for m, N in zip(meshes, npolys))
    t0 = time.perf_counter()
    with solver.Batch(cfg, meshes, npolys) as b:
        for r, m in enumerate(meshes):
            b.upload(r, setups.planar_riemann(m, npolys[r], rng))
        t_setup = time.perf_counter() - t0
        tuned = b.autotune(steps=2)
        table = [dict(t, ms_per_step=round(t["ms_per_step"], 3)) for t in b.tune_table()]
        b.run(2)  # warm-up (builds the graph)
        t1 = time.perf_counter()
        b.run(steps, use_graph=True)
        wall = time.perf_counter() - t1
    return {"workload": f"{nruns} planar Riemann problems, 8^3..32^3 elements, N in {{2,3,4}}, seed {seed}",
            "runs": nruns, "elements": int(sum(m.nelem for m in meshes)), "dof_per_var": int(dof),
            "runs_per_N": {str(N): npolys.count(N) for N in (2, 3, 4)},
            "value": dof * NSTAGES * steps / wall, "unit": "DOF-updates/s", "steps": steps,
            "ms_per_step": wall / steps * 1e3,
            "autotuned": dict(tuned, ms_per_step=round(tuned["ms_per_step"], 3)),
            "best_one_launch_ms_per_step": min(t["ms_per_step"] for t in table if t["one_launch"]),
            "best_per_degree_ms_per_step": min(t["ms_per_step"] for t in table if not t["one_launch"]),
            "tuning_table": table,
            "setup_s": round(t_setup, 1)}


def amr_sedov(steps: int, seed: int, lmin: int = 3, lmax: int = 6, every: int = 8):
    """Sedov blast (E = 1 in a Gaussian of radius 0.02, rho = 1, p = 1e-5,
    seeded centre jitter) on an adaptive forest of levels lmin..lmax, N = 3,
    refinement driven by the blending factor of the indicator every `every`
    steps.  Legs: (a) FP64 operator + Persson-Peraire indicator, adapting on
    its own factors, recorded; (b) the chosen form (FP32 operator + learned
    indicator) adapting on its own factors, timed; (c) the chosen form
    replaying (a)'s meshes and time steps, for its L1 density error against
    (a) on the same final mesh."""
    from nemo_amd import _lib, amr, amr_run, indicator_mlp, operators, solver
    N = 3
    kw = dict(N=N, volume=_lib.VOL_EC, surface=_lib.SURF_HLLC, timedisc=_lib.TD_LSRK54, repair=1, cfl=0.5,
              alpha_max=0.5, alpha_min=0.001)
    cfg64 = solver.make_config(indicator=_lib.IND_PP, precision=0, **kw)
    cfg32 = solver.make_config(indicator=_lib.IND_MLP, precision=1, **kw)
    params = indicator_mlp.load_params()
    ic = lambda f: indicator_mlp.sedov(f, N, seed=seed, radius=0.02, p0=1e-5)
    f0 = amr_run.initial_forest(lambda: solver.Solver(cfg64), amr.Forest.uniform(lmin, minlevel=lmin, maxlevel=lmax),
                                ic)
    u0 = ic(f0)
    M3 = (N + 1) ** 3
    uniform = (1 << lmax) ** 3
    res = {"workload": f"Sedov blast E=1, Gaussian radius 0.02, p0=1e-5, levels {lmin}-{lmax}, N={N}, "
                       f"adapt every {every} steps, seed {seed}",
           "steps": steps, "initial_elements": f0.nelem, "uniform_elements": uniform,
           "initial_levels": np.bincount(f0.level, minlength=lmax + 1)[lmin:].tolist()}
    legs = {}
    for name, cfg in (("fp64_pp", cfg64), ("chosen", cfg32)):
        with solver.Solver(cfg) as s:
            s.upload_mesh(f0)
            s.set_state(u0)
            if cfg is cfg32:
                s.mlp_set_params(params)
            run = amr_run.AdaptiveRun(s, f0, every=every)
            t0 = time.perf_counter()
            run.advance(steps, record=(name == "fp64_pp"))
            s.synchronize()
            wall = time.perf_counter() - t0
            leg = {"value": run.dof_updates / wall, "unit": "DOF-updates/s",
                   "uniform_equivalent": uniform * M3 * NSTAGES * steps / wall,
                   "ms_per_step": wall / steps * 1e3, "remeshes": run.remeshes, "fallbacks": run.fallbacks,
                   "remesh_ms_total": round(run.remesh_seconds * 1e3, 1),
                   "elements_min": int(min(run.elements)), "elements_max": int(max(run.elements)),
                   "final_elements": run.forest.nelem,
                   "final_levels": np.bincount(run.forest.level, minlength=lmax + 1)[lmin:].tolist()}
            if cfg is cfg32:
                j, o, fb = s.mlp_stats()
                leg["judge"] = {"judged": j, "overridden": o, "fallback": fb}
            else:
                history, f_end, u_end = run.history, run.forest, s.get_state()
            legs[name] = leg
    _, w = operators.lgl(N)
    w3 = (w[:, None, None] * w[None, :, None] * w[None, None, :]).reshape(-1)
    vol = np.prod(f_end.delta, 1)[:, None] * w3[None, :] / 8.0
    res.update(legs)
    # replays of the FP64 + PP run's meshes and time steps: the chosen form,
    # and each of its two departures alone (FP32 operator; learned indicator)
    for name, cfg in (("chosen", cfg32), ("fp32_pp", solver.make_config(indicator=_lib.IND_PP, precision=1, **kw)),
                      ("fp64_mlp", solver.make_config(indicator=_lib.IND_MLP, precision=0, **kw))):
        with solver.Solver(cfg) as s:
            s.upload_mesh(f0)
            s.set_state(u0)
            if cfg.indicator == _lib.IND_MLP:
                s.mlp_set_params(params)
            amr_run.replay(s, history)
            u_rep = s.get_state()
        err = float(np.sum(np.abs(u_rep[:, 0] - u_end[:, 0]) * vol) / np.sum(np.abs(u_end[:, 0]) * vol))
        res[f"{name}_vs_fp64_same_meshes"] = {"l1_density_err": err, "final_elements": f_end.nelem}
    return res


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n: int, N: int, seed: int, steps: int, cfg_kw, u_gpu=None):
    """FP64 oracle on the host cores, timed as the reference times itself
    (its performance index, driver_mod.f90:83,142; OpenMP over the cores
    with OMP_PROC_BIND=close and OMP_PLACES=cores as in its run script):
    one untimed warm-up step on an 8^3 mesh, then `steps` full steps of the
    same n^3 workload on every core this process may use (its CPU affinity,
    or OMP_NUM_THREADS where the environment sets it lower), then one step
    of a 32^3 sample on one core.  With u_gpu (the GPU state after the same
    steps from the same initial data) it also reports the relative Linf
    difference at full size."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    threads = min(cores, int(env)) if env and env.isdigit() and int(env) > 0 else cores
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle import oracle as O
    kw = dict(cfg_kw)
    ocfg = O.make_cfg(N=N, volume=kw["volume"], surface=kw["surface"], indicator=kw["indicator"],
                      timedisc=kw["timedisc"], repair=kw["repair"], cfl=kw["cfl"])
    O.set_num_threads(threads)
    mw, uw = workload(8, N, seed)
    O.run(ocfg, mw.nbr, mw.delta, uw, 1)  # warm-up
    m, u = workload(n, N, seed)
    t0 = time.perf_counter()
    rc, _ = O.run(ocfg, m.nbr, m.delta, u, steps)
    wall = time.perf_counter() - t0
    dof = m.nelem * (N + 1) ** 3
    value = dof * NSTAGES * steps / wall
    used = O.num_threads()
    # one core, smaller sample of the same workload
    O.set_num_threads(1)
    m1, u1 = workload(32, N, seed)
    t0 = time.perf_counter()
    O.run(ocfg, m1.nbr, m1.delta, u1, 1)
    wall1 = time.perf_counter() - t0
    v1 = m1.nelem * (N + 1) ** 3 * NSTAGES / wall1
    O.set_num_threads(threads)
    out = {"value": value, "unit": "DOF-updates/s", "cores": used, "kind": "port",
           "sample": f"{n}^3 elements, N={N}, {steps} full LSRK 5-4 step(s) of the same KH workload after a "
                     f"one-step warm-up on 8^3, {wall:.1f} s wall, FP64 C oracle, OpenMP on {used} threads",
           "cpu_model": _cpu_model(), "cores_available": cores,
           "perf_index_us_per_dof_stage_core": 1e6 * used / value,
           "one_core": {"value": v1, "unit": "DOF-updates/s", "sample": f"32^3 elements, 1 step, {wall1:.1f} s",
                        "perf_index_us_per_dof_stage_core": 1e6 / v1},
           "status_ok": rc == 0}
    if u_gpu is not None:
        out["parity_full_size"] = {"steps": steps,
                                   "rel_linf_gpu_vs_oracle": float(np.abs(u_gpu - u).max() / np.abs(u).max()),
                                   "bar": 1e-10}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--elems", type=int, default=64, help="elements per direction")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-chosen", action="store_true", help="skip the FP32-operator (chosen form) leg")
    ap.add_argument("--mhd-elems", type=int, default=48, help="elements per direction of the MHD precision leg")
    ap.add_argument("--mhd-steps", type=int, default=100)
    ap.add_argument("--no-mhd", action="store_true", help="skip the GLM-MHD precision leg")
    ap.add_argument("--ensemble-runs", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=10)
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--amr-steps", type=int, default=200)
    ap.add_argument("--no-amr", action="store_true", help="skip the adaptive Sedov leg")
    ap.add_argument("--calibrate", action="store_true",
                    help="end with one plain device copy of 1 GiB: the known-size kernel that scales "
                         "FETCH_SIZE/WRITE_SIZE in counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the library runs on one GPU (asked for --gpus {args.gpus}); "
              "sharding over several GPUs is a design only", file=sys.stderr)
        sys.exit(2)

    from nemo_amd import _lib
    N = 3
    cfg_kw = dict(volume=_lib.VOL_EC, surface=_lib.SURF_HLLC, indicator=_lib.IND_PP, timedisc=_lib.TD_LSRK54,
                  repair=1, cfl=0.5, alpha_max=0.5, alpha_min=0.001)
    m, u0 = workload(args.elems, N, args.seed)
    wall, launches, kernel_ms, ok, u64, dts, _ = gpu_run(m, u0, N, args.steps, args.warmup, cfg_kw, precision=0)
    dof = m.nelem * (N + 1) ** 3
    value = dof * NSTAGES * args.steps / wall

    # algorithmic bytes of the stage launches: stage 1 = 3 arrays, 2..5 = 4 arrays
    bytes_per_step = dof * NVARS * 8 * (3 + 4 * 4)
    total_bytes = bytes_per_step * args.steps
    stage_launches = NSTAGES * args.steps
    assert launches == stage_launches, (launches, stage_launches)
    avg_launch_s = kernel_ms / 1e3 / launches
    achieved = total_bytes / (kernel_ms / 1e3) / 1e9
    # HBM bytes per launch from the calibrated FETCH_SIZE/WRITE_SIZE counter
    # runs of this workload (profiles/traffic_stage_kernel_kh64_n3.json)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_stage_kernel_kh64_n3.json")
    if os.path.exists(tpath) and args.elems == 64:
        with open(tpath) as f:
            traffic = json.load(f)["traffic_bytes_per_launch"]
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "stage_kernel<4,4,EC,HLLC,double>", "avg_launch_us": round(avg_launch_s * 1e6, 2),
                "bytes_per_launch_avg": int(bytes_per_step / NSTAGES)}

    chosen = None
    if not args.no_chosen:
        # chosen form: FP32 operator (FP64 state, RK accumulation, CFL, repair)
        # with the learned indicator (Persson-Peraire judging 5% per step)
        from nemo_amd import indicator_mlp
        params = indicator_mlp.load_params() if os.path.exists(indicator_mlp.DEFAULT_PARAMS) else None
        kw = dict(cfg_kw, indicator=_lib.IND_MLP) if params is not None else cfg_kw
        wall32, l32, ms32, ok32, _, _, st = gpu_run(m, u0, N, args.steps, args.warmup, kw, precision=1,
                                                   mlp_params=params)
        u32 = gpu_replay(m, u0, N, dts, kw, precision=1, mlp_params=params)
        err = l1_rel(u32[:, 0], u64[:, 0], N)
        v32 = dof * NSTAGES * args.steps / wall32
        # the same FP32 operator with the plain indicator, for reference
        wpp, lpp, mspp, _, _, _, _ = gpu_run(m, u0, N, args.steps, 1, cfg_kw, precision=1)
        upp = gpu_replay(m, u0, N, dts, cfg_kw, precision=1)
        chosen = {"form": ("FP32 operator and LSRK register, FP64 state/update/CFL/repair, learned MLP indicator "
                           "at the first stage of each step (PP judge on 5%/step), stored factor at stages 2-5")
                          if params is not None else
                          "FP32 operator and LSRK register, FP64 state/update/CFL/repair, Persson-Peraire indicator "
                          "(no trained network found)",
                  "value": v32, "unit": "DOF-updates/s", "ms_per_step": wall32 / args.steps * 1e3,
                  "l1_density_err_vs_fp64": err, "l1_bar": 1e-6, "meets_l1_bar": err <= 1e-6,
                  "speedup_vs_plain": v32 / value,
                  "stage_kernel_avg_us": round(ms32 / 1e3 / l32 * 1e6, 2),
                  "hbm_frac": round(total_bytes / (ms32 / 1e3) / 1e9 / HBM_PEAK_GBS, 4), "status_ok": ok32,
                  "judge": None if st is None else {"judged": st[0], "overridden": st[1], "fallback": st[2],
                                                     "elements_x_steps": m.nelem * args.steps},
                  "fp32_pp": {"value": dof * NSTAGES * args.steps / wpp,
                              "stage_kernel_avg_us": round(mspp / 1e3 / lpp * 1e6, 2),
                              "l1_density_err_vs_fp64": l1_rel(upp[:, 0], u64[:, 0], N)}}

    out = {
        "metric": "DOF-updates/s per RK stage: plain FP64 vs chosen form at fixed L1 err",
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"euler3d-kh-{args.elems}^3-N3-ec-hllc-pp-blend-lsrk54-fp64",
                   "elements": m.nelem, "N": N, "dof_per_var": dof, "seed": args.seed,
                   "form": "plain FP64"},
        "roofline": roofline,
        "chosen": chosen,
        "status_ok": ok,
    }
    if not args.no_mhd:
        out["precision_mhd"] = mhd_precision(args.mhd_elems, args.mhd_steps, args.seed)
    if not args.no_ensemble:
        out["ensemble"] = ensemble(args.ensemble_runs, args.ensemble_steps, args.seed)
    if not args.no_amr:
        out["amr_sedov"] = amr_sedov(args.amr_steps, args.seed)
    if not args.no_cpu:
        # the GPU state after the same steps from the same data: a parity
        # point at full size for the CPU leg to compare with
        from nemo_amd import solver
        with solver.Solver(solver.make_config(N=N, precision=0, **cfg_kw)) as s:
            s.upload_mesh(m)
            s.set_state(u0)
            s.run(args.cpu_steps)
            u_gpu = s.get_state()
        out["cpu_baseline"] = cpu_baseline(args.elems, N, args.seed, args.cpu_steps, cfg_kw, u_gpu)
        del u_gpu
    print(json.dumps(out), flush=True)
    if args.calibrate:
        from nemo_amd import solver
        with solver.Solver(solver.make_config(N=N)) as s:
            s._call("nemo_gpu_copy_calibration", s._h, ctypes.c_int64(1 << 30))


if __name__ == "__main__":
    main()
