#!/usr/bin/env python3
"""Throughput of the MI355X spatial operator + RK stage update.

Workload (the configuration the metric is quoted on): 3-D Euler on the
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
  cpu_baseline  the FP64 C oracle on the host cores (OpenMP), timed on a
                bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
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


def gpu_run(m, u0, N, steps, warmup, cfg_kw, precision=0):
    """Warm up, then time `steps` device-resident steps from u0.  Returns the
    wall time, stage-launch timing, the final state and the dt sequence."""
    from nemo_amd import solver
    cfg = solver.make_config(N=N, precision=precision, **cfg_kw)
    with solver.Solver(cfg) as s:
        s.upload_mesh(m)
        s.set_state(u0)
        if warmup:
            s.run(warmup)
        s.set_state(u0)
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
    return t1 - t0, launches, kernel_ms, status, u, dts


def gpu_replay(m, u0, N, dts, cfg_kw, precision):
    """Advance u0 with a prescribed dt sequence (same times as another run)."""
    from nemo_amd import solver
    cfg = solver.make_config(N=N, precision=precision, **cfg_kw)
    with solver.Solver(cfg) as s:
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


def cpu_baseline(n: int, N: int, seed: int, steps: int, cfg_kw):
    """FP64 oracle on the host cores: one or more full steps of the same
    64^3 workload, after one untimed warm-up stage sequence on a small mesh."""
    threads = min(16, os.cpu_count() or 1)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle import oracle as O
    kw = dict(cfg_kw)
    ocfg = O.make_cfg(N=N, volume=kw["volume"], surface=kw["surface"], indicator=kw["indicator"],
                      timedisc=kw["timedisc"], repair=kw["repair"], cfl=kw["cfl"])
    m, u = workload(n, N, seed)
    t0 = time.perf_counter()
    rc, _ = O.run(ocfg, m.nbr, m.delta, u, steps)
    wall = time.perf_counter() - t0
    dof = m.nelem * (N + 1) ** 3
    return {"value": dof * NSTAGES * steps / wall, "unit": "DOF-updates/s", "cores": O.num_threads(),
            "kind": "port",
            "sample": f"{n}^3 elements, N={N}, {steps} full LSRK 5-4 step(s) of the same KH workload, "
                      f"{wall:.1f} s wall, FP64 C oracle, OpenMP"}


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
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is design-only", file=sys.stderr)
        sys.exit(2)

    from nemo_amd import _lib
    N = 3
    cfg_kw = dict(volume=_lib.VOL_EC, surface=_lib.SURF_HLLC, indicator=_lib.IND_PP, timedisc=_lib.TD_LSRK54,
                  repair=1, cfl=0.5, alpha_max=0.5, alpha_min=0.001)
    m, u0 = workload(args.elems, N, args.seed)
    wall, launches, kernel_ms, ok, u64, dts = gpu_run(m, u0, N, args.steps, args.warmup, cfg_kw, precision=0)
    dof = m.nelem * (N + 1) ** 3
    value = dof * NSTAGES * args.steps / wall

    # algorithmic bytes of the stage launches: stage 1 = 3 arrays, 2..5 = 4 arrays
    bytes_per_step = dof * NVARS * 8 * (3 + 4 * 4)
    total_bytes = bytes_per_step * args.steps
    stage_launches = NSTAGES * args.steps
    assert launches == stage_launches, (launches, stage_launches)
    avg_launch_s = kernel_ms / 1e3 / launches
    achieved = total_bytes / (kernel_ms / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "stage_kernel<4,4,EC,HLLC,double>", "avg_launch_us": round(avg_launch_s * 1e6, 2),
                "bytes_per_launch_avg": int(bytes_per_step / NSTAGES)}

    chosen = None
    if not args.no_chosen:
        # chosen form: FP32 operator (FP64 state, RK accumulation, CFL, repair)
        wall32, l32, ms32, ok32, _, _ = gpu_run(m, u0, N, args.steps, args.warmup, cfg_kw, precision=1)
        u32 = gpu_replay(m, u0, N, dts, cfg_kw, precision=1)
        err = l1_rel(u32[:, 0], u64[:, 0], N)
        v32 = dof * NSTAGES * args.steps / wall32
        chosen = {"form": "FP32 operator + FP64 RK/CFL/repair, Persson-Peraire indicator, per-stage launches",
                  "value": v32, "unit": "DOF-updates/s", "ms_per_step": wall32 / args.steps * 1e3,
                  "l1_density_err_vs_fp64": err, "l1_bar": 1e-6, "meets_l1_bar": err <= 1e-6,
                  "speedup_vs_plain": v32 / value,
                  "stage_kernel_avg_us": round(ms32 / 1e3 / l32 * 1e6, 2),
                  "hbm_frac": round(total_bytes / (ms32 / 1e3) / 1e9 / HBM_PEAK_GBS, 4), "status_ok": ok32}

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
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.elems, N, args.seed, args.cpu_steps, cfg_kw)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
