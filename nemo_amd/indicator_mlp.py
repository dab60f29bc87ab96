"""Learned troubled-cell indicator: initialisation, training data from the
method's own Persson-Peraire runs, GPU training, and the parameter file.

The network (4 -> 64 -> 64 -> 64 -> 1, GELU, sigmoid; layout in
include/nemo_gpu.h) predicts the DG/FV blending factor from the log modal
energy fractions of rho and p.  Training data are (features, alpha_PP) pairs
that nemo_gpu_indicator_features dumps during plain-indicator runs of Sedov
blasts with seeded centre jitter and of the Kelvin-Helmholtz shear layer;
training runs on the GPU (nemo_gpu_mlp_train: the library's own forward,
backward and Adam kernels).

Run as a script to regenerate the committed parameters:
    python -m nemo_amd.indicator_mlp --out nemo_amd/data/mlp_indicator.npz
"""
from __future__ import annotations

import argparse
import os
import time

import numpy as np

from . import _lib, mesh as meshmod, setups, solver

H, IN = 64, 4
NPARAMS = _lib.MLP_NPARAMS
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PARAMS = os.path.join(HERE, "data", "mlp_indicator.npz")


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights for the GELU layers, zero biases; the output bias
    starts at logit(0.01) so that an untrained network blends little."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2.0 / IN), (H, IN))
    W2 = rng.normal(0, np.sqrt(2.0 / H), (H, H))
    W3 = rng.normal(0, np.sqrt(2.0 / H), (H, H))
    w4 = rng.normal(0, np.sqrt(1.0 / H), H)
    b4 = np.log(0.01 / 0.99)
    p = np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), W3.ravel(), np.zeros(H), w4, [b4]])
    assert p.size == NPARAMS
    return p.astype(np.float32)


def sedov(mesh: meshmod.Mesh, N: int, gamma: float = 1.4, seed: int = 1234, energy: float = 1.0,
          radius: float = 0.02, jitter: float = 0.05, p0: float = 1.0e-5) -> np.ndarray:
    """Sedov blast: energy E deposited as pressure in a Gaussian of width
    `radius` around the cube centre shifted by a seeded jitter of up to
    `jitter` per coordinate; ambient rho = 1, p = p0, at rest.
    p = p0 + (gamma-1) E exp(-r^2/(2 radius^2)) / ((2 pi)^(3/2) radius^3)
    (the Gaussian integrates to one on the periodic cube)."""
    from . import operators
    rng = np.random.default_rng(seed)
    c = 0.5 + jitter * (2.0 * rng.random(3) - 1.0)
    xi, _ = operators.lgl(N)
    X = mesh.node_coords(xi)
    d = X - c[None, :, None]
    d = d - np.round(d)  # periodic distance
    r2 = (d ** 2).sum(1)
    g = np.exp(-r2 / (2.0 * radius ** 2)) / ((2.0 * np.pi) ** 1.5 * radius ** 3)
    p = p0 + (gamma - 1.0) * energy * g
    rho = np.ones_like(p)
    z = np.zeros_like(p)
    return np.ascontiguousarray(setups.prim2cons(rho, z, z, z, p, gamma))


SCHEME = dict(N=3, volume=_lib.VOL_EC, surface=_lib.SURF_HLLC, timedisc=_lib.TD_LSRK54, repair=1, cfl=0.5,
              alpha_max=0.5, alpha_min=0.001)


def collect(runs, steps: int, every: int, precision: int = 1, log=print):
    """Training pairs from plain-indicator runs: `runs` is a list of
    (mesh, u0) pairs; every `every` steps of each run the features and PP
    factors of all elements are recorded.  Returns (features, alpha)."""
    feats, alphas = [], []
    for m, u0 in runs:
        cfg = solver.make_config(indicator=_lib.IND_PP, precision=precision, **SCHEME)
        with solver.Solver(cfg) as s:
            s.upload_mesh(m)
            s.set_state(u0)
            for it in range(0, steps, every):
                f, a = s.indicator_features()
                feats.append(f)
                alphas.append(a.astype(np.float32))
                s.run(every)
        log(f"  collected {sum(len(a) for a in alphas)} pairs")
    return np.concatenate(feats), np.concatenate(alphas)


def default_runs(n_sedov: int = 24, sedov_runs: int = 6, n_kh: int = 24, seed: int = 1234,
                 amr_runs: int = 2, amr_seed: int = 2001, amr_levels=(3, 6)):
    """Uniform Sedov blasts with jittered centres, one KH layer, and Sedov
    blasts on the adaptive forests (levels 3-6) the indicator builds for
    their initial data (blast centres other than the uniform runs')."""
    from . import amr, amr_run
    runs = []
    m = meshmod.uniform_periodic(n_sedov)
    for k in range(sedov_runs):
        runs.append((m, sedov(m, 3, seed=seed + k)))
    mk = meshmod.uniform_periodic(n_kh)
    runs.append((mk, setups.kelvin_helmholtz(mk, 3, seed=seed)))
    cfg = solver.make_config(indicator=_lib.IND_PP, precision=0, **SCHEME)
    for k in range(amr_runs):
        ic = lambda f, sd=amr_seed + k: sedov(f, 3, seed=sd)
        f0 = amr_run.initial_forest(lambda: solver.Solver(cfg),
                                    amr.Forest.uniform(amr_levels[0], minlevel=amr_levels[0],
                                                       maxlevel=amr_levels[1]), ic)
        runs.append((f0, ic(f0)))
    return runs


def train(feat: np.ndarray, alpha: np.ndarray, epochs: int = 30, batch: int = 4096, lr: float = 2e-3,
          seed: int = 7, params: np.ndarray | None = None):
    """Adam training on the GPU; returns (params, loss per epoch)."""
    cfg = solver.make_config(indicator=_lib.IND_PP, **SCHEME)
    with solver.Solver(cfg) as s:
        s.mlp_set_params(init_params(seed) if params is None else params)
        hist = s.mlp_train(feat, alpha, epochs=epochs, batch=batch, lr=lr, seed=seed)
        return s.mlp_get_params(), hist


def load_params(path: str = DEFAULT_PARAMS) -> np.ndarray:
    with np.load(path, allow_pickle=False) as z:
        return np.asarray(z["params"], dtype=np.float32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_PARAMS)
    ap.add_argument("--steps", type=int, default=240)
    ap.add_argument("--every", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=30)
    args = ap.parse_args()
    t0 = time.time()
    feat, alpha = collect(default_runs(), args.steps, args.every)
    print(f"{len(alpha)} pairs, {np.mean(alpha > 0):.3f} blended, {time.time() - t0:.1f} s", flush=True)
    t1 = time.time()
    params, hist = train(feat, alpha, epochs=args.epochs)
    print("loss per epoch:", " ".join(f"{h:.3e}" for h in hist), f"({time.time() - t1:.1f} s)", flush=True)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    np.savez_compressed(args.out, params=params, loss=hist, pairs=np.int64(len(alpha)),
                        blended_fraction=np.float64(np.mean(alpha > 0)))
    print("wrote", args.out)


if __name__ == "__main__":
    main()
