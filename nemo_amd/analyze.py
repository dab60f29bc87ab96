"""Analysis layer: conserved integrals over time and errors against exact data.

The reference's hook_analyze (setups/euler/2d/_common_/setup_analyze_mod.f90:
34-164; the 3-D convergence test's setups/euler/3d/convergence-test/
setup_analyze_mod.f90:40-151) sums quadrature-weighted conserved quantities
over all quads every dt_analyze and appends a line to analysis.dat; the
convergence test adds L2/Linf errors against the exact solution.  Here the
integrals come from the device (nemo_gpu_analyze, no state download) and the
error norms from a downloaded state on the host.
"""
from __future__ import annotations

import os

import numpy as np

from . import operators


def quadrature_volumes(mesh, N: int) -> np.ndarray:
    """[nelem, M^3] LGL quadrature volume of every node (node = i + M j + M^2 k)."""
    _, w = operators.lgl(N)
    M = N + 1
    w3 = (w[None, None, :] * w[None, :, None] * w[:, None, None]).reshape(-1)  # k slowest, i fastest
    return np.prod(np.asarray(mesh.delta), axis=1)[:, None] * w3[None, :] / 8.0


def error_norms(mesh, N: int, u: np.ndarray, u_exact: np.ndarray) -> dict:
    """Per-variable L1, L2 (volume-weighted, normalised by the volume) and
    Linf norms of u - u_exact, u of shape [nelem, V, M^3]."""
    vol = quadrature_volumes(mesh, N)
    d = np.asarray(u) - np.asarray(u_exact)
    V = d.shape[1]
    total = vol.sum()
    out = {}
    for h in range(V):
        e = d[:, h, :]
        out[h] = {"l1": float(np.sum(np.abs(e) * vol) / total),
                  "l2": float(np.sqrt(np.sum(e * e * vol) / total)),
                  "linf": float(np.abs(e).max())}
    return out


class AnalysisLog:
    """analysis.dat writer: one line per call of ``record`` with simulation
    time, time step, step count and the device integrals, the columns of the
    reference's resBuf (time, dt, steps, mass, momenta, total energy, ...)."""

    COLUMNS = ["time", "dt", "steps", "dens", "momx", "momy", "momz", "ener", "ekin", "eint", "rho_min", "p_min"]

    def __init__(self, path: str = "analysis.dat", every: float = 0.0):
        self.path = path
        self.every = every
        self.next_time = 0.0
        self.rows = []
        d = os.path.dirname(os.path.abspath(path))
        os.makedirs(d, exist_ok=True)
        with open(path, "w") as f:
            f.write("# " + " ".join(self.COLUMNS) + "\n")

    def record(self, solver, time: float, dt: float, steps: int, force: bool = False) -> dict | None:
        """hook_analyze_500: integrals of the solver's current state, if
        ``time`` has reached the next analysis time (or ``force``)."""
        if not force and time < self.next_time:
            return None
        self.next_time = time + self.every
        a = solver.analyze()
        row = [time, dt, steps] + [a.get(k, 0.0) for k in self.COLUMNS[3:]]
        self.rows.append(row)
        with open(self.path, "a") as f:
            f.write(" ".join(f"{v:.16e}" for v in row) + "\n")
        return a
