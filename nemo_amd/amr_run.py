"""Indicator-driven mesh adaptation around the device solver.

The reference adapts at the start of every ``amr_each_ntimesteps``-th step
(hook_pre_timestep_300, setups/euler/3d/sedov/setup_amr_mod.f90:7-19) through
mesh_adapt (mesh/p4est/mesh_utils_mod.f90:879-1009) with per-setup probe
callbacks.  Here the probes read the blending factor of the troubled-cell
indicator the operator itself uses (Persson-Peraire or the learned one):
an element is refined where the factor is above ``refine_above`` (and so are
its face neighbours, one element of buffer for the front to move into), and
a family of 8 is coarsened where all factors are at most ``coarsen_below``.
The forest and the replacement units are built on the host (nemo_amd.amr);
the state moves on the device (nemo_gpu_remesh).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import amr


def face_neighbours_any(forest: amr.Forest, flag: np.ndarray) -> np.ndarray:
    """flag | (flag of any face neighbour), across conforming and 2:1 faces."""
    nbr, mortars = forest.connectivity()
    flag = np.asarray(flag, dtype=bool)
    n = flag.shape[0]
    conf = nbr >= 0
    out = flag | (conf & flag[np.where(conf, nbr, 0)]).any(axis=1)
    if len(mortars):
        big, small = mortars[:, 0], mortars[:, 2:]
        # a leaf may sit on several mortars: count the flagged ones
        out |= np.bincount(big, weights=flag[small].any(axis=1), minlength=n) > 0
        out |= np.bincount(small.ravel(), weights=np.repeat(flag[big], 4), minlength=n) > 0
    return out


def adapt_flags(forest: amr.Forest, alpha: np.ndarray, refine_above: float = 0.0, coarsen_below: float = 0.0):
    refine = face_neighbours_any(forest, alpha > refine_above)
    coarsen = (alpha <= coarsen_below) & ~refine
    return refine, coarsen


def initial_forest(solver_factory, forest: amr.Forest, ic, refine_above: float = 0.0, max_rounds: int = 8):
    """Refine the initial mesh where the indicator fires on the initial data
    (evaluated afresh on every new mesh) until the mesh stops changing."""
    for _ in range(max_rounds):
        u = ic(forest)
        with solver_factory() as s:
            s.upload_mesh(forest)
            s.set_state(u)
            alpha = s.indicator()
        refine = face_neighbours_any(forest, alpha > refine_above)
        new, units = forest.adapt(refine, np.zeros(forest.nelem, bool))
        if units is None:
            break
        forest = new
    return forest


@dataclass
class Segment:
    forest: amr.Forest          # mesh of the segment
    units: np.ndarray | None    # replacement units that produced it (None: unchanged)
    dts: np.ndarray             # time steps taken on it


@dataclass
class AdaptiveRun:
    solver: object
    forest: amr.Forest
    every: int = 8
    refine_above: float = 0.0
    coarsen_below: float = 0.0
    history: list = field(default_factory=list)
    remeshes: int = 0
    fallbacks: int = 0
    dof_updates: int = 0        # element-node updates summed over steps (times stages)
    elements: list = field(default_factory=list)
    remesh_seconds: float = 0.0
    # the indicator masks of the last pass that left the forest unchanged:
    # the flags, and so the outcome, are a function of the forest and these
    # masks, so an equal pair on the same forest is steady again
    _steady: tuple = field(default=None, repr=False)

    def adapt(self) -> np.ndarray | None:
        t0 = time.perf_counter()
        alpha = self.solver.indicator()
        masks = (alpha > self.refine_above, alpha <= self.coarsen_below)
        if self._steady is not None and all(np.array_equal(m, k) for m, k in zip(masks, self._steady)):
            self.remesh_seconds += time.perf_counter() - t0
            return None
        refine, coarsen = adapt_flags(self.forest, alpha, self.refine_above, self.coarsen_below)
        new, units = self.forest.adapt(refine, coarsen)
        if units is not None:
            self.fallbacks += self.solver.remesh(new, units)
            self.forest = new
            self.remeshes += 1
            self._steady = None
        else:
            self._steady = masks
        self.remesh_seconds += time.perf_counter() - t0
        return units

    def advance(self, nsteps: int, record: bool = False) -> None:
        """nsteps steps, adapting before every ``every``-th (as the reference)."""
        done = 0
        M3 = self.solver.M3
        while done < nsteps:
            units = self.adapt()
            k = min(self.every, nsteps - done)
            dts = self.solver.run(k, record_dt=record)
            if record:
                self.history.append(Segment(self.forest, units, dts))
            self.dof_updates += k * 5 * self.forest.nelem * M3
            self.elements.append(self.forest.nelem)
            done += k


def replay(solver, history) -> None:
    """Repeat a recorded run's meshes and time steps on another solver (the
    same data, e.g. another precision or indicator): remesh where the record
    did, then step with the recorded dt."""
    for seg in history:
        if seg.units is not None:
            solver.remesh(seg.forest, seg.units)
        for dt in seg.dts:
            solver.step(float(dt))
    solver.synchronize()
