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
    cols, big, small = forest.neighbour_columns()
    flag = np.asarray(flag, dtype=bool)
    ext = np.append(flag, False)     # the slot of a face without a conforming neighbour
    out = flag.copy()
    for c in cols:
        out |= ext[c]
    if len(big):
        # a leaf may sit on several mortars: any flagged side marks it
        out[big[flag[small].any(axis=1)]] = True
        out[small[flag[big]].ravel()] = True
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
    # the flags of that pass as Forest.adapt sees them (refine below the
    # finest level, coarsen): its outcome is a function of the forest and
    # these, so equal flags on the same forest leave it unchanged again --
    # the masks move from pass to pass where a factor hovers at the
    # threshold while the flags after the neighbour layer stay put
    _steady_flags: tuple = field(default=None, repr=False)
    steady_shortcuts: bool = True   # False: run every pass in full (tests)
    mask_shortcuts: int = 0         # passes decided by the mask test
    flag_shortcuts: int = 0         # passes decided by the flag test

    def adapt(self) -> np.ndarray | None:
        t0 = time.perf_counter()
        alpha = self.solver.indicator()
        masks = (alpha > self.refine_above, alpha <= self.coarsen_below)
        if self.steady_shortcuts and self._steady is not None and \
                all(np.array_equal(m, k) for m, k in zip(masks, self._steady)):
            self.mask_shortcuts += 1
            self.remesh_seconds += time.perf_counter() - t0
            return None
        refine, coarsen = adapt_flags(self.forest, alpha, self.refine_above, self.coarsen_below)
        flags = (refine & (self.forest.level < self.forest.maxlevel), coarsen)
        if self.steady_shortcuts and self._steady_flags is not None and \
                all(np.array_equal(m, k) for m, k in zip(flags, self._steady_flags)):
            self._steady = masks
            self.flag_shortcuts += 1
            self.remesh_seconds += time.perf_counter() - t0
            return None
        new, units = self.forest.adapt(refine, coarsen)
        if units is not None:
            self.fallbacks += self.solver.remesh(new, units)
            self.forest = new
            self.remeshes += 1
            self._steady = None
            self._steady_flags = None
        else:
            self._steady = masks
            self._steady_flags = flags
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
