"""Host-side mirror of nemo's kernel / timedisc / driver interfaces over the
C-ABI of libnemo_gpu.so.

The reference drives one time step as (driver_mod.f90:62-150):

    hook_calc_timestep  -> CFL reduction            (kernel_timestep_mod.f90:19-63)
    hook_evolve_500     -> for stage in 1..nstages:  (timedisc_evolve_mod.f90:20-84)
                             rktimestep = dt*dt_coeffs(stage)
                             hook_pre_rk_timestep   (halo exchange, face fluxes)
                             hook_mid_rk_timestep   (rkstep_cb per quad)
                             status check

The class below keeps those names and their meaning; each hook is one C-ABI
call.  Errors follow the reference's convention: a failed call sets
``runtime.stat`` like ``raise_error`` (utils/status_mod.f90:7-21) and raises
``NemoGpuError``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from ._lib import Config, NemoGpuError, check, dptr, fptr, iptr, lib
from .mesh import Mesh

# timedisc_const_mod::dt_coeffs (low-storage/4-5:29 = b_rk, ssp/4-5:36)
LSRK54_DT_COEFFS = (0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                    2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0)
SSPRK54_DT_COEFFS = (1.0, 0.39175222700392, 0.58607968896780, 0.47454236302687, 0.93501063100924)
NSTAGES = 5


@dataclass
class Status:
    """status_t of runtime_mod.f90:55-61."""
    ok: bool = True
    errcode: int = 0
    message: str = ""


@dataclass
class Runtime:
    """runtime_t of runtime_mod.f90:63-84 (the fields the hot path uses)."""
    simtime: float = 0.0
    timestep: float = 0.0
    maxtimestep: float = 0.0
    rktimestep: float = 0.0
    rksimtime: float = 0.0
    rkstage: int = 0
    timesteps: int = 0
    nextdump: float = 0.0
    chkptcount: int = 0
    stat: Status = field(default_factory=Status)


def make_config(N: int = 3, volume: int = _lib.VOL_WEAK, surface: int = _lib.SURF_RUSANOV,
                indicator: int = _lib.IND_NONE, timedisc: int = _lib.TD_SSPRK54, repair: int = 0,
                gamma: float = 1.4, alpha_max: float = 0.5, alpha_min: float = 0.001, cfl: float = 0.5,
                precision: int = 0, equations: int = _lib.EQ_EULER, mu0: float = 1.0) -> Config:
    return Config(N, equations, volume, surface, indicator, timedisc, repair, precision, gamma,
                  alpha_max, alpha_min, cfl, mu0)


class Solver:
    """One device-resident solver (an opaque nemo_gpu handle)."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self.M = cfg.n_poly + 1
        self.M3 = self.M ** 3
        self.nvars = 9 if cfg.equations == _lib.EQ_GLM_MHD else 5
        self.runtime = Runtime()
        self._h = ctypes.c_void_p()
        self._call("nemo_gpu_create", ctypes.byref(cfg), ctypes.byref(self._h))
        self.nelem = 0

    # -- plumbing ----------------------------------------------------------
    def _call(self, name: str, *args) -> None:
        rc = getattr(lib(), name)(*args)
        if rc != _lib.NEMO_OK:
            msg = _lib.last_error()
            self.runtime.stat.ok = False
            self.runtime.stat.errcode = -1
            self.runtime.stat.message += msg + "\n"
            raise NemoGpuError(rc, name, msg)

    def close(self) -> None:
        if self._h:
            lib().nemo_gpu_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- mesh and state (mesh_alloc / payload upload) -----------------------
    def upload_mesh(self, mesh: Mesh) -> None:
        """Connectivity of a uniform mesh (periodic, or with the boundary faces
        of nemo_amd.mesh.with_boundaries) or of an adaptive forest
        (nemo_amd.amr.Forest)."""
        nbr = np.ascontiguousarray(mesh.nbr, dtype=np.int32)
        delta = np.ascontiguousarray(mesh.delta, dtype=np.float64)
        mortars = np.ascontiguousarray(getattr(mesh, "mortars", np.zeros((0, 6))), dtype=np.int32)
        bf = getattr(mesh, "bfaces", None)
        bfaces = np.ascontiguousarray(bf if bf is not None else np.zeros((0, 3)), dtype=np.int32)
        self._call("nemo_gpu_upload_mesh_bc", self._h, ctypes.c_int64(mesh.nelem), iptr(nbr), dptr(delta),
                   ctypes.c_int64(mortars.shape[0]), iptr(mortars), ctypes.c_int64(bfaces.shape[0]), iptr(bfaces))
        self.nelem = mesh.nelem

    def remesh(self, forest, units: np.ndarray) -> int:
        """Move the device state onto a new forest (kernel_amr_refine/coarsen on
        the GPU); units from nemo_amd.amr.replacement_units.  Returns the
        number of units that fell back to the element mean."""
        nbr = np.ascontiguousarray(forest.nbr, dtype=np.int32)
        delta = np.ascontiguousarray(forest.delta, dtype=np.float64)
        mortars = np.ascontiguousarray(forest.mortars, dtype=np.int32)
        units = np.ascontiguousarray(units, dtype=np.int32)
        fb = ctypes.c_int64(0)
        self._call("nemo_gpu_remesh", self._h, ctypes.c_int64(forest.nelem), iptr(nbr), dptr(delta),
                   ctypes.c_int64(mortars.shape[0]), iptr(mortars), ctypes.c_int64(units.shape[0]), iptr(units),
                   ctypes.byref(fb))
        self.nelem = forest.nelem
        return int(fb.value)

    def set_state(self, u: np.ndarray) -> None:
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.shape != (self.nelem, self.nvars, self.M3):
            raise ValueError(f"state of shape {u.shape}, expected {(self.nelem, self.nvars, self.M3)}")
        self._call("nemo_gpu_upload_state", self._h, dptr(u))

    def get_state(self) -> np.ndarray:
        u = np.empty((self.nelem, self.nvars, self.M3))
        self._call("nemo_gpu_download_state", self._h, dptr(u))
        return u

    # -- kernel interface ----------------------------------------------------
    def kernel_calc_flux(self) -> np.ndarray:
        """du/dt of the current state (kernel_calc_flux after the face pass)."""
        r = np.empty((self.nelem, self.nvars, self.M3))
        self._call("nemo_gpu_rhs", self._h, dptr(r))
        return r

    def indicator(self) -> np.ndarray:
        a = np.empty(self.nelem)
        self._call("nemo_gpu_indicator", self._h, dptr(a))
        return a

    def analyze(self) -> dict:
        """Volume integrals and extrema of the current state (nemo_gpu_analyze)."""
        out = np.empty(16)
        self._call("nemo_gpu_analyze", self._h, dptr(out))
        names = ["dens", "momx", "momy", "momz", "ener", "magx", "magy", "magz", "glmp"][:self.nvars]
        res = {n: float(out[k]) for k, n in enumerate(names)}
        res.update(ekin=float(out[9]), eint=float(out[10]), volume=float(out[12]), rho_min=float(out[13]),
                   p_min=float(out[14]), v_max=float(out[15]))
        if self.nvars == 9:
            res["emag"] = float(out[11])
        return res

    def hook_calc_timestep(self):
        """CFL time step; sets runtime.maxtimestep (kernel_timestep_mod.f90:19-63)."""
        dt, lm, vm = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._call("nemo_gpu_calc_timestep", self._h, ctypes.byref(dt), ctypes.byref(lm), ctypes.byref(vm))
        self.runtime.maxtimestep = dt.value
        return dt.value, lm.value, vm.value

    def set_glm(self, ch: float, alpha: float) -> None:
        """equations_share_mod::glm_ch / glm_alpha (GLM-MHD)."""
        self._call("nemo_gpu_set_glm", self._h, ctypes.c_double(ch), ctypes.c_double(alpha))

    def get_glm(self):
        ch, alpha = ctypes.c_double(), ctypes.c_double()
        self._call("nemo_gpu_get_glm", self._h, ctypes.byref(ch), ctypes.byref(alpha))
        return ch.value, alpha.value

    # -- learned indicator ------------------------------------------------------
    def indicator_features(self):
        """Features [nelem, 4] (float32) and Persson-Peraire factors of the
        current state: the training data of the learned indicator."""
        f = np.empty((self.nelem, 4), dtype=np.float32)
        a = np.empty(self.nelem)
        self._call("nemo_gpu_indicator_features", self._h, fptr(f), dptr(a))
        return f, a

    def mlp_set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != _lib.MLP_NPARAMS:
            raise ValueError(f"{p.size} network parameters, expected {_lib.MLP_NPARAMS}")
        self._call("nemo_gpu_mlp_set_params", self._h, fptr(p))

    def mlp_get_params(self) -> np.ndarray:
        p = np.empty(_lib.MLP_NPARAMS, dtype=np.float32)
        self._call("nemo_gpu_mlp_get_params", self._h, fptr(p))
        return p

    def mlp_set_policy(self, sample_fraction: float = 0.05, tolerance: float = 0.05) -> None:
        self._call("nemo_gpu_mlp_set_policy", self._h, ctypes.c_double(sample_fraction), ctypes.c_double(tolerance))

    def mlp_stats(self):
        j, o, f = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._call("nemo_gpu_mlp_stats", self._h, ctypes.byref(j), ctypes.byref(o), ctypes.byref(f))
        return j.value, o.value, f.value

    def mlp_forward(self, feat: np.ndarray) -> np.ndarray:
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        y = np.empty(feat.shape[0], dtype=np.float32)
        self._call("nemo_gpu_mlp_forward", self._h, ctypes.c_int64(feat.shape[0]), fptr(feat), fptr(y))
        return y

    def mlp_grad(self, feat: np.ndarray, target: np.ndarray):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        g = np.empty(_lib.MLP_NPARAMS, dtype=np.float32)
        self._call("nemo_gpu_mlp_grad", self._h, ctypes.c_int64(feat.shape[0]), fptr(feat), fptr(t), fptr(loss), fptr(g))
        return float(loss[0]), g

    def mlp_train(self, feat: np.ndarray, target: np.ndarray, epochs: int = 10, batch: int = 4096,
                  lr: float = 1e-3, seed: int = 1) -> np.ndarray:
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        hist = np.zeros(epochs, dtype=np.float32)
        self._call("nemo_gpu_mlp_train", self._h, ctypes.c_int64(feat.shape[0]), fptr(feat), fptr(t),
                   ctypes.c_int32(epochs), ctypes.c_int32(batch), ctypes.c_float(lr), ctypes.c_uint64(seed), fptr(hist))
        return hist

    # -- timedisc interface ----------------------------------------------------
    def dt_coeffs(self):
        return LSRK54_DT_COEFFS if self.cfg.timedisc == _lib.TD_LSRK54 else SSPRK54_DT_COEFFS

    def hook_evolve(self) -> None:
        """All stages of one step with runtime.timestep (timedisc_evolve_mod.f90:20-84)."""
        rt = self.runtime
        for stage in range(1, NSTAGES + 1):
            rt.rkstage = stage
            rt.rktimestep = rt.timestep * self.dt_coeffs()[stage - 1]
            rt.rksimtime = rt.simtime + rt.rktimestep
            self._call("nemo_gpu_stage", self._h, ctypes.c_int32(stage), ctypes.c_double(rt.timestep))
        self._call("nemo_gpu_synchronize", self._h)

    def stage(self, stage: int, dt: float) -> None:
        """One RK stage (1..5) with time step dt, asynchronous."""
        self._call("nemo_gpu_stage", self._h, ctypes.c_int32(stage), ctypes.c_double(dt))

    def step(self, dt: float) -> None:
        self._call("nemo_gpu_step", self._h, ctypes.c_double(dt))

    def run(self, nsteps: int, record_dt: bool = False) -> Optional[np.ndarray]:
        """nsteps CFL-limited steps entirely on the device."""
        hist = np.zeros(max(nsteps, 1)) if record_dt else None
        self._call("nemo_gpu_run", self._h, ctypes.c_int32(nsteps), dptr(hist) if record_dt else None)
        self.runtime.timesteps += nsteps
        return hist[:nsteps] if record_dt else None

    def synchronize(self) -> None:
        self._call("nemo_gpu_synchronize", self._h)

    # -- timing ------------------------------------------------------------------
    def set_timing(self, enable: bool) -> None:
        self._call("nemo_gpu_set_timing", self._h, ctypes.c_int32(1 if enable else 0))

    def read_timing(self):
        n, ms = ctypes.c_int64(), ctypes.c_double()
        self._call("nemo_gpu_read_timing", self._h, ctypes.byref(n), ctypes.byref(ms))
        return n.value, ms.value

    def check_arithmetic(self, n: int, seed: int = 1):
        """(div_rn, rcp_nr) results whose bits differ from IEEE, over n
        generated operand pairs (nemo_gpu_check_arithmetic)."""
        d, r = ctypes.c_int64(), ctypes.c_int64()
        self._call("nemo_gpu_check_arithmetic", self._h, ctypes.c_int64(n), ctypes.c_uint64(seed), ctypes.byref(d),
                   ctypes.byref(r))
        return d.value, r.value

    # -- driver loop -------------------------------------------------------------
    def mainloop(self, stop_time: float, max_steps: int = 1 << 30, *, dt_checkpoint: float = 0.0,
                 checkpoint_dir: str = None, mesh=None, analysis=None, log=None) -> Runtime:
        """Time loop of driver_mod.f90:62-150: CFL step, analysis
        (hook_analyze, an AnalysisLog), a FLASH checkpoint of ``mesh`` into
        ``checkpoint_dir`` whenever the time reaches the next dump time
        (every ``dt_checkpoint``, and at the end), then the step with
        dt = min(maxtimestep, dt_checkpoint, |nextdump - simtime|)."""
        rt = self.runtime
        tol = 4.0e-16
        chk = checkpoint_dir is not None and dt_checkpoint > 0.0
        rt.nextdump = rt.simtime if chk else float("inf")
        while rt.stat.ok and rt.timesteps < max_steps:
            self.hook_calc_timestep()
            if analysis is not None:
                analysis.record(self, rt.simtime, rt.maxtimestep, rt.timesteps)
            if chk and abs(rt.simtime - rt.nextdump) < 10 * tol * max(1.0, abs(rt.simtime)):
                self.hook_checkpoint(checkpoint_dir, mesh)
                if log:
                    log(f"[CHECKPOINT] {rt.chkptcount - 1:6d} {rt.timesteps:8d} {rt.simtime:12.3e} {rt.maxtimestep:12.4e}")
                rt.nextdump = min(stop_time, rt.simtime + dt_checkpoint)
            if abs(rt.simtime - stop_time) < 100 * tol:
                break
            rt.timestep = min(rt.maxtimestep, abs(stop_time - rt.simtime))
            if chk:
                rt.timestep = min(rt.timestep, dt_checkpoint, abs(rt.nextdump - rt.simtime))
            self.hook_evolve()
            rt.simtime += rt.timestep
            rt.timesteps += 1
        return rt

    def hook_checkpoint(self, directory: str, mesh) -> str:
        """FLASH checkpoint of the current state (hook_checkpoint_100..500,
        checkpoint/flash/checkpoint_mod.f90:11-364) as
        <directory>/chkpt_NNNNNN.f5; the blending factor goes to /blnd."""
        from . import flash
        rt = self.runtime
        u = self.get_state()
        blend = self.indicator() if self.cfg.indicator != _lib.IND_NONE else None
        path = flash.write(flash.checkpoint_path(directory, rt.chkptcount), mesh, self.M - 1, u, time=rt.simtime,
                           dt=rt.maxtimestep, nstep=rt.timesteps, number=rt.chkptcount, gamma=self.cfg.gamma,
                           cfl=self.cfg.cfl, blend=blend)
        rt.chkptcount += 1
        return path


class Batch:
    """Batched ensemble of independent runs (nemo_gpu_batch_*): each run has
    its own mesh, state, polynomial degree and CFL time step."""

    def __init__(self, cfg: Config, meshes, n_polys):
        self.cfg = cfg
        self.meshes = list(meshes)
        self.n_polys = [int(n) for n in n_polys]
        self.nruns = len(self.meshes)
        npoly = np.array(self.n_polys, dtype=np.int32)
        nel = np.array([m.nelem for m in self.meshes], dtype=np.int64)
        self._h = ctypes.c_void_p()
        self._call("nemo_gpu_batch_create", ctypes.byref(cfg), ctypes.c_int32(self.nruns), iptr(npoly),
                   nel.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(self._h))

    def _call(self, name, *args):
        rc = getattr(lib(), name)(*args)
        if rc != _lib.NEMO_OK:
            raise NemoGpuError(rc, name, _lib.last_error())

    def close(self):
        if self._h:
            lib().nemo_gpu_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, run: int, u: np.ndarray) -> None:
        m = self.meshes[run]
        M3 = (self.n_polys[run] + 1) ** 3
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.shape != (m.nelem, 5, M3):
            raise ValueError(f"state of run {run} has shape {u.shape}, expected {(m.nelem, 5, M3)}")
        nbr = np.ascontiguousarray(m.nbr, dtype=np.int32)
        delta = np.ascontiguousarray(m.delta, dtype=np.float64)
        self._call("nemo_gpu_batch_upload", self._h, ctypes.c_int32(run), iptr(nbr), dptr(delta), dptr(u))

    def download(self, run: int) -> np.ndarray:
        m = self.meshes[run]
        u = np.empty((m.nelem, 5, (self.n_polys[run] + 1) ** 3))
        self._call("nemo_gpu_batch_download", self._h, ctypes.c_int32(run), dptr(u))
        return u

    def run(self, nsteps: int, use_graph: bool = True, record_dt: bool = False):
        hist = np.zeros((max(nsteps, 1), self.nruns)) if record_dt else None
        self._call("nemo_gpu_batch_run", self._h, ctypes.c_int32(nsteps), ctypes.c_int32(1 if use_graph else 0),
                   dptr(hist) if record_dt else None)
        return hist[:nsteps] if record_dt else None

    def set_launch(self, grid_per_cu: int, xcd_map: bool = True, one_launch: bool = True, order: int = 0) -> None:
        """Launch geometry (nemo_gpu_batch_set_launch): one launch per stage
        over all degrees or one per degree, persistent grid per CU (0: one
        workgroup per block), XCD map, order of the degree groups."""
        self._call("nemo_gpu_batch_set_launch", self._h, ctypes.c_int32(grid_per_cu), ctypes.c_int32(int(xcd_map)),
                   ctypes.c_int32(int(one_launch)), ctypes.c_int32(order))

    def autotune(self, steps: int = 2) -> dict:
        """Time every candidate launch geometry, keep the fastest; returns it
        with its ms per step."""
        g, x, one, order = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        ms = ctypes.c_double()
        self._call("nemo_gpu_batch_autotune", self._h, ctypes.c_int32(steps), ctypes.byref(g), ctypes.byref(x),
                   ctypes.byref(one), ctypes.byref(order), ctypes.byref(ms))
        return {"grid_per_cu": g.value, "xcd_map": bool(x.value), "one_launch": bool(one.value),
                "order": order.value, "ms_per_step": ms.value}

    def tune_table(self) -> list:
        """Every candidate of the last autotuning with its ms per step."""
        n = ctypes.c_int32()
        self._call("nemo_gpu_batch_tune_table", self._h, ctypes.c_int32(0), ctypes.byref(n), None, None)
        geo = np.zeros((max(n.value, 1), 4), dtype=np.int32)
        ms = np.zeros(max(n.value, 1))
        self._call("nemo_gpu_batch_tune_table", self._h, ctypes.c_int32(n.value), ctypes.byref(n), iptr(geo),
                   dptr(ms))
        return [{"one_launch": bool(g[0]), "order": int(g[1]), "grid_per_cu": int(g[2]), "xcd_map": bool(g[3]),
                 "ms_per_step": float(t)} for g, t in zip(geo[:n.value], ms[:n.value])]
