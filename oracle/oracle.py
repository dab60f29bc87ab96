"""ctypes view of the FP64 CPU oracle (oracle/libnemo_oracle.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module; the product
package ``nemo_amd`` never does.  It is the checker, never the thing
measured or shipped.

Arrays follow the product's C-ABI layout: ``u[elem][var][node]`` with
``node = i + M*j + M*M*k`` (x fastest, ``M = N+1``).
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libnemo_oracle.so")
_REF_A = os.path.join(_HERE, "_ref", "libnemo_ref_a.so")
_REF_B = os.path.join(_HERE, "_ref", "libnemo_ref_b.so")
_REF_C = os.path.join(_HERE, "_ref", "libnemo_ref_c.so")

EQ_EULER, EQ_MHD = 0, 1
VOL_WEAK, VOL_CENTRAL, VOL_EC = 0, 1, 2
SURF_RUSANOV, SURF_HLLC, SURF_HLLD, SURF_EC_TEST = 0, 1, 2, 3
TD_LSRK54, TD_SSPRK54 = 0, 1
IND_NONE, IND_PP = 0, 1


class Cfg(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int32),
        ("eq", ctypes.c_int32),
        ("volume", ctypes.c_int32),
        ("surface", ctypes.c_int32),
        ("indicator", ctypes.c_int32),
        ("timedisc", ctypes.c_int32),
        ("repair", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("gamma", ctypes.c_double),
        ("alpha_max", ctypes.c_double),
        ("alpha_min", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("mu0", ctypes.c_double),
        ("glm_ch", ctypes.c_double),
        ("glm_alpha", ctypes.c_double),
        ("nmortar", ctypes.c_int64),
        ("mortars", ctypes.c_void_p),
        ("nbound", ctypes.c_int64),
        ("bfaces", ctypes.c_void_p),
    ]


def make_cfg(N=3, eq=EQ_EULER, volume=VOL_WEAK, surface=SURF_RUSANOV,
             indicator=IND_NONE, timedisc=TD_SSPRK54, repair=0, gamma=1.4,
             alpha_max=0.5, alpha_min=0.001, cfl=0.5, mu0=1.0, glm_ch=0.0,
             glm_alpha=0.0) -> Cfg:
    return Cfg(N, eq, volume, surface, indicator, timedisc, repair, 0, gamma,
               alpha_max, alpha_min, cfl, mu0, glm_ch, glm_alpha, 0, None, 0, None)


def with_mortars(cfg: Cfg, mortars) -> Cfg:
    """Copy of cfg that carries a mortar table (kept alive with the copy)."""
    c = Cfg.from_buffer_copy(cfg)
    arr = np.ascontiguousarray(mortars if mortars is not None else np.zeros((0, 6)), dtype=np.int32)
    c._mortars = arr
    c.nmortar = arr.shape[0]
    c.mortars = arr.ctypes.data if arr.shape[0] else None
    c._bfaces = getattr(cfg, "_bfaces", None)
    return c


BC_TRANSMISSIVE, BC_SLIP_WALL = 0, 1


def with_boundaries(cfg: Cfg, bfaces) -> Cfg:
    """Copy of cfg that carries a boundary-face table [nbound][3] = (element,
    face, kind) (kept alive with the copy)."""
    c = Cfg.from_buffer_copy(cfg)
    c._mortars = getattr(cfg, "_mortars", None)
    arr = np.ascontiguousarray(bfaces if bfaces is not None else np.zeros((0, 3)), dtype=np.int32)
    c._bfaces = arr
    c.nbound = arr.shape[0]
    c.bfaces = arr.ctypes.data if arr.shape[0] else None
    return c


def slip_wall_pressure(cfg: Cfg, rho: float, v: float, p: float) -> float:
    return lib().orc_slip_wall_pressure(ctypes.byref(cfg), rho, v, p)


_P = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int32)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(_PI)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError(f"oracle library not built: {_LIB} (run make -C oracle)")
        L = ctypes.CDLL(_LIB)
        cp = ctypes.POINTER(Cfg)
        L.orc_pressure.restype = ctypes.c_double
        L.orc_lambda_max.restype = ctypes.c_double
        L.orc_logmean.restype = ctypes.c_double
        L.orc_logmean.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_entropy.restype = ctypes.c_double
        L.orc_calc_timestep.restype = ctypes.c_double
        L.orc_calc_timestep.argtypes = [cp, ctypes.c_int64, _P, _P, _P, _P]
        L.orc_rhs.argtypes = [cp, ctypes.c_int64, _PI, _P, _P, _P, _P]
        L.orc_indicator.argtypes = [cp, ctypes.c_int64, _P, _P, _P]
        L.orc_step.argtypes = [cp, ctypes.c_int64, _PI, _P, _P, _P, _P, ctypes.c_double]
        L.orc_stage.argtypes = [cp, ctypes.c_int64, _PI, _P, _P, _P, _P, ctypes.c_int, ctypes.c_double]
        L.orc_run.argtypes = [cp, ctypes.c_int64, _PI, _P, _P, ctypes.c_int, _P]
        L.orc_legendre.argtypes = [ctypes.c_int, ctypes.c_double, _P, _P]
        L.orc_mortar_ops.argtypes = [ctypes.c_int, _P, _P]
        L.orc_mortar_fluxes.argtypes = [cp, ctypes.c_int64, _P, _P]
        L.orc_boundary_fluxes.argtypes = [cp, ctypes.c_int64, _P, _P]
        L.orc_slip_wall_pressure.restype = ctypes.c_double
        L.orc_slip_wall_pressure.argtypes = [cp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_mhd_pressure.restype = ctypes.c_double
        L.orc_mhd_fastspeed.restype = ctypes.c_double
        L.orc_phys_lambda_max.restype = ctypes.c_double
        L.orc_phys_isvalid.restype = ctypes.c_int
        _lib = L
    return _lib


# ---- operators -----------------------------------------------------------

def lgl(N: int):
    x = np.zeros(N + 1)
    w = np.zeros(N + 1)
    lib().orc_lgl_nodes_weights(ctypes.c_int(N), _ptr(x), _ptr(w))
    return x, w


def diff_matrix(N: int, x: np.ndarray) -> np.ndarray:
    D = np.zeros((N + 1, N + 1))
    lib().orc_diff_matrix(ctypes.c_int(N), _ptr(np.ascontiguousarray(x)), _ptr(D))
    return D


def vandermonde(N: int, x: np.ndarray):
    V = np.zeros((N + 1, N + 1))
    Vi = np.zeros((N + 1, N + 1))
    lib().orc_vandermonde(ctypes.c_int(N), _ptr(np.ascontiguousarray(x)), _ptr(V), _ptr(Vi))
    return V, Vi


def legendre(n: int, x: float):
    L = np.zeros(1)
    dL = np.zeros(1)
    lib().orc_legendre(n, x, _ptr(L), _ptr(dL))
    return float(L[0]), float(dL[0])


# ---- pointwise -----------------------------------------------------------

def pressure(cfg: Cfg, u) -> float:
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().orc_pressure(ctypes.byref(cfg), _ptr(u))


def flux(cfg: Cfg, d: int, u) -> np.ndarray:
    u = np.ascontiguousarray(u, dtype=np.float64)
    f = np.zeros_like(u)
    lib().orc_flux(ctypes.byref(cfg), ctypes.c_int(d), _ptr(u), _ptr(f))
    return f


def lambda_max(cfg: Cfg, d: int, u) -> float:
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().orc_lambda_max(ctypes.byref(cfg), ctypes.c_int(d), _ptr(u))


def riemann(cfg: Cfg, kind: int, d: int, uL, uR) -> np.ndarray:
    uL = np.ascontiguousarray(uL, dtype=np.float64)
    uR = np.ascontiguousarray(uR, dtype=np.float64)
    f = np.zeros_like(uL)
    lib().orc_riemann(ctypes.byref(cfg), ctypes.c_int(kind), ctypes.c_int(d), _ptr(uL), _ptr(uR), _ptr(f))
    return f


def two_point(cfg: Cfg, kind: int, d: int, u, v) -> np.ndarray:
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    f = np.zeros_like(u)
    lib().orc_two_point(ctypes.byref(cfg), ctypes.c_int(kind), ctypes.c_int(d), _ptr(u), _ptr(v), _ptr(f))
    return f


def logmean(a: float, b: float) -> float:
    return lib().orc_logmean(a, b)


def cons2evec(cfg: Cfg, u) -> np.ndarray:
    u = np.ascontiguousarray(u, dtype=np.float64)
    w = np.zeros_like(u)
    lib().orc_cons2evec(ctypes.byref(cfg), _ptr(u), _ptr(w))
    return w


def entropy(cfg: Cfg, u) -> float:
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().orc_entropy(ctypes.byref(cfg), _ptr(u))


def isvalid(cfg: Cfg, u) -> bool:
    u = np.ascontiguousarray(u, dtype=np.float64)
    return bool(lib().orc_isvalid(ctypes.byref(cfg), _ptr(u)))


# ---- pointwise, dispatched on cfg.eq (Euler: 5 vars, GLM-MHD: 9) ----------

def _vec(u):
    return np.ascontiguousarray(u, dtype=np.float64)


def mhd_pressure(cfg: Cfg, u) -> float:
    return lib().orc_mhd_pressure(ctypes.byref(cfg), _ptr(_vec(u)))


def mhd_fastspeed(cfg: Cfg, u) -> float:
    return lib().orc_mhd_fastspeed(ctypes.byref(cfg), _ptr(_vec(u)))


def phys_flux(cfg: Cfg, d: int, u) -> np.ndarray:
    u = _vec(u)
    f = np.zeros_like(u)
    lib().orc_phys_flux(ctypes.byref(cfg), ctypes.c_int(d), _ptr(u), _ptr(f))
    return f


def phys_riemann(cfg: Cfg, kind: int, d: int, uL, uR) -> np.ndarray:
    uL, uR = _vec(uL), _vec(uR)
    f = np.zeros_like(uL)
    lib().orc_phys_riemann(ctypes.byref(cfg), ctypes.c_int(kind), ctypes.c_int(d), _ptr(uL), _ptr(uR), _ptr(f))
    return f


def phys_two_point(cfg: Cfg, kind: int, d: int, u, v) -> np.ndarray:
    u, v = _vec(u), _vec(v)
    f = np.zeros_like(u)
    lib().orc_phys_two_point(ctypes.byref(cfg), ctypes.c_int(kind), ctypes.c_int(d), _ptr(u), _ptr(v), _ptr(f))
    return f


def phys_nc(cfg: Cfg, d: int, uI, uL, uR) -> np.ndarray:
    uI, uL, uR = _vec(uI), _vec(uL), _vec(uR)
    f = np.zeros_like(uI)
    lib().orc_phys_nc(ctypes.byref(cfg), ctypes.c_int(d), _ptr(uI), _ptr(uL), _ptr(uR), _ptr(f))
    return f


def phys_lambda_max(cfg: Cfg, d: int, u) -> float:
    return lib().orc_phys_lambda_max(ctypes.byref(cfg), ctypes.c_int(d), _ptr(_vec(u)))


def phys_isvalid(cfg: Cfg, u) -> bool:
    return bool(lib().orc_phys_isvalid(ctypes.byref(cfg), _ptr(_vec(u))))


# ---- element operators ---------------------------------------------------

def nvars(cfg: Cfg) -> int:
    return 9 if cfg.eq == EQ_MHD else 5


def indicator(cfg: Cfg, u: np.ndarray):
    nelem = u.shape[0]
    a = np.zeros(nelem)
    e = np.zeros(nelem)
    lib().orc_indicator(ctypes.byref(cfg), nelem, _ptr(np.ascontiguousarray(u)), _ptr(a), _ptr(e))
    return a, e


def rhs(cfg: Cfg, nbr: np.ndarray, delta: np.ndarray, u: np.ndarray,
        alpha: Optional[np.ndarray] = None) -> np.ndarray:
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros_like(u)
    ap = _ptr(np.ascontiguousarray(alpha, dtype=np.float64)) if alpha is not None else None
    lib().orc_rhs(ctypes.byref(cfg), u.shape[0], _iptr(np.ascontiguousarray(nbr, dtype=np.int32)),
                  _ptr(np.ascontiguousarray(delta, dtype=np.float64)), _ptr(u), ap, _ptr(out))
    return out


def mortar_ops(N: int):
    """(P, R), each [2, M, M]: 2:1 interpolation and its projection."""
    M = N + 1
    P = np.zeros((2, M, M))
    R = np.zeros((2, M, M))
    lib().orc_mortar_ops(ctypes.c_int(N), _ptr(P), _ptr(R))
    return P, R


def mortar_fluxes(cfg: Cfg, u: np.ndarray) -> np.ndarray:
    """Face fluxes of every mortar slot, [5*nmortar, nvars, M*M]."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    M = cfg.N + 1
    out = np.zeros((5 * cfg.nmortar, nvars(cfg), M * M))
    lib().orc_mortar_fluxes(ctypes.byref(cfg), u.shape[0], _ptr(u), _ptr(out))
    return out


UNIT_COPY, UNIT_REFINE, UNIT_COARSEN = 0, 1, 2


def _valid_euler(gamma: float, v: np.ndarray) -> bool:
    """rho > 0 and p > 0 at every node of v [..., 5, M^3] (equations_mod.f90:337-344)."""
    rho = v[..., 0, :]
    p = (gamma - 1.0) * (v[..., 4, :] - 0.5 * (v[..., 1, :] ** 2 + v[..., 2, :] ** 2 + v[..., 3, :] ** 2) / rho)
    return bool(np.all(rho > 0.0) and np.all(p > 0.0))


def _valid_mhd(gamma: float, mu0: float, v: np.ndarray) -> bool:
    """rho > 0 and gas pressure > 0 at every node of v [..., 9, M^3]
    (equations/mhd/equations_mod.f90:223-234: B^2 and psi^2 over 2 mu0 taken
    out of the energy)."""
    rho = v[..., 0, :]
    mag = v[..., 5, :] ** 2 + v[..., 6, :] ** 2 + v[..., 7, :] ** 2 + v[..., 8, :] ** 2
    p = (gamma - 1.0) * (v[..., 4, :] - 0.5 / rho * (v[..., 1, :] ** 2 + v[..., 2, :] ** 2 + v[..., 3, :] ** 2)
                         - 0.5 / mu0 * mag)
    return bool(np.all(rho > 0.0) and np.all(p > 0.0))


def remap(N: int, gamma: float, u_old: np.ndarray, units: np.ndarray, nnew: int, mu0: Optional[float] = None):
    """State transfer of one adaptation: kernel_amr_refine / kernel_amr_coarsen
    (kernel/mcfv/3d/kernel_amr_mod.f90:12-70) as tensor products of the 2:1
    operators per direction, child c = x + 2y + 4z; a unit whose result has a
    nonpositive density or pressure takes the element mean instead (the
    positivity fallback of setup_amr_replace_refine/coarsen,
    setups/euler/3d/sedov/setup_amr_mod.f90:76-130, reduced to its last,
    first-order step).  A state of 9 variables is GLM-MHD (gas pressure with
    magnetic constant mu0).  Returns (u_new [nnew, V, M^3], number of
    fallbacks)."""
    M = N + 1
    V = u_old.shape[1]
    if V == 9:
        valid = lambda v: _valid_mhd(gamma, 1.0 if mu0 is None else mu0, v)
    else:
        valid = lambda v: _valid_euler(gamma, v)
    P, R = mortar_ops(N)
    _, w = lgl(N)
    w3 = (w[:, None, None] * w[None, :, None] * w[None, None, :]).reshape(-1) / 8.0  # (k, j, i) order
    nold = u_old.shape[0]
    uo = u_old.reshape(nold, V, M, M, M)       # node axes (k, j, i)
    out = np.zeros((nnew, V, M, M, M))
    fallbacks = 0
    for kind, io, jn in units:
        if kind == UNIT_COPY:
            out[jn] = uo[io]
        elif kind == UNIT_REFINE:
            ch = np.stack([np.einsum("ka,jb,ic,vabc->vkji", P[c >> 2], P[(c >> 1) & 1], P[c & 1], uo[io])
                           for c in range(8)])
            if not valid(ch.reshape(8, V, -1)):
                mean = uo[io].reshape(V, -1) @ w3
                ch[:] = mean[None, :, None, None, None]
                fallbacks += 1
            out[jn:jn + 8] = ch
        else:
            par = sum(np.einsum("ka,jb,ic,vabc->vkji", R[c >> 2], R[(c >> 1) & 1], R[c & 1], uo[io + c])
                      for c in range(8))
            if not valid(par.reshape(V, -1)):
                mean = par.reshape(V, -1) @ w3
                par = np.broadcast_to(mean[:, None, None, None], par.shape).copy()
                fallbacks += 1
            out[jn] = par
    return out.reshape(nnew, V, M ** 3), fallbacks


def calc_timestep(cfg: Cfg, delta: np.ndarray, u: np.ndarray):
    lm = np.zeros(1)
    vm = np.zeros(1)
    dt = lib().orc_calc_timestep(ctypes.byref(cfg), u.shape[0],
                                 _ptr(np.ascontiguousarray(delta, dtype=np.float64)),
                                 _ptr(np.ascontiguousarray(u)), _ptr(lm), _ptr(vm))
    return dt, float(lm[0]), float(vm[0])


def repair(cfg: Cfg, ue: np.ndarray) -> bool:
    ue = np.ascontiguousarray(ue, dtype=np.float64)
    ok = lib().orc_repair(ctypes.byref(cfg), _ptr(ue))
    return bool(ok), ue


def step(cfg: Cfg, nbr, delta, u: np.ndarray, r0: np.ndarray, r1: np.ndarray, dt: float) -> int:
    """One time step in place on u (registers r0/r1 updated in place)."""
    for a in (u, r0, r1):
        assert a.flags.c_contiguous and a.dtype == np.float64
    return lib().orc_step(ctypes.byref(cfg), u.shape[0], _iptr(np.ascontiguousarray(nbr, dtype=np.int32)),
                          _ptr(np.ascontiguousarray(delta, dtype=np.float64)), _ptr(u), _ptr(r0), _ptr(r1),
                          ctypes.c_double(dt))


def stage(cfg: Cfg, nbr, delta, u: np.ndarray, r0: np.ndarray, r1: np.ndarray, s: int, dt: float) -> int:
    """One RK stage in place on u (registers r0/r1 updated in place)."""
    for a in (u, r0, r1):
        assert a.flags.c_contiguous and a.dtype == np.float64
    return lib().orc_stage(ctypes.byref(cfg), u.shape[0], _iptr(np.ascontiguousarray(nbr, dtype=np.int32)),
                           _ptr(np.ascontiguousarray(delta, dtype=np.float64)), _ptr(u), _ptr(r0), _ptr(r1),
                           int(s), ctypes.c_double(dt))


def run(cfg: Cfg, nbr, delta, u: np.ndarray, nsteps: int):
    """nsteps CFL-limited steps in place on u; returns (rc, dt history)."""
    assert u.flags.c_contiguous and u.dtype == np.float64
    hist = np.zeros(nsteps)
    rc = lib().orc_run(ctypes.byref(cfg), u.shape[0], _iptr(np.ascontiguousarray(nbr, dtype=np.int32)),
                       _ptr(np.ascontiguousarray(delta, dtype=np.float64)), _ptr(u), nsteps, _ptr(hist))
    return rc, hist


def lsrk54_coeffs():
    a, b, c = np.zeros(5), np.zeros(5), np.zeros(5)
    lib().orc_lsrk54_coeffs(_ptr(a), _ptr(b), _ptr(c))
    return a, b, c


def ssprk54_coeffs():
    k = np.zeros(23)
    lib().orc_ssprk54_coeffs(_ptr(k))
    return k


def num_threads() -> int:
    return lib().orc_num_threads()


def set_num_threads(n: int) -> None:
    """OpenMP thread count of the oracle's following parallel regions."""
    lib().orc_set_num_threads(int(n))


# ---- the reference's own Fortran leaf modules (oracle/_ref) ---------------

class Ref:
    """The reference's Fortran modules compiled in place (oracle/_ref)."""

    def __init__(self):
        if not (os.path.exists(_REF_A) and os.path.exists(_REF_B)):
            raise FileNotFoundError("oracle/_ref not built (make -C oracle ref)")
        self.a = ctypes.CDLL(_REF_A)
        self.b = ctypes.CDLL(_REF_B)
        self.a.ref_pressure.restype = ctypes.c_double
        self.a.ref_lambda_max.restype = ctypes.c_double
        self.a.ref_lambda_max.argtypes = [ctypes.c_int, _P]
        self.a.ref_flux.argtypes = [ctypes.c_int, _P, _P]
        self.a.ref_hllc.argtypes = [ctypes.c_int, _P, _P, _P]
        self.a.ref_ec.argtypes = [ctypes.c_int, _P, _P, _P]
        self.a.ref_lgl.argtypes = [ctypes.c_int, _P, _P]
        self.a.ref_diff_matrix.argtypes = [ctypes.c_int, _P, _P]
        self.a.ref_legendre.argtypes = [ctypes.c_int, ctypes.c_double, _P, _P]
        self.a.ref_vandermonde.argtypes = [ctypes.c_int, _P, _P, _P]
        self.a.ref_minmod3.restype = ctypes.c_double
        self.a.ref_minmod3.argtypes = [ctypes.c_double] * 3
        self.a.ref_isvalid.restype = ctypes.c_int
        self.b.ref_rusanov.argtypes = [ctypes.c_int, _P, _P, _P]
        self.b.ref_central.argtypes = [ctypes.c_int, _P, _P, _P]

    @staticmethod
    def _v(u):
        return np.ascontiguousarray(u, dtype=np.float64)

    def pressure(self, u):
        return self.a.ref_pressure(_ptr(self._v(u)))

    def flux(self, d, u):
        f = np.zeros(5)
        self.a.ref_flux(d, _ptr(self._v(u)), _ptr(f))
        return f

    def lambda_max(self, d, u):
        return self.a.ref_lambda_max(d, _ptr(self._v(u)))

    def cons2evec(self, u):
        w = np.zeros(5)
        self.a.ref_cons2evec(_ptr(self._v(u)), _ptr(w))
        return w

    def isvalid(self, u):
        return bool(self.a.ref_isvalid(_ptr(self._v(u))))

    def hllc(self, d, uL, uR):
        f = np.zeros(5)
        self.a.ref_hllc(d, _ptr(self._v(uL)), _ptr(self._v(uR)), _ptr(f))
        return f

    def rusanov(self, d, uL, uR):
        f = np.zeros(5)
        self.b.ref_rusanov(d, _ptr(self._v(uL)), _ptr(self._v(uR)), _ptr(f))
        return f

    def ec(self, d, u, v):
        f = np.zeros(5)
        self.a.ref_ec(d, _ptr(self._v(u)), _ptr(self._v(v)), _ptr(f))
        return f

    def central(self, d, u, v):
        f = np.zeros(5)
        self.b.ref_central(d, _ptr(self._v(u)), _ptr(self._v(v)), _ptr(f))
        return f

    def lsrk_coeffs(self):
        a, b, c = np.zeros(5), np.zeros(5), np.zeros(5)
        self.a.ref_lsrk_coeffs(_ptr(a), _ptr(b), _ptr(c))
        return a, b, c

    def ssp_coeffs(self):
        k = np.zeros(23)
        self.b.ref_ssp_coeffs(_ptr(k))
        return k

    def lgl(self, N):
        x, w = np.zeros(N + 1), np.zeros(N + 1)
        self.a.ref_lgl(N, _ptr(x), _ptr(w))
        return x, w

    def diff_matrix(self, N, x):
        D = np.zeros((N + 1, N + 1), order="F")
        self.a.ref_diff_matrix(N, _ptr(self._v(x)), D.ctypes.data_as(_P))
        return np.ascontiguousarray(D)

    def legendre(self, n, x):
        L, dL = np.zeros(1), np.zeros(1)
        self.a.ref_legendre(n, x, _ptr(L), _ptr(dL))
        return float(L[0]), float(dL[0])

    def vandermonde(self, x):
        m = len(x)
        V = np.zeros((m, m), order="F")
        Vi = np.zeros((m, m), order="F")
        self.a.ref_vandermonde(m, _ptr(self._v(x)), V.ctypes.data_as(_P), Vi.ctypes.data_as(_P))
        return np.ascontiguousarray(V), np.ascontiguousarray(Vi)

    def minmod3(self, a, b, c):
        return self.a.ref_minmod3(a, b, c)


def ref_available() -> bool:
    return os.path.exists(_REF_A) and os.path.exists(_REF_B)


class RefMHD:
    """The reference's GLM-MHD modules compiled in place (oracle/_ref, library
    C: equations/mhd/{equations_mod, riemann/rusanov, non_conservative_flux_mod}
    with the Orszag-Tang setup_config_mod, kappa = 5/3, mu0 = 1)."""

    def __init__(self):
        if not os.path.exists(_REF_C):
            raise FileNotFoundError("oracle/_ref not built (make -C oracle ref)")
        c = self.c = ctypes.CDLL(_REF_C)
        c.ref_mhd_set_glm.argtypes = [ctypes.c_double]
        c.ref_mhd_pressure.restype = ctypes.c_double
        c.ref_mhd_pressure.argtypes = [_P]
        c.ref_mhd_flux.argtypes = [ctypes.c_int, _P, _P]
        c.ref_mhd_lambda_max.restype = ctypes.c_double
        c.ref_mhd_lambda_max.argtypes = [ctypes.c_int, _P]
        c.ref_mhd_fastspeed.restype = ctypes.c_double
        c.ref_mhd_fastspeed.argtypes = [_P]
        c.ref_mhd_isvalid.restype = ctypes.c_int
        c.ref_mhd_isvalid.argtypes = [_P]
        c.ref_mhd_rusanov.argtypes = [ctypes.c_int, _P, _P, _P]
        c.ref_mhd_nc.argtypes = [ctypes.c_int, _P, _P, _P, _P]

    def set_glm(self, ch):
        self.c.ref_mhd_set_glm(float(ch))

    def pressure(self, u):
        return self.c.ref_mhd_pressure(_ptr(_vec(u)))

    def flux(self, d, u):
        f = np.zeros(9)
        self.c.ref_mhd_flux(d, _ptr(_vec(u)), _ptr(f))
        return f

    def lambda_max(self, d, u):
        return self.c.ref_mhd_lambda_max(d, _ptr(_vec(u)))

    def fastspeed(self, u):
        return self.c.ref_mhd_fastspeed(_ptr(_vec(u)))

    def isvalid(self, u):
        return bool(self.c.ref_mhd_isvalid(_ptr(_vec(u))))

    def rusanov(self, d, uL, uR):
        f = np.zeros(9)
        self.c.ref_mhd_rusanov(d, _ptr(_vec(uL)), _ptr(_vec(uR)), _ptr(f))
        return f

    def nc(self, d, uI, uL, uR):
        f = np.zeros(9)
        self.c.ref_mhd_nc(d, _ptr(_vec(uI)), _ptr(_vec(uL)), _ptr(_vec(uR)), _ptr(f))
        return f


def ref_mhd_available() -> bool:
    return os.path.exists(_REF_C)
