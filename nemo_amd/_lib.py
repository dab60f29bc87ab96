"""Loader for the in-tree HIP library nemo_amd/libnemo_gpu.so (C-ABI of
include/nemo_gpu.h).  There is no fallback: if the library is missing the
import fails loudly."""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnemo_gpu.so")

NEMO_OK = 0
NEMO_ERR_ARG = 1
NEMO_ERR_HIP = 2
NEMO_ERR_STATE = 3
NEMO_ERR_NOT_READY = 4

EQ_EULER, EQ_GLM_MHD = 0, 1
BC_TRANSMISSIVE, BC_SLIP_WALL = 0, 1
VOL_WEAK, VOL_CENTRAL, VOL_EC = 0, 1, 2
SURF_RUSANOV, SURF_HLLC, SURF_HLLD = 0, 1, 2
IND_NONE, IND_PP, IND_MLP = 0, 1, 2
MLP_NPARAMS = 8705
TD_LSRK54, TD_SSPRK54 = 0, 1


class Config(ctypes.Structure):
    """nemo_gpu_config of include/nemo_gpu.h."""
    _fields_ = [
        ("n_poly", ctypes.c_int32),
        ("equations", ctypes.c_int32),
        ("volume", ctypes.c_int32),
        ("surface", ctypes.c_int32),
        ("indicator", ctypes.c_int32),
        ("timedisc", ctypes.c_int32),
        ("repair", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("gamma", ctypes.c_double),
        ("alpha_max", ctypes.c_double),
        ("alpha_min", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("mu0", ctypes.c_double),
    ]


class NemoGpuError(RuntimeError):
    def __init__(self, code: int, where: str, message: str):
        super().__init__(f"{where} failed with code {code}: {message}")
        self.code = code


_lib = None

_P = ctypes.POINTER(ctypes.c_double)
_PF = ctypes.POINTER(ctypes.c_float)
_PI32 = ctypes.POINTER(ctypes.c_int32)
_PI64 = ctypes.POINTER(ctypes.c_int64)
_VP = ctypes.c_void_p

SIGNATURES = {
    "nemo_gpu_create": [ctypes.POINTER(Config), ctypes.POINTER(_VP)],
    "nemo_gpu_destroy": [_VP],
    "nemo_gpu_upload_mesh": [_VP, ctypes.c_int64, _PI32, _P],
    "nemo_gpu_upload_mesh_amr": [_VP, ctypes.c_int64, _PI32, _P, ctypes.c_int64, _PI32],
    "nemo_gpu_upload_mesh_bc": [_VP, ctypes.c_int64, _PI32, _P, ctypes.c_int64, _PI32, ctypes.c_int64, _PI32],
    "nemo_gpu_remesh": [_VP, ctypes.c_int64, _PI32, _P, ctypes.c_int64, _PI32, ctypes.c_int64, _PI32,
                        ctypes.POINTER(ctypes.c_int64)],
    "nemo_gpu_upload_state": [_VP, _P],
    "nemo_gpu_download_state": [_VP, _P],
    "nemo_gpu_rhs": [_VP, _P],
    "nemo_gpu_indicator": [_VP, _P],
    "nemo_gpu_analyze": [_VP, _P],
    "nemo_gpu_analyze_noalloc": [_VP, _VP],
    "nemo_gpu_calc_timestep": [_VP, _P, _P, _P],
    "nemo_gpu_set_glm": [_VP, ctypes.c_double, ctypes.c_double],
    "nemo_gpu_get_glm": [_VP, _P, _P],
    "nemo_gpu_stage": [_VP, ctypes.c_int32, ctypes.c_double],
    "nemo_gpu_step": [_VP, ctypes.c_double],
    "nemo_gpu_run": [_VP, ctypes.c_int32, _P],
    "nemo_gpu_synchronize": [_VP],
    "nemo_gpu_get_status": [_VP, _PI32],
    "nemo_gpu_indicator_features": [_VP, _PF, _P],
    "nemo_gpu_mlp_set_params": [_VP, _PF],
    "nemo_gpu_mlp_get_params": [_VP, _PF],
    "nemo_gpu_mlp_set_policy": [_VP, ctypes.c_double, ctypes.c_double],
    "nemo_gpu_mlp_stats": [_VP, _PI64, _PI64, _PI64],
    "nemo_gpu_mlp_forward": [_VP, ctypes.c_int64, _PF, _PF],
    "nemo_gpu_mlp_grad": [_VP, ctypes.c_int64, _PF, _PF, _PF, _PF],
    "nemo_gpu_mlp_train": [_VP, ctypes.c_int64, _PF, _PF, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                           ctypes.c_uint64, _PF],
    "nemo_gpu_batch_create": [ctypes.POINTER(Config), ctypes.c_int32, _PI32, _PI64, ctypes.POINTER(_VP)],
    "nemo_gpu_batch_destroy": [_VP],
    "nemo_gpu_batch_upload": [_VP, ctypes.c_int32, _PI32, _P, _P],
    "nemo_gpu_batch_download": [_VP, ctypes.c_int32, _P],
    "nemo_gpu_batch_run": [_VP, ctypes.c_int32, ctypes.c_int32, _P],
    "nemo_gpu_batch_set_launch": [_VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32],
    "nemo_gpu_batch_autotune": [_VP, ctypes.c_int32, _PI32, _PI32, _PI32, _PI32, _P],
    "nemo_gpu_batch_tune_table": [_VP, ctypes.c_int32, _PI32, _PI32, _P],
    "nemo_gpu_set_timing": [_VP, ctypes.c_int32],
    "nemo_gpu_copy_calibration": [_VP, ctypes.c_int64],
    "nemo_gpu_check_arithmetic": [_VP, ctypes.c_int64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int64),
                                  ctypes.POINTER(ctypes.c_int64)],
    "nemo_gpu_read_timing": [_VP, _PI64, _P],
    "nemo_gpu_operators": [ctypes.c_int32, _P, _P, _P, _P],
    "nemo_gpu_error_message": [ctypes.c_char_p, ctypes.c_int32],
}


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"nemo_amd: HIP library {LIB_PATH} is not built; run __graft_entry__.build() "
                "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, argtypes in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argtypes
            fn.restype = ctypes.c_int
        _lib = L
    return _lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().nemo_gpu_error_message(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc: int, where: str) -> None:
    if rc != NEMO_OK:
        raise NemoGpuError(rc, where, last_error())


def dptr(a):
    return a.ctypes.data_as(_P)


def fptr(a):
    return a.ctypes.data_as(_PF)


def iptr(a):
    return a.ctypes.data_as(_PI32)
