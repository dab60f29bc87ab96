"""Allocation-free analysis entry point (nemo_gpu_analyze_noalloc) against
nemo_gpu_analyze through the C-ABI: the same kernels and reduction order,
so the sums must agree bit for bit, on two calls in a row (the second on a
smaller mesh, so that stale partial rows from the first call would show),
and on a mesh of another degree.  The device output buffer is allocated by
the test with hipMalloc (the caller's memory)."""
import ctypes
import os

import numpy as np
import pytest

from nemo_amd import _lib, mesh as meshmod, setups, solver

pytestmark = pytest.mark.gpu

NSUMS = 16
HIP_D2H = 2


def _hip():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    hip = ctypes.CDLL(os.path.join(rocm, "lib", "libamdhip64.so"))
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


def _noalloc(s, hip, dev):
    s._call("nemo_gpu_analyze_noalloc", s._h, dev)
    s._call("nemo_gpu_synchronize", s._h)
    out = np.empty(NSUMS)
    assert hip.hipMemcpy(out.ctypes.data, dev, NSUMS * 8, HIP_D2H) == 0
    return out


def _host(s):
    out = np.empty(NSUMS)
    s._call("nemo_gpu_analyze", s._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def test_noalloc_analyze_matches_allocating_entry_point():
    hip = _hip()
    dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev), NSUMS * 8) == 0
    try:
        cfg = solver.make_config(N=3, volume=_lib.VOL_EC, surface=_lib.SURF_HLLC, indicator=_lib.IND_PP)
        with solver.Solver(cfg) as s:
            for n, seed in ((8, 3), (4, 7)):
                m = meshmod.uniform_periodic(n)
                u = setups.kelvin_helmholtz(m, 3, seed=seed)
                s.upload_mesh(m)
                s.set_state(u)
                print(f"noalloc analyze: {n}^3 elements, N=3, seed {seed}", flush=True)
                a = _noalloc(s, hip, dev)
                b = _host(s)
                assert np.array_equal(a, b), (a, b)
                assert abs(b[12] - 1.0) < 1e-13   # volume of the unit cube
        cfg2 = solver.make_config(N=2, volume=_lib.VOL_EC, surface=_lib.SURF_HLLC, indicator=_lib.IND_PP)
        with solver.Solver(cfg2) as s:
            m = meshmod.uniform_periodic(5)
            s.upload_mesh(m)
            s.set_state(setups.kelvin_helmholtz(m, 2, seed=11))
            print("noalloc analyze: 5^3 elements, N=2", flush=True)
            a = _noalloc(s, hip, dev)
            assert np.array_equal(a, _host(s))
    finally:
        hip.hipFree(dev)
