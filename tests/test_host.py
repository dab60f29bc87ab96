"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point include/nemo_gpu.h declares, argument checking, the mesh
upload format and the initial data."""
import ctypes
import os
import re

import numpy as np

from nemo_amd import _lib, mesh as meshmod, setups, solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "nemo_gpu.h")).read()
    return sorted(set(re.findall(r"\bint\s+(nemo_gpu_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"


def test_flash_library_exports_its_header():
    """libnemo_flash.so exports what include/nemo_flash.h declares, and the
    ctypes mirrors of its structs have the C layout."""
    import pytest
    from nemo_amd import flash
    if not os.path.exists(flash.LIB_PATH):
        pytest.skip("libnemo_flash.so not built (no HDF5 C library)")
    txt = open(os.path.join(ROOT, "include", "nemo_flash.h")).read()
    names = sorted(set(re.findall(r"\bint\s+(nemo_flash_\w+)\s*\(", txt)))
    assert names == ["nemo_flash_error_message", "nemo_flash_write"]
    L = flash.lib()
    for n in names:
        assert hasattr(L, n), n
    assert ctypes.sizeof(flash.Meta) == 8 * 4 + 14 * 8
    assert ctypes.sizeof(flash.Tree) == 2 * 8 + 8 * 8


def test_config_struct_layout():
    assert ctypes.sizeof(_lib.Config) == 8 * 4 + 5 * 8


def test_create_rejects_bad_arguments_without_device():
    h = ctypes.c_void_p()
    cfg = solver.make_config(N=7)
    rc = _lib.lib().nemo_gpu_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == _lib.NEMO_ERR_ARG
    assert "n_poly" in _lib.last_error()
    cfg = solver.make_config(N=3, gamma=1.0)
    assert _lib.lib().nemo_gpu_create(ctypes.byref(cfg), ctypes.byref(h)) == _lib.NEMO_ERR_ARG
    # combinations that are not built are refused before any device call
    bad = [
        (dict(equations=_lib.EQ_GLM_MHD, volume=_lib.VOL_WEAK, surface=_lib.SURF_HLLD), "flux differencing"),
        (dict(equations=_lib.EQ_GLM_MHD, volume=_lib.VOL_EC, surface=_lib.SURF_HLLC), "GLM-MHD"),
        (dict(equations=_lib.EQ_GLM_MHD, volume=_lib.VOL_EC, surface=_lib.SURF_HLLD, mu0=0.0), "mu0"),
        (dict(volume=_lib.VOL_EC, surface=_lib.SURF_HLLD), "Euler"),
        (dict(equations=_lib.EQ_GLM_MHD, volume=_lib.VOL_EC, surface=_lib.SURF_HLLD, indicator=_lib.IND_MLP),
         "learned indicator"),
        (dict(volume=_lib.VOL_WEAK, surface=_lib.SURF_HLLC, indicator=_lib.IND_MLP), "learned indicator"),
        (dict(precision=2), "precision"),
    ]
    for kw, word in bad:
        cfg = solver.make_config(N=3, **kw)
        assert _lib.lib().nemo_gpu_create(ctypes.byref(cfg), ctypes.byref(h)) == _lib.NEMO_ERR_ARG, kw
        assert word in _lib.last_error(), (kw, _lib.last_error())


def test_operators_entry_point():
    from nemo_amd import operators
    x, w, D, Vi = operators.operators(3)
    assert abs(w.sum() - 2.0) < 1e-15
    assert np.allclose(D @ x, 1.0)


def test_mesh_connectivity():
    for n in (1, 2, 4, 5):
        m = meshmod.uniform_periodic(n)
        e = np.arange(m.nelem)
        for f in range(3):
            assert np.array_equal(m.nbr[m.nbr[e, 2 * f + 1], 2 * f], e)
            assert np.array_equal(m.nbr[m.nbr[e, 2 * f], 2 * f + 1], e)
        assert np.array_equal(np.sort(m.lex_id()), e)
        # neighbours are one lattice step away (periodic)
        d = (m.ijk[m.nbr[:, 1]] - m.ijk) % n
        assert np.all(d[:, 0] == (1 % n)) and np.all(d[:, 1:] == 0)
    m = meshmod.uniform_periodic(4)
    assert m.morton and list(m.ijk[:4, 0]) == [0, 1, 0, 1]


def test_initial_data_deterministic_and_order_independent():
    m1 = meshmod.uniform_periodic(4, order="morton")
    m2 = meshmod.uniform_periodic(4, order="lex")
    u1 = setups.kelvin_helmholtz(m1, 3, seed=5)
    u2 = setups.kelvin_helmholtz(m2, 3, seed=5)
    perm = np.argsort(m1.lex_id())
    assert np.array_equal(u1[perm], u2[np.argsort(m2.lex_id())])
    u3 = setups.kelvin_helmholtz(m1, 3, seed=6)
    assert not np.array_equal(u1, u3)
    v = setups.isentropic_vortex(m1, 3)
    assert v[:, 0].min() > 0.45 and v[:, 0].max() <= 1.0
