"""FLASH-format checkpoints (nemo_amd.flash + libnemo_flash.so), read back
with the HDF5 tools' own h5dump: dataset names, types and shapes of the
reference's layout, the block tree (parents level by level, leaves last,
parent/child links, refine levels, which child) and the leaf data bit for
bit, parent blocks zero."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from nemo_amd import amr, flash, mesh as meshmod

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
pytestmark = pytest.mark.skipif(not (H5DUMP and os.path.exists(flash.LIB_PATH)),
                                reason="h5dump or libnemo_flash.so not available")


def _dump_binary(path, dset, tmp, dtype):
    out = os.path.join(tmp, "d.bin")
    subprocess.run([H5DUMP, "-d", dset, "-b", "LE", "-o", out, path], check=True, capture_output=True)
    return np.fromfile(out, dtype=dtype)


def _forest():
    f = amr.Forest.uniform(1, maxlevel=3)
    f = f.refine(np.arange(8) == 3).balance()
    f = f.refine(f.level == 2).balance()
    return f


def test_checkpoint_layout_tree_and_data(tmp_path):
    f = _forest()
    N, M = 3, 4
    rng = np.random.default_rng(0)
    u = rng.random((f.nelem, 5, M ** 3))
    blend = rng.random(f.nelem)
    path = flash.write(flash.checkpoint_path(str(tmp_path / "checkpoints"), 7), f, N, u, time=0.25, dt=1e-3,
                       nstep=12, number=7, blend=blend, dtmin=rng.random(f.nelem))
    assert path.endswith("chkpt_000007.f5") and os.path.exists(path) and not os.path.exists(path + ".partial")
    head = subprocess.run([H5DUMP, "-H", path], check=True, capture_output=True, text=True).stdout
    t = flash.block_tree(f)
    nb = t.numblocks
    for name, shape in (("coordinates", f"( {nb}, 3 )"), ("block size", f"( {nb}, 3 )"),
                        ("bounding box", f"( {nb}, 3, 2 )"), ("gid", f"( {nb}, 15 )"),
                        ("refine level", f"( {nb} )"), ("which child", f"( {nb} )"), ("node type", f"( {nb} )"),
                        ("dens", f"( {nb}, 4, 4, 4 )"), ("ener", f"( {nb}, 4, 4, 4 )"),
                        ("blnd", f"( {nb}, 4, 4, 4 )"), ("unknown names", "( 5, 1 )")):
        i = head.index(f'DATASET "{name}"')
        assert shape in head[i:i + 400], name
    for name in ("sim info", "integer runtime parameters", "real runtime parameters", "string runtime parameters",
                 "logical runtime parameters", "integer scalars", "real scalars", "string scalars", "logical scalars"):
        assert f'DATASET "{name}"' in head, name
    assert 'GROUP "mesh"' in head
    # tree: root, parents by level, then the leaves in element order
    lev = _dump_binary(path, "/refine level", str(tmp_path), "<i4")
    assert lev[0] == 1 and np.all(np.diff(lev[: nb - f.nelem]) >= 0)
    assert np.array_equal(lev[nb - f.nelem:], f.level + 1)
    nt = _dump_binary(path, "/node type", str(tmp_path), "<i4")
    assert np.all(nt[: nb - f.nelem] == 2) and np.all(nt[nb - f.nelem:] == 1)
    gid = _dump_binary(path, "/gid", str(tmp_path), "<i4").reshape(nb, 15)
    wc = _dump_binary(path, "/which child", str(tmp_path), "<i4")
    bb = _dump_binary(path, "/bounding box", str(tmp_path), "<f8").reshape(nb, 3, 2)
    for b in range(nb):
        kids = gid[b, 7:]
        if nt[b] == 2:
            assert np.all(kids > 0)
            for c, k in enumerate(kids):
                assert gid[k - 1, 6] == b + 1 and wc[k - 1] == c + 1
                off = np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1])
                half = 0.5 * (bb[b, :, 1] - bb[b, :, 0])
                assert np.allclose(bb[k - 1, :, 0], bb[b, :, 0] + off * half)
        else:
            assert np.all(kids == -1)
    assert gid[0, 6] == -1 and wc[0] == -1
    # data: leaves bit for bit in node order i + 4j + 16k, parents zero
    for h, name in enumerate(flash.EULER_NAMES):
        d = _dump_binary(path, "/" + name, str(tmp_path), "<f8").reshape(nb, M ** 3)
        assert np.all(d[: nb - f.nelem] == 0.0)
        assert np.array_equal(d[nb - f.nelem:], u[:, h])
    bl = _dump_binary(path, "/blnd", str(tmp_path), "<f8").reshape(nb, M ** 3)
    assert np.array_equal(bl[nb - f.nelem:], np.repeat(blend[:, None], M ** 3, 1))
    txt = subprocess.run([H5DUMP, "-d", "/integer scalars", path], check=True, capture_output=True, text=True).stdout
    assert '"nxb"' in txt and '"globalnumblocks"' in txt and f"{nb}" in txt
    txt = subprocess.run([H5DUMP, "-d", "/unknown names", path], check=True, capture_output=True, text=True).stdout
    for n in flash.EULER_NAMES:
        assert f'"{n}"' in txt


def test_uniform_mesh_checkpoint_is_full_tree(tmp_path):
    m = meshmod.uniform_periodic(4)
    u = np.random.default_rng(1).random((m.nelem, 5, 27))
    path = flash.write(str(tmp_path / "c.f5"), m, 2, u)
    nb = 1 + 8 + 64
    lev = _dump_binary(path, "/refine level", str(tmp_path), "<i4")
    assert lev.tolist() == [1] + [2] * 8 + [3] * 64
    d = _dump_binary(path, "/momy", str(tmp_path), "<f8").reshape(nb, 27)
    assert np.array_equal(d[9:], u[:, 2])


def test_boundary_types_per_box_side(tmp_path):
    """A box with boundary faces records each side's boundary type: periodic,
    outflow (transmissive faces) or reflecting (slip walls)."""
    m = meshmod.with_boundaries(meshmod.uniform_periodic(4), [meshmod.BC_TRANSMISSIVE, meshmod.BC_TRANSMISSIVE,
                                                              meshmod.BC_SLIP_WALL, meshmod.BC_SLIP_WALL, None, None])
    u = np.random.default_rng(2).random((m.nelem, 5, 27))
    path = flash.write(str(tmp_path / "b.f5"), m, 2, u)
    txt = subprocess.run([H5DUMP, "-d", "/string runtime parameters", path], check=True, capture_output=True,
                         text=True).stdout
    found = {}
    for side in ("xl", "xr", "yl", "yr", "zl", "zr"):
        i = txt.index(f"{side}_boundary_type")
        found[side] = txt[i:i + 120].split('"')[2]
    assert found == {"xl": "outflow", "xr": "outflow", "yl": "reflecting", "yr": "reflecting",
                     "zl": "periodic", "zr": "periodic"}, found
    path = flash.write(str(tmp_path / "p.f5"), meshmod.uniform_periodic(4), 2, u)
    txt = subprocess.run([H5DUMP, "-d", "/string runtime parameters", path], check=True, capture_output=True,
                         text=True).stdout
    assert txt.count('"periodic') == 6
