"""FLASH-format checkpoints of the device state (host side).

The reference writes FLASH 4.3 checkpoint files
(checkpoint/flash/checkpoint_mod.f90:11-364) with the block tree that
mesh_alloc_tree builds (mesh/p4est/mesh_utils_mod.f90:443-586): the root,
then the parent blocks of each level in Morton order, then every leaf in
Morton order; gid holds parent and children (1-based, -1 where none; the
face-neighbour entries stay -1 as in the reference), refine level = level + 1,
which child = 1 + x + 2y + 4z.  Leaves carry the state (one dataset per
variable, nodes of a block in x-fastest order); parent blocks stay zero.
The HDF5 writing is in libnemo_flash.so (nemo_amd/csrc/flash_checkpoint.cpp).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import amr
from .mesh import morton_key

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnemo_flash.so")

EULER_NAMES = ["dens", "momx", "momy", "momz", "ener"]
MHD_NAMES = EULER_NAMES + ["magx", "magy", "magz", "glmp"]


class Meta(ctypes.Structure):
    _fields_ = [("nxb", ctypes.c_int32), ("nstep", ctypes.c_int32), ("file_number", ctypes.c_int32),
                ("lrefine_min", ctypes.c_int32), ("lrefine_max", ctypes.c_int32),
                ("file_format_version", ctypes.c_int32), ("boundary", ctypes.c_int8 * 6), ("pad", ctypes.c_int16),
                ("time", ctypes.c_double), ("dt", ctypes.c_double), ("next_checkpoint_time", ctypes.c_double),
                ("xmin", ctypes.c_double), ("xmax", ctypes.c_double), ("ymin", ctypes.c_double),
                ("ymax", ctypes.c_double), ("zmin", ctypes.c_double), ("zmax", ctypes.c_double),
                ("cfl", ctypes.c_double), ("gamma", ctypes.c_double), ("tinitial", ctypes.c_double),
                ("tmax", ctypes.c_double), ("checkpoint_interval_time", ctypes.c_double)]


class Tree(ctypes.Structure):
    _fields_ = [("numblocks", ctypes.c_int64), ("nleaves", ctypes.c_int64),
                ("coordinates", ctypes.c_void_p), ("block_size", ctypes.c_void_p),
                ("bounding_box", ctypes.c_void_p), ("gid", ctypes.c_void_p), ("node_type", ctypes.c_void_p),
                ("refine_level", ctypes.c_void_p), ("which_child", ctypes.c_void_p), ("dtmin", ctypes.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (nemo_amd.build.build_flash needs the HDF5 C library)")
        L = ctypes.CDLL(LIB_PATH)
        L.nemo_flash_write.argtypes = [ctypes.c_char_p, ctypes.POINTER(Meta), ctypes.POINTER(Tree), ctypes.c_int32,
                                       ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p]
        L.nemo_flash_write.restype = ctypes.c_int
        L.nemo_flash_error_message.argtypes = [ctypes.c_char_p, ctypes.c_int32]
        _lib = L
    return _lib


@dataclass
class BlockTree:
    level: np.ndarray         # [numblocks]
    anchor: np.ndarray        # [numblocks, 3] lattice units
    nleaves: int
    gid: np.ndarray           # [numblocks, 15] int32
    which_child: np.ndarray   # [numblocks] int32

    @property
    def numblocks(self) -> int:
        return int(self.level.shape[0])


def block_tree(forest: amr.Forest) -> BlockTree:
    """Parents (levels 0..max-1, Morton order within a level) then leaves."""
    plev, panc = [], []
    for lv in range(int(forest.level.max())):
        sel = forest.level > lv
        mask = ~(np.int64(amr.ROOT >> lv) - 1)
        a = np.unique(forest.anchor[sel] & mask, axis=0)
        key = morton_key(a[:, 0], a[:, 1], a[:, 2])
        a = a[np.argsort(key, kind="stable")]
        plev.append(np.full(len(a), lv, dtype=np.int64))
        panc.append(a)
    level = np.concatenate(plev + [forest.level.astype(np.int64)])
    anchor = np.concatenate(panc + [forest.anchor.astype(np.int64)])
    nb = len(level)
    index = {(int(l), int(x), int(y), int(z)): i for i, (l, (x, y, z)) in enumerate(zip(level, anchor))}
    gid = np.full((nb, 15), -1, dtype=np.int32)
    which = np.full(nb, -1, dtype=np.int32)
    for i in range(nb):
        lv = int(level[i])
        if lv > 0:
            h = amr.ROOT >> lv
            pa = anchor[i] & ~np.int64(2 * h - 1)
            p = index[(lv - 1, int(pa[0]), int(pa[1]), int(pa[2]))]
            c = int(((anchor[i] - pa) // h) @ np.array([1, 2, 4]))
            gid[i, 6] = p + 1
            gid[p, 7 + c] = i + 1
            which[i] = c + 1
    return BlockTree(level, anchor, forest.nelem, gid, which)


def checkpoint_path(directory: str, number: int, prefix: str = "chkpt_", extension: str = "f5") -> str:
    """checkpoints/chkpt_NNNNNN.f5 (checkpoint_types_mod.f90:6-23, checkpoint_mod.f90:28-31)."""
    return os.path.join(directory, f"{prefix}{number:06d}.{extension}")


def write(path: str, forest, N: int, u: np.ndarray, *, time: float = 0.0, dt: float = 0.0, nstep: int = 0,
          number: int = 0, gamma: float = 1.4, cfl: float = 0.5, tmax: float = 0.0, blend=None, dtmin=None,
          names=None) -> str:
    """Write the state u [nelem, V, M^3] (element order of the forest) as a
    FLASH checkpoint; a uniform power-of-two mesh is taken as a forest of one
    level.  Returns the path."""
    # boundary type of each box side: periodic, or outflow / reflecting for
    # the transmissive / slip-wall faces of a box with boundary faces
    sides = [0] * 6
    bf = getattr(forest, "bfaces", None)
    if bf is not None:
        for _, f, kind in np.asarray(bf).reshape(-1, 3):
            sides[int(f)] = 2 if int(kind) == 1 else 1
    if not isinstance(forest, amr.Forest):
        n = forest.shape[0]
        if not (forest.morton and n & (n - 1) == 0 and forest.shape == (n, n, n)):
            raise ValueError("FLASH checkpoints need an octree mesh (adaptive forest or uniform 2^k cube)")
        forest = amr.Forest.uniform(int(np.log2(n)), extent=forest.extent)
    u = np.asarray(u, dtype=np.float64)
    M = N + 1
    V = u.shape[1]
    names = names or (EULER_NAMES if V == 5 else MHD_NAMES)
    t = block_tree(forest)
    size = (amr.ROOT >> t.level).astype(np.float64)[:, None] / amr.ROOT * np.asarray(forest.extent)[None, :]
    lower = np.asarray(forest.origin)[None, :] + t.anchor / amr.ROOT * np.asarray(forest.extent)[None, :]
    coords = np.ascontiguousarray(lower + 0.5 * size)
    bbox = np.ascontiguousarray(np.stack([lower, lower + size], axis=2))
    node_type = np.where(np.arange(t.numblocks) >= t.numblocks - t.nleaves, 1, 2).astype(np.int32)
    refine = (t.level + 1).astype(np.int32)
    leaf = np.ascontiguousarray(np.transpose(u, (1, 0, 2)))            # [V][nleaves][M^3]
    blend_arr = None
    if blend is not None:
        blend_arr = np.ascontiguousarray(np.repeat(np.asarray(blend, dtype=np.float64)[:, None], M ** 3, axis=1))
    dtmin_arr = None if dtmin is None else np.ascontiguousarray(dtmin, dtype=np.float64)
    lo, hi = forest.origin, tuple(o + e for o, e in zip(forest.origin, forest.extent))
    meta = Meta(nxb=M, nstep=nstep, file_number=number, lrefine_min=forest.minlevel + 1,
                lrefine_max=forest.maxlevel + 1, file_format_version=9, boundary=(ctypes.c_int8 * 6)(*sides), pad=0,
                time=time, dt=dt,
                next_checkpoint_time=0.0, xmin=lo[0], xmax=hi[0], ymin=lo[1], ymax=hi[1], zmin=lo[2], zmax=hi[2],
                cfl=cfl, gamma=gamma, tinitial=0.0, tmax=tmax, checkpoint_interval_time=0.0)
    tree = Tree(numblocks=t.numblocks, nleaves=t.nleaves, coordinates=coords.ctypes.data,
                block_size=np.ascontiguousarray(size).ctypes.data, bounding_box=bbox.ctypes.data,
                gid=t.gid.ctypes.data, node_type=node_type.ctypes.data, refine_level=refine.ctypes.data,
                which_child=t.which_child.ctypes.data,
                dtmin=None if dtmin_arr is None else dtmin_arr.ctypes.data)
    keep = (size, coords, bbox, node_type, refine, leaf, blend_arr, dtmin_arr)  # alive during the call
    nm = b"".join(n.encode()[:4].ljust(4, b"\0") for n in names)
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    rc = lib().nemo_flash_write(path.encode(), ctypes.byref(meta), ctypes.byref(tree), V, nm, leaf.ctypes.data,
                                None if blend_arr is None else blend_arr.ctypes.data)
    del keep
    if rc != 0:
        buf = ctypes.create_string_buffer(512)
        lib().nemo_flash_error_message(buf, 512)
        raise RuntimeError(f"FLASH checkpoint failed: {buf.value.decode()}")
    return path
