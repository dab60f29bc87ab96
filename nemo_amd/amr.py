"""Host-side adaptive octree mesh: 2:1-balanced leaves of one periodic tree.

Stands in for the p4est forest of the reference (mesh/p4est/mesh_utils_mod.f90)
on the configurations it runs: a single periodic cube tree refined to levels
between ``minlevel`` and ``maxlevel``.  Leaves are kept in Morton (Z-order)
order of their anchors, which is the order p4est gives the leaves of one tree,
so families (the 8 children of one parent) are consecutive with child 0 first
and child c = x + 2y + 4z.

Connectivity follows the reference's side model (quad_t.isides, side_t with
nquads = 1 or 4, mesh/p4est/mesh_types_mod.f90:30-100, built in mesh_alloc,
mesh_utils_mod.f90:340-376), flattened into the device upload format:

    nbr[e, f] >= 0        conforming neighbour element across face f
    nbr[e, f] = -(s + 1)  face f of e lies on a non-conforming (2:1) interface;
                          its flux is read from face slot s
    mortars[m] = (big, face, small_0, small_1, small_2, small_3)

Mortar m owns face slots 5m (the big element's face) and 5m + 1 + k (the face
of small_k).  small_k covers the half k & 1 of the big face along its first
tangential axis and the half k >> 1 along its second one (tangential axes in
increasing order: x-faces (y, z), y-faces (x, z), z-faces (x, y)); this is the
face-child order of the reference's 3-D face arrays
(kernel/mcfv/3d/kernel_flux_mod.f90:100-121, rfaceMat(:,:,iquad) on the first
and rfaceTam(:,:,iquad) on the second face index).

Balance is 2:1 across faces only: DG couples elements through faces, so edge
and corner neighbours may differ by more than one level.

Adaptation (``adapt``) follows mesh_adapt_wrapped (mesh_utils_mod.f90:879-1009):
flag leaves for refinement, flag whole families for coarsening, coarsen, refine,
balance, then pair old and new leaves in Morton order into replacement units
(copy, refine one into eight, coarsen eight into one) for the state transfer.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .mesh import morton_key

GRID_BITS = 16           # anchors are integers on a 2^16 lattice per direction
ROOT = 1 << GRID_BITS

UNIT_COPY, UNIT_REFINE, UNIT_COARSEN = 0, 1, 2

# tangential axes of a face normal to axis d, in increasing order
TANGENTIAL = ((1, 2), (0, 2), (0, 1))


def _mkeys(anchor: np.ndarray) -> np.ndarray:
    """Morton codes (uint64) of lattice anchors [n, 3]."""
    return morton_key(anchor[:, 0], anchor[:, 1], anchor[:, 2])


# Moves on the Morton code itself (x in bits 0, 3, 6, ..., y in 1, 4, ...,
# z in 2, 5, ...).  Each axis is a dilated integer: with M its bit mask,
# x + y = ((x | ~M) + y) & M (the carry runs through the other axes' bits,
# set to 1) and x - y = ((x & M) - y) & M; bits past the lattice drop out,
# which is the periodic wrap of the box.
_AXIS_MASK = [np.uint64(sum(1 << (3 * i + d) for i in range(GRID_BITS))) for d in range(3)]


def _dilated(size: np.ndarray, d: int) -> np.ndarray:
    """A power-of-two lattice length along axis d, in Morton-code bits."""
    e = np.log2(size.astype(np.float64)).astype(np.uint64)
    return np.left_shift(np.uint64(1), np.uint64(3) * e + np.uint64(d))


def _key_move(key: np.ndarray, d: int, step: np.ndarray, s: int) -> np.ndarray:
    """Morton codes of the anchors moved by s * (length) along axis d
    (periodic), step = that length from _dilated."""
    M = _AXIS_MASK[d]
    moved = (((key | ~M) + step) if s > 0 else ((key & M) - step)) & M
    return (key & ~M) | moved


def _key_floor(key: np.ndarray, size: np.ndarray) -> np.ndarray:
    """Morton codes of the anchors rounded down to multiples of size (a power
    of two, per anchor) in every axis."""
    e = np.log2(size.astype(np.float64)).astype(np.uint64)
    return key & ~(np.left_shift(np.uint64(1), np.uint64(3) * e) - np.uint64(1))


class _Index:
    """Sorted-key lookup of leaves by (level, Morton code of the anchor)."""

    def __init__(self, level: np.ndarray, mkey: np.ndarray):
        k = mkey.astype(np.int64) * 32 + level.astype(np.int64)
        self.order = np.argsort(k, kind="stable")
        self.sorted = k[self.order]

    def find(self, level: np.ndarray, mkey: np.ndarray) -> np.ndarray:
        """Leaf index of each (level, Morton code of the anchor), or -1."""
        k = mkey.astype(np.int64) * 32 + level.astype(np.int64)
        pos = np.searchsorted(self.sorted, k)
        pos = np.minimum(pos, len(self.sorted) - 1)
        hit = self.sorted[pos] == k
        return np.where(hit, self.order[pos], -1)


def _size(level: np.ndarray) -> np.ndarray:
    return np.left_shift(np.int64(1), (GRID_BITS - level.astype(np.int64)))


def _morton_sort(level: np.ndarray, anchor: np.ndarray):
    key = morton_key(anchor[:, 0], anchor[:, 1], anchor[:, 2])
    p = np.argsort(key, kind="stable")
    return level[p], anchor[p]


def _children(level: np.ndarray, anchor: np.ndarray):
    """The 8 children of each leaf, child c = x + 2y + 4z, grouped per parent."""
    h = _size(level + 1)
    c = np.arange(8)
    off = np.stack([c & 1, (c >> 1) & 1, (c >> 2) & 1], axis=1)             # [8, 3]
    ach = anchor[:, None, :] + off[None, :, :] * h[:, None, None]          # [n, 8, 3]
    lch = np.repeat(level + 1, 8)
    return lch, ach.reshape(-1, 3)


@dataclass
class Forest:
    """Leaves of one periodic tree over the box [origin, origin + extent]."""
    level: np.ndarray           # int64 [n]
    anchor: np.ndarray          # int64 [n, 3], lattice units (2^16 per box edge)
    minlevel: int
    maxlevel: int
    extent: tuple = (1.0, 1.0, 1.0)
    origin: tuple = (0.0, 0.0, 0.0)
    _conn: tuple = field(default=None, repr=False, compare=False)
    _fam: np.ndarray = field(default=None, repr=False, compare=False)
    _cols: tuple = field(default=None, repr=False, compare=False)

    # ---- construction ----------------------------------------------------
    @classmethod
    def uniform(cls, level: int, minlevel: int = None, maxlevel: int = None, extent=(1.0, 1.0, 1.0),
                origin=(0.0, 0.0, 0.0)) -> "Forest":
        n = 1 << level
        ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        h = ROOT >> level
        anchor = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1).astype(np.int64) * h
        lv, an = _morton_sort(np.full(n ** 3, level, dtype=np.int64), anchor)
        return cls(lv, an, level if minlevel is None else minlevel, level if maxlevel is None else maxlevel,
                   tuple(extent), tuple(origin))

    @property
    def nelem(self) -> int:
        return int(self.level.shape[0])

    def _replace(self, level, anchor) -> "Forest":
        lv, an = _morton_sort(level, anchor)
        return Forest(lv, an, self.minlevel, self.maxlevel, self.extent, self.origin)

    def refine(self, flags: np.ndarray) -> "Forest":
        """Replace every flagged leaf by its 8 children (no balancing)."""
        flags = np.asarray(flags, dtype=bool)
        lch, ach = _children(self.level[flags], self.anchor[flags])
        level = np.concatenate([self.level[~flags], lch])
        anchor = np.concatenate([self.anchor[~flags], ach])
        return self._replace(level, anchor)

    def family_starts(self) -> np.ndarray:
        """Boolean mask: leaf i is child 0 of a complete family i..i+7 (a
        copy of the forest's cached mask; leaves are never changed in
        place)."""
        if self._fam is None:
            self._fam = self._family_starts()
        return self._fam.copy()

    def _family_starts(self) -> np.ndarray:
        n = self.nelem
        out = np.zeros(n, dtype=bool)
        if n < 8:
            return out
        lv = self.level
        cand = np.nonzero(lv[: n - 7] > 0)[0]
        h = _size(lv[cand])
        first = np.all((self.anchor[cand] & (2 * h)[:, None] - 1) == 0, axis=1)
        cand = cand[first]
        ok = np.ones(len(cand), dtype=bool)
        base = self.anchor[cand]
        hh = _size(lv[cand])
        for c in range(8):
            off = np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1], dtype=np.int64)
            j = cand + c
            ok &= lv[j] == lv[cand]
            ok &= np.all(self.anchor[j] == base + off[None, :] * hh[:, None], axis=1)
        out[cand[ok]] = True
        return out

    def coarsen(self, starts: np.ndarray) -> "Forest":
        """Replace each family starting at a True entry of ``starts`` by its parent."""
        starts = np.asarray(starts, dtype=bool)
        if not starts.any():
            return self
        drop = np.zeros(self.nelem, dtype=bool)
        idx = np.nonzero(starts)[0]
        for c in range(8):
            drop[idx + c] = True
        level = np.concatenate([self.level[~drop], self.level[idx] - 1])
        anchor = np.concatenate([self.anchor[~drop], self.anchor[idx]])
        return self._replace(level, anchor)

    def balance(self) -> "Forest":
        """Refine until face neighbours differ by at most one level (2:1).

        The leaves tile the box in Morton order, so the leaf containing a
        lattice point is the last one whose anchor key is not above the
        point's key.  A leaf needs refining when a face neighbour point of
        a leaf two or more levels finer lies in it.  After the first round a
        violation can only come from a leaf just made (refining a leaf only
        makes the leaves around it satisfy the rule more easily) or persist
        at a leaf whose neighbour was refined but is still too coarse, so
        later rounds look from those leaves alone, and the Morton codes of
        the other leaves are carried over."""
        level, anchor = self.level, self.anchor
        mk = _mkeys(anchor)
        look = None  # the leaves to look from: all, then the new and the failed ones
        while True:
            keys = mk.astype(np.int64)
            lv = level if look is None else level[look]
            mq = mk if look is None else mk[look]
            h = _size(lv)
            need = np.zeros(len(level), dtype=bool)
            failed = np.zeros(len(lv), dtype=bool)
            for d in range(3):
                step = _dilated(h, d)
                for s in (-1, 1):
                    c = np.searchsorted(keys, _key_move(mq, d, step, s).astype(np.int64), side="right") - 1
                    bad = level[c] < lv - 1
                    need[c[bad]] = True
                    failed |= bad
            if not need.any():
                if look is None:
                    return self
                return Forest(level, anchor, self.minlevel, self.maxlevel, self.extent, self.origin)
            again = np.zeros(len(level), dtype=bool)
            again[np.arange(len(level)) if look is None else look] = failed
            lch, ach = _children(level[need], anchor[need])
            level = np.concatenate([level[~need], lch])
            anchor = np.concatenate([anchor[~need], ach])
            mk = np.concatenate([mk[~need], _mkeys(ach)])
            again = np.concatenate([again[~need], np.ones(len(lch), dtype=bool)])
            p = np.argsort(mk, kind="stable")
            level, anchor, mk, again = level[p], anchor[p], mk[p], again[p]
            look = np.nonzero(again)[0]

    # ---- geometry and connectivity ---------------------------------------
    @property
    def delta(self) -> np.ndarray:
        h = _size(self.level).astype(np.float64) / ROOT
        return np.ascontiguousarray(h[:, None] * np.asarray(self.extent)[None, :])

    @property
    def lower(self) -> np.ndarray:
        return np.ascontiguousarray(np.asarray(self.origin)[None, :]
                                    + self.anchor.astype(np.float64) / ROOT * np.asarray(self.extent)[None, :])

    def node_coords(self, xi: np.ndarray) -> np.ndarray:
        """Physical coordinates [nelem, 3, M^3] of the LGL nodes xi (on [-1,1])."""
        M = len(xi)
        ii, jj, kk = np.meshgrid(np.arange(M), np.arange(M), np.arange(M), indexing="ij")
        ref = np.stack([xi[ii].ravel(order="F"), xi[jj].ravel(order="F"), xi[kk].ravel(order="F")])
        return self.lower[:, :, None] + 0.5 * (ref[None, :, :] + 1.0) * self.delta[:, :, None]

    def neighbour_columns(self):
        """The conforming face neighbours as six index arrays, one per face,
        holding n (a slot past the last leaf) where the face has none, and the
        big sides [nm] and small sides [nm, 4] of the mortars; cached with the
        forest (the gathers of amr_run.face_neighbours_any)."""
        if self._cols is None:
            nbr, mortars = self.connectivity()
            n = self.nelem
            cols = tuple(np.ascontiguousarray(np.where(nbr[:, f] >= 0, nbr[:, f], n)) for f in range(6))
            self._cols = (cols, np.ascontiguousarray(mortars[:, 0]), np.ascontiguousarray(mortars[:, 2:]))
        return self._cols

    def connectivity(self):
        """(nbr [n, 6] int32, mortars [nm, 6] int32) of the upload format."""
        if self._conn is not None:
            return self._conn
        n = self.nelem
        mk = _mkeys(self.anchor)
        index = _Index(self.level, mk)
        h = _size(self.level)
        nbr = np.full((n, 6), np.iinfo(np.int32).min, dtype=np.int64)
        big_faces = []  # (element, face, children[4])
        for d in range(3):
            t1, t2 = TANGENTIAL[d]
            for side in (0, 1):
                f = 2 * d + side
                s = 1 if side else -1
                kp = _key_move(mk, d, _dilated(h, d), s)   # anchor of the same-size neighbour
                same = index.find(self.level, kp)
                nbr[same >= 0, f] = same[same >= 0]
                rest = np.nonzero(same < 0)[0]
                if len(rest) == 0:
                    continue
                # coarser neighbour: this element is a small side (filled from the big side below)
                lvl = self.level[rest]
                coarse = index.find(lvl - 1, _key_floor(kp[rest], 2 * h[rest]))
                fine_idx = rest[coarse < 0]
                if len(fine_idx) == 0:
                    continue
                # finer neighbours: this element is the big side of a mortar;
                # the same-size anchor is a multiple of h, so adding half of h
                # along an axis sets one bit
                hh = h[fine_idx] // 2
                base = kp[fine_idx] | (_dilated(hh, d) if side == 0 else np.uint64(0))
                kids = np.empty((len(fine_idx), 4), dtype=np.int64)
                for k in range(4):
                    a = base.copy()
                    if k & 1:
                        a |= _dilated(hh, t1)
                    if k >> 1:
                        a |= _dilated(hh, t2)
                    kids[:, k] = index.find(self.level[fine_idx] + 1, a)
                if (kids < 0).any():
                    raise ValueError("mesh is not 2:1 balanced across faces")
                big_faces.append(np.concatenate([fine_idx[:, None], np.full((len(fine_idx), 1), f), kids], axis=1))
        if big_faces:
            mortars = np.concatenate(big_faces).astype(np.int64)
            mortars = mortars[np.lexsort((mortars[:, 1], mortars[:, 0]))]
        else:
            mortars = np.zeros((0, 6), dtype=np.int64)
        m = np.arange(len(mortars))
        nbr[mortars[:, 0], mortars[:, 1]] = -(5 * m + 1)
        fo = mortars[:, 1] ^ 1
        for k in range(4):
            nbr[mortars[:, 2 + k], fo] = -(5 * m + 2 + k)
        mortars = np.ascontiguousarray(mortars.astype(np.int32))
        if (nbr == np.iinfo(np.int32).min).any():
            raise ValueError("unresolved face (mesh is not 2:1 balanced)")
        self._conn = (np.ascontiguousarray(nbr.astype(np.int32)), mortars)
        return self._conn

    @property
    def nbr(self) -> np.ndarray:
        return self.connectivity()[0]

    @property
    def mortars(self) -> np.ndarray:
        return self.connectivity()[1]

    # ---- adaptation --------------------------------------------------------
    def _drop_rerefined(self, starts: np.ndarray, refine: np.ndarray) -> np.ndarray:
        """Remove the families whose parent balance would split again at once.

        The parent P (level L) of a coarsened family is refined back into
        the same 8 children when a face neighbour of P ends up at level
        L + 2: a conforming neighbour of a child that is refined now, or a
        finer neighbour (mortar big side) not all of whose four small
        elements are coarsened now.  A coarser leaf never makes another leaf
        refine, so leaving such a family alone gives the same forest; it
        only skips building and balancing one that is undone."""
        nbr, mortars = self.connectivity()
        starts = starts.copy()
        while True:
            idx = np.nonzero(starts)[0]
            if len(idx) == 0:
                return starts
            coarsened = np.zeros(self.nelem, dtype=bool)
            for c in range(8):
                coarsened[idx + c] = True
            nb = nbr[idx[:, None] + np.arange(8)[None, :]].astype(np.int64)      # [nf, 8, 6]
            conf = nb >= 0
            nbc = np.where(conf, nb, 0)
            lo = idx[:, None, None]
            outside = conf & ((nbc < lo) | (nbc >= lo + 8))
            doomed = (outside & refine[nbc]).any(axis=(1, 2))
            big = (nb < 0) & ((-nb - 1) % 5 == 0)
            if len(mortars):
                m = np.where(big, (-nb - 1) // 5, 0)
                small = mortars[m][..., 2:6].astype(np.int64)                    # [nf, 8, 6, 4]
                stay = ~coarsened[small].all(axis=-1)
                doomed |= (big & stay).any(axis=(1, 2))
            if not doomed.any():
                return starts
            starts[idx[doomed]] = False

    def adapt(self, refine: np.ndarray, coarsen: np.ndarray, prune: bool = True):
        """One adaptation pass (mesh_utils_mod.f90:879-1009).

        ``refine[i]``: leaf i wants refinement (honoured below maxlevel).
        ``coarsen[i]``: leaf i may be merged with its siblings; a family is
        coarsened when all 8 members agree, none wants refinement and the
        family is above minlevel.  Returns (new forest, units) where units is
        int32 [nunits, 3] = (kind, first old leaf, first new leaf), kinds
        UNIT_COPY (1 -> 1), UNIT_REFINE (1 -> 8 children in child order) and
        UNIT_COARSEN (8 siblings -> 1), in Morton order; or (self, None) if
        nothing changes.
        """
        refine = np.asarray(refine, dtype=bool) & (self.level < self.maxlevel)
        coarsen = np.asarray(coarsen, dtype=bool) & ~refine
        starts = self.family_starts() & (self.level > self.minlevel)
        idx = np.nonzero(starts)[0]
        ok = np.ones(len(idx), dtype=bool)
        for c in range(8):
            ok &= coarsen[idx + c]
        starts[:] = False
        starts[idx[ok]] = True
        if prune and starts.any():
            starts = self._drop_rerefined(starts, refine)
        if not refine.any() and not starts.any():
            return self, None
        new = self.coarsen_refine(starts, refine)
        new = new.balance()
        if new.nelem == self.nelem and np.array_equal(new.level, self.level) and np.array_equal(new.anchor,
                                                                                                 self.anchor):
            return self, None
        return new, replacement_units(self, new)

    def coarsen_refine(self, starts: np.ndarray, refine: np.ndarray) -> "Forest":
        """Coarsen the flagged families, then refine the flagged leaves."""
        drop = np.zeros(self.nelem, dtype=bool)
        idx = np.nonzero(starts)[0]
        for c in range(8):
            drop[idx + c] = True
        keep = ~drop & ~refine
        lch, ach = _children(self.level[refine & ~drop], self.anchor[refine & ~drop])
        level = np.concatenate([self.level[keep], self.level[idx] - 1, lch])
        anchor = np.concatenate([self.anchor[keep], self.anchor[idx], ach])
        return self._replace(level, anchor)


def replacement_units(old: Forest, new: Forest) -> np.ndarray:
    """Pair old and new leaves (both in Morton order) into transfer units.

    The pairing of the replace loop of mesh_adapt_wrapped
    (mesh_utils_mod.f90:967-1000), which walks both leaf sequences in
    lockstep: a new leaf one level finer than the old leaf containing its
    anchor starts a refine unit (8 new), one level coarser a coarsen unit
    (8 old), equal level a copy.  Computed at once from the Morton keys of
    the anchors (a leaf covers the key range [key, key + 8^(bits - level))).
    """
    ko = morton_key(old.anchor[:, 0], old.anchor[:, 1], old.anchor[:, 2]).astype(np.int64)
    kn = morton_key(new.anchor[:, 0], new.anchor[:, 1], new.anchor[:, 2]).astype(np.int64)
    if ko[0] != 0 or kn[0] != 0 or np.any(np.diff(ko) <= 0) or np.any(np.diff(kn) <= 0):
        raise ValueError("leaves are not in Morton order")
    io = np.searchsorted(ko, kn, side="right") - 1          # old leaf containing each new anchor
    dl = new.level - old.level[io]
    if np.any(np.abs(dl) > 1):
        raise ValueError("a leaf changed by more than one level")
    first = ko[io] == kn                                     # new anchor starts the old leaf
    copy = (dl == 0) & first
    refine = (dl == 1) & first
    coarsen = dl == -1                                       # io is then the parent's first child
    start = copy | refine | coarsen
    if np.any((dl == 0) & ~first) or np.any((dl == -1) & ~first):
        raise ValueError("old and new leaves out of step")
    j = np.nonzero(start)[0]
    kind = np.where(copy[j], UNIT_COPY, np.where(refine[j], UNIT_REFINE, UNIT_COARSEN))
    units = np.stack([kind, io[j], j], axis=1).astype(np.int32)
    # every old leaf consumed once, every new leaf produced once, in order
    n_old = np.where(kind == UNIT_COARSEN, 8, 1)
    n_new = np.where(kind == UNIT_REFINE, 8, 1)
    if (not np.array_equal(units[:, 1], np.concatenate([[0], np.cumsum(n_old)[:-1]]))
            or not np.array_equal(units[:, 2], np.concatenate([[0], np.cumsum(n_new)[:-1]]))
            or n_old.sum() != old.nelem or n_new.sum() != new.nelem):
        raise ValueError("old and new leaves do not cover the same domain")
    ref = units[units[:, 0] == UNIT_REFINE]
    for c in range(8):
        if np.any(new.level[ref[:, 2] + c] != old.level[ref[:, 1]] + 1):
            raise ValueError("refined leaf without its 8 children")
    crs = units[units[:, 0] == UNIT_COARSEN]
    for c in range(8):
        if np.any(old.level[crs[:, 1] + c] != new.level[crs[:, 2]] + 1):
            raise ValueError("coarsened leaf without its 8 children")
    return units
