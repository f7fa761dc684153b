"""CPU oracle for the jztree tree hot path.  TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may import this
module; the product package ``jztree_amd`` never does.  It is the checker, never the thing
measured or shipped.

Two layers:
  * ctypes bindings to ``libjzoracle.so`` (``jzoracle.c``): the z-order math, the tree kernels
    (FlagLeafBoundaries / FindNodeBoundaries / GetNodeGeometry / CenterOfMass), brute-force and
    grid kNN, FoF.  CenterOfMass: parity with the reference's compiled bits unpinned (it depends
    on the FMA contraction its compiler applies); pinned by closed forms, long-double sums and
    the plane recursion.
  * a numpy restatement of the reference's JAX tree assembly (``src/jztree/tree.py:55-329``,
    ``:600-829``, ``data.py:251-452``), integer-exact.

Parity pinning of each part is described next to the function and checked in ``tests/``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_c = ctypes


def build():
    """Compile libjzoracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libjzoracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_msb_xor_f32.argtypes = [_c.c_float, _c.c_float]
        L.orc_msb_xor_f32.restype = _c.c_int32
        L.orc_msb_xor_f64.argtypes = [_c.c_double, _c.c_double]
        L.orc_msb_xor_f64.restype = _c.c_int32
        L.orc_msb_diff_level_f32.argtypes = [_f32p, _f32p, _c.c_int]
        L.orc_msb_diff_level_f32.restype = _c.c_int32
        L.orc_z_pos_less_f32.argtypes = [_f32p, _f32p, _c.c_int]
        L.orc_z_pos_less_f64.argtypes = [_f64p, _f64p, _c.c_int]
        L.orc_z_exact_cmp_f32.argtypes = [_f32p, _f32p, _c.c_int]
        L.orc_z_exact_cmp_f64.argtypes = [_f64p, _f64p, _c.c_int]
        L.orc_zsort_f32.argtypes = [_f32p, _c.c_int64, _c.c_int, _i64p]
        L.orc_zsort_f64.argtypes = [_f64p, _c.c_int64, _c.c_int, _i64p]
        L.orc_zsort_i32.argtypes = [_i32p, _c.c_int64, _c.c_int, _i64p]
        L.orc_zsort_i64.argtypes = [_i64p, _c.c_int64, _c.c_int, _i64p]
        L.orc_flag_leaf_boundaries_f32.argtypes = [
            _f32p, _u8p, _i32p, _c.c_int32, _c.c_int64, _c.c_int, _c.c_int, _c.c_int, _c.c_int,
            _i8p, _i32p]
        L.orc_find_node_boundaries_f32.argtypes = [
            _f32p, _f32p, _c.c_int32, _c.c_int64, _c.c_int, _c.c_int, _c.c_int, _i32p, _i32p, _i32p]
        L.orc_get_node_geometry_f32.argtypes = [
            _f32p, _i32p, _i32p, _c.c_int32, _c.c_int64, _c.c_int64, _c.c_int, _c.c_int,
            _c.c_uint32, _c.c_int, _i32p, _f32p, _f32p]
        L.orc_r2_f32.argtypes = [_f32p, _f32p, _c.c_int, _c.c_float]
        L.orc_r2_f32.restype = _c.c_float
        knn_args = [_f32p, _c.c_void_p, _c.c_int64, _f32p, _c.c_int64, _c.c_int, _c.c_int,
                    _c.c_float, _f32p, _i64p]
        L.orc_knn_brute_f32.argtypes = knn_args
        L.orc_knn_grid_f32.argtypes = knn_args
        L.orc_knn_kdtree_f32.argtypes = knn_args
        L.orc_r2_f64.argtypes = [_f64p, _f64p, _c.c_int, _c.c_float]
        L.orc_r2_f64.restype = _c.c_double
        knn_nd_args = [_c.c_void_p, _c.c_void_p, _c.c_int64, _c.c_void_p, _c.c_int64, _c.c_int, _c.c_int,
                       _c.c_float, _c.c_int, _f64p, _i64p]
        L.orc_knn_brute_nd.argtypes = knn_nd_args
        L.orc_knn_kdtree_nd.argtypes = knn_nd_args
        L.orc_knn_kdtree_nd.restype = _c.c_int
        fof_args = [_f32p, _c.c_int64, _c.c_int, _c.c_float, _c.c_float, _c.c_void_p, _i64p]
        L.orc_fof_f32.argtypes = fof_args
        L.orc_fof_brute_f32.argtypes = fof_args
        L.orc_fof_cells_f32.argtypes = fof_args
        L.orc_num_threads.restype = _c.c_int
        L.orc_center_of_mass_f32.argtypes = [_i32p, _f32p, _f32p, _c.c_int64, _c.c_int, _c.c_int, _f32p]
        L.orc_center_of_mass_f64.argtypes = [_i32p, _f64p, _f64p, _c.c_int64, _c.c_int, _c.c_int, _f64p]
        _LIB = L
    return _LIB


def num_threads() -> int:
    return lib().orc_num_threads()


# ------------------------------------------------------------------------------------------------
# z-order math (math.cuh:241-402) and sort (sort.cuh:40-76)
# ------------------------------------------------------------------------------------------------

def msb_xor(a: float, b: float, dtype=np.float32) -> int:
    if np.dtype(dtype) == np.float64:
        return lib().orc_msb_xor_f64(float(a), float(b))
    return lib().orc_msb_xor_f32(float(np.float32(a)), float(np.float32(b)))


def msb_diff_level(p1, p2) -> int:
    p1 = np.ascontiguousarray(p1, np.float32)
    p2 = np.ascontiguousarray(p2, np.float32)
    return lib().orc_msb_diff_level_f32(p1, p2, len(p1))


def z_pos_less(p1, p2) -> bool:
    """Literal reference comparator (math.cuh:381-402)."""
    if np.asarray(p1).dtype == np.float64:
        return bool(lib().orc_z_pos_less_f64(np.ascontiguousarray(p1), np.ascontiguousarray(p2), len(p1)))
    p1 = np.ascontiguousarray(p1, np.float32)
    p2 = np.ascontiguousarray(p2, np.float32)
    return bool(lib().orc_z_pos_less_f32(p1, p2, len(p1)))


def z_exact_cmp(p1, p2) -> int:
    """This build's exact Morton comparator (-1/0/1); see jzoracle.c for the quirks it drops."""
    if np.asarray(p1).dtype == np.float64:
        return lib().orc_z_exact_cmp_f64(np.ascontiguousarray(p1), np.ascontiguousarray(p2), len(p1))
    p1 = np.ascontiguousarray(p1, np.float32)
    p2 = np.ascontiguousarray(p2, np.float32)
    return lib().orc_z_exact_cmp_f32(p1, p2, len(p1))


def zsort(pos: np.ndarray):
    """(posz, idz) with posz = pos[idz]; order = (exact Morton order, input index)."""
    pos = np.ascontiguousarray(pos)
    n, dim = pos.shape
    idz = np.empty(n, np.int64)
    if pos.dtype == np.int32:
        lib().orc_zsort_i32(pos, n, dim, idz)
    elif pos.dtype == np.int64:
        lib().orc_zsort_i64(pos, n, dim, idz)
    elif pos.dtype == np.float64:
        lib().orc_zsort_f64(pos, n, dim, idz)
    else:
        pos = pos.astype(np.float32)
        lib().orc_zsort_f32(pos, n, dim, idz)
    return pos[idz], idz


# ------------------------------------------------------------------------------------------------
# tree kernels (tree.cuh)
# ------------------------------------------------------------------------------------------------

def level_info(dim: int, dtype=np.float32):
    """data.py:352-379: (min_lvl, max_lvl) of the Morton levels."""
    lo, hi = {np.dtype(np.float32): (-150, 128), np.dtype(np.float64): (-1075, 1024)}[np.dtype(dtype)]
    return dim * lo, dim * (hi + 1)


def flag_leaf_boundaries(posz, npart=None, leaf_size=32, lvl_bound=None, ptype=None, num_types=1):
    """tree.cuh:13-98 restated.  Returns (split_flags int8 (N+1,), lvl int32 (N+1,))."""
    posz = np.ascontiguousarray(posz, np.float32)
    n, dim = posz.shape
    if npart is None:
        npart = n
    if lvl_bound is None:
        mx = level_info(dim)[1]
        lvl_bound = (mx, mx)
    lb = np.ascontiguousarray(lvl_bound, np.int32)
    pt = np.zeros(max(n, 1), np.uint8) if ptype is None else np.ascontiguousarray(ptype, np.uint8)
    flags = np.zeros(n + 1, np.int8)
    lvl = np.zeros(n + 1, np.int32)
    scan = (leaf_size + 1) * num_types
    lib().orc_flag_leaf_boundaries_f32(posz if n else np.zeros((1, dim), np.float32), pt, lb,
                                        int(npart), n, dim, leaf_size, scan, num_types, flags, lvl)
    return flags, lvl


def find_node_boundaries(pos_in, nleaves, pos_boundary=None, lvl_invalid=-2000):
    """tree.cuh:104-199 restated; pos_boundary defaults to NaN as in tree.py:275."""
    pos_in = np.ascontiguousarray(pos_in, np.float32)
    size, dim = pos_in.shape
    if pos_boundary is None:
        pos_boundary = np.full((2, dim), np.nan, np.float32)
    pos_boundary = np.ascontiguousarray(pos_boundary, np.float32)
    out = [np.zeros(size + 1, np.int32) for _ in range(3)]
    lib().orc_find_node_boundaries_f32(pos_in, pos_boundary, int(min(nleaves, size)), size + 1, dim,
                                       level_info(dim)[1], lvl_invalid, *out)
    return tuple(out)


def get_node_geometry(posz, lbound, rbound, num=None, upper_extent=False, lvl_invalid=-2000):
    """tree.cuh:312-372 restated.  Returns (lvl, center, extent)."""
    posz = np.ascontiguousarray(posz, np.float32)
    lbound = np.ascontiguousarray(lbound, np.int32)
    rbound = np.ascontiguousarray(rbound, np.int32)
    size = len(lbound)
    n, dim = posz.shape
    if num is None:
        num = size
    lvl = np.zeros(size, np.int32)
    cent = np.zeros((size, dim), np.float32)
    ext = np.zeros((size, dim), np.float32)
    lib().orc_get_node_geometry_f32(posz, lbound, rbound, int(num), size, n, dim, lvl_invalid, 7,
                                    int(upper_extent), lvl, cent, ext)
    return lvl, cent, ext


# ------------------------------------------------------------------------------------------------
# tree assembly (tree.py:55-329, :600-829; data.py:251-452), numpy, integer-exact
# ------------------------------------------------------------------------------------------------

@dataclass
class PackedArray:
    """data.py:251-345: several per-plane arrays stacked in one buffer."""
    data: np.ndarray
    ispl: np.ndarray
    fill_values: np.ndarray

    @classmethod
    def from_data(cls, data, ispl, fill_values):
        fv = np.asarray(fill_values)
        if fv.ndim == 0:
            fv = np.full(len(ispl) - 1, fv, dtype=data.dtype)
        return cls(data, np.asarray(ispl), fv.astype(data.dtype))

    def get(self, level, size=None):
        if size is None:
            size = len(self.data)
        idx = np.arange(size) + self.ispl[level]
        valid = idx < self.ispl[level + 1]
        vals = self.data[np.minimum(idx, len(self.data) - 1)]
        fill = self.fill_values[level]
        if self.data.ndim > 1:
            valid = valid[:, None]
        return np.where(valid, vals, fill)

    def num(self, level):
        return int(self.ispl[level + 1] - self.ispl[level])

    def nfilled(self):
        return int(self.ispl[-1])


@dataclass
class TreeHierarchy:
    """data.py:381-452 (numpy mirror)."""
    size_leaves: int
    ispl_n2n: PackedArray
    ispl_n2l: PackedArray
    ispl_l2p_per_type: np.ndarray
    lvl: PackedArray
    geom_cent: PackedArray

    def size(self):
        return len(self.ispl_n2n.data) - 1

    def num_planes(self):
        return len(self.ispl_n2l.ispl) - 1

    def num(self, level):
        return self.lvl.num(level)

    def splits_leaf_to_part(self, ptype=0, size=None):
        if size is None:
            size = self.size() + 1
        return self.ispl_l2p_per_type[ptype][:size]

    def npart(self, level, ptype=0, size=None):
        if size is None:
            size = self.size()
        spl = self.splits_leaf_to_part(ptype)
        n2l = self.ispl_n2l.get(level, size + 1)
        n2p = spl[np.minimum(n2l, len(spl) - 1)]
        return n2p[1:] - n2p[:-1]


def div_ceil(a, b):
    return (a + b - 1) // b


def estimate_node_number(npart, max_node_size, alloc_fac=1.0):
    """tree.py:600-609"""
    return int(div_ceil(npart * alloc_fac, np.maximum(max_node_size // 2, 1))) + 1


def tree_level_node_sizes(npart, max_leaf_size, coarse_fac, stop_coarsen):
    """tree.py:611-619"""
    max_num_leaves = estimate_node_number(npart, max_leaf_size)
    nlevels = np.log(max_num_leaves / stop_coarsen) / np.log(coarse_fac)
    nlevels = np.maximum(int(np.ceil(nlevels)), 1)
    return [int(max_leaf_size * (coarse_fac ** i)) for i in range(0, nlevels)]


def find_regularization_level(lvl, weights, dim, max_vol_fac=20.0, min_percentile=90.0):
    """tree.py:55-68, :160-187, FP32 throughout, sums in ascending level order.

    The reference sums under XLA's reduction order, so borderline decisions (a power of two
    within an ulp of max_vol_fac*avg_vol) are parity-unpinned."""
    mn, mx = level_info(dim)
    nlvl = mx - mn + 1
    idx = np.maximum(np.asarray(lvl, np.int64) - mn, 0)  # jnp.bincount: clip below, drop above
    keep = idx < nlvl
    counts = np.zeros(nlvl, np.int64)
    np.add.at(counts, idx[keep], np.asarray(weights, np.int64)[keep])
    counts = counts.astype(np.float32)
    levels = np.arange(nlvl, dtype=np.int32) + mn
    csum = np.cumsum(counts, dtype=np.float32)
    thresh = np.float32(0.01 * min_percentile) * csum[-1]
    consider = csum >= thresh
    ref_level = int(np.min(np.where(consider, levels, mx)))
    with np.errstate(over="ignore", under="ignore"):
        vol_rel = np.ldexp(np.float32(1.0), (levels - ref_level).astype(np.int32)).astype(np.float32)
        below = levels <= ref_level
        vol_sum = np.float32(0.0)
        cts_sum = np.float32(0.0)
        for v, c, b in zip(vol_rel, counts, below):
            if b:
                vol_sum = np.float32(vol_sum + np.float32(v * c))
                cts_sum = np.float32(cts_sum + c)
        avg_vol = np.float32(vol_sum / cts_sum) if cts_sum > 0 else np.float32(np.nan)
        too_large = consider & (vol_rel >= np.float32(max_vol_fac) * avg_vol)
    first = int(np.argmax(too_large)) if np.any(too_large) else nlvl - 1
    return int(levels[first])


def detect_leaf_boundaries(posz, leaf_size, npart, alloc_size, regularize=None):
    """tree.py:189-262 (single type).  regularize = (max_volume_fac, regularize_percentile) or None."""
    n, dim = posz.shape
    flags, lvl = flag_leaf_boundaries(posz, npart=npart, leaf_size=leaf_size)
    flags = flags.astype(bool)
    if regularize is not None:
        splits = np.nonzero(flags)[0][:alloc_size]
        splits = np.concatenate([splits, np.full(alloc_size - len(splits), npart, np.int64)])
        leaf_npart = splits[1:] - splits[:-1]
        leaf_lvl = np.minimum(lvl[splits[1:]], lvl[splits[:-1]]) - 1
        lvl_max = find_regularization_level(leaf_lvl, leaf_npart, dim, *regularize)
        flags = flags | ((lvl >= lvl_max) & (np.arange(n + 1) < npart + 1))
    nfilled = int(np.sum(flags))
    if nfilled > alloc_size:
        raise RuntimeError(f"Tree-Leaf allocation too small: filled={nfilled} size={alloc_size}.")
    splits = np.nonzero(flags)[0]
    splits = np.concatenate([splits, np.full(alloc_size - len(splits), npart, np.int64)])
    return splits.astype(np.int32)


def define_split_hierarchy(posz, node_sizes, alloc_size, npart, regularize=None):
    """tree.py:628-734 (single type, single device: lvl_bound = (max_lvl, max_lvl))."""
    n, dim = posz.shape
    nlevels = len(node_sizes)
    mx = level_info(dim)[1]
    ispl = detect_leaf_boundaries(posz, node_sizes[0], npart, alloc_size, regularize)
    nleaves = int(np.argmax(ispl))
    reps = posz[np.minimum(ispl[:-1], n - 1)] if n else np.zeros((alloc_size - 1, dim), np.float32)
    lvl, lbound, rbound = find_node_boundaries(reps, nleaves)
    npins = ispl[rbound] - ispl[lbound]
    isl = npins[None, :] > np.array(node_sizes, np.int32)[:, None]
    isl = isl | ((lbound <= 0) & (lvl > mx))[None, :]
    isl = isl | ((rbound >= nleaves) & (lvl > mx))[None, :]
    if regularize is not None:
        new = [(npins > 0) | isl[0]]
        lvl_up = np.minimum(lvl[lbound], lvl[rbound]) - 1
        ar = np.arange(len(lvl))
        for i in range(1, nlevels):
            is_node = (~isl[i]) & isl[i, lbound] & isl[i, rbound]
            is_node = is_node & (ar < nleaves)
            lvl_max = find_regularization_level(lvl_up, np.where(is_node, npins, 0), dim, *regularize)
            new.append(isl[i] | ((lvl > lvl_max) & (ar <= nleaves)))
        isl = np.array(new)
    offsets = np.cumsum(isl.flatten()).reshape(isl.shape)
    level_spl = np.concatenate([[0], offsets[:, -1]]).astype(np.int32)
    nnodes_on_level = level_spl[1:] - level_spl[:-1] - 1
    if level_spl[-1] > alloc_size:
        raise RuntimeError(f"Tree-node allocation too small: filled={level_spl[-1]} size={alloc_size}.")
    offsets = np.where(isl, offsets - isl, alloc_size).astype(np.int64)
    ispl_n2l = np.zeros(alloc_size, np.int32)
    v0 = np.broadcast_to(offsets[0:1, :], offsets.shape)
    m = offsets < alloc_size
    ispl_n2l[offsets[m]] = v0[m]
    ispl_n2l = PackedArray.from_data(ispl_n2l, level_spl, nnodes_on_level[0])
    value = np.empty(offsets.shape, np.int64)
    value[0] = ispl
    for li in range(1, nlevels):
        value[li] = offsets[li - 1] - level_spl[li - 1]
    ispl_n2n = np.zeros(alloc_size, np.int32)
    ispl_n2n[offsets[m]] = value[m]
    fill_val = np.concatenate([[ispl[-1]], nnodes_on_level[:-1]]).astype(np.int32)
    ispl_n2n = PackedArray.from_data(ispl_n2n, level_spl, fill_val)
    return ispl, ispl_n2l, ispl_n2n


def build_tree_hierarchy(posz, max_leaf_size=32, coarse_fac=6.0, stop_coarsen=1024,
                         alloc_fac_nodes=1.0, regularize=None, npart=None):
    """tree.py:756-829 (single device, geometric centres)."""
    posz = np.ascontiguousarray(posz, np.float32)
    n, dim = posz.shape
    if npart is None:
        npart = int(np.sum(~np.isnan(posz[:, 0])))
    node_sizes = tree_level_node_sizes(n, max_leaf_size, coarse_fac, stop_coarsen)
    alloc_size = estimate_node_number(n, max_leaf_size, alloc_fac_nodes)
    ispl, ispl_n2l, ispl_n2n = define_split_hierarchy(posz, node_sizes, alloc_size, npart, regularize)
    n2p = ispl[ispl_n2l.data]
    lvl, cent, _ = get_node_geometry(posz, n2p[:-1], n2p[1:], num=ispl_n2l.nfilled() - 1)
    seams = ispl_n2l.ispl[1:-1] - 1
    lvl = np.delete(lvl, seams)
    cent = np.delete(cent, seams, axis=0)
    prop = np.concatenate([[0], np.cumsum(ispl_n2n.ispl[1:] - ispl_n2n.ispl[:-1] - 1)]).astype(np.int32)
    size_leaves = estimate_node_number(n, max_leaf_size)
    return TreeHierarchy(size_leaves, ispl_n2n, ispl_n2l, ispl.reshape(1, -1),
                         PackedArray.from_data(lvl, prop, -2000),
                         PackedArray.from_data(cent, prop, np.nan))


# ------------------------------------------------------------------------------------------------
# kNN and FoF
# ------------------------------------------------------------------------------------------------

def r2(a, b, boxsize=0.0):
    """Canonical r^2: fma(dz,dz, fma(dy,dy, dx*dx)) of wrapped differences, in the positions' type
    (float32, or float64 with the half box taken as a float, math.cuh:471-522)."""
    if np.asarray(a).dtype == np.float64:
        a = np.ascontiguousarray(a, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        return np.float64(lib().orc_r2_f64(a, b, len(a), float(boxsize)))
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.float32(lib().orc_r2_f32(a, b, len(a), float(boxsize)))


def knn(xs, k, xq=None, boxsize=0.0, ids=None, method="kdtree"):
    """Exact kNN: per query the k smallest (canonical r2, id).  Returns (r2 (nq,k), id (nq,k) i64),
    r2 in the positions' type: float32 (any other input is rounded to it) or float64.

    ids: the id reported (and used for tie-breaks) for each source point; default = its index.
    float32 in 2-3 dims runs the float32 routines (brute, grid, kdtree); float64, or more than 3
    dims (up to 10), the any-dim routines (brute, kdtree; grid maps to kdtree)."""
    is64 = np.asarray(xs).dtype == np.float64
    ty = np.float64 if is64 else np.float32
    xs = np.ascontiguousarray(xs, ty)
    xq = xs if xq is None else np.ascontiguousarray(xq, ty)
    ns, dim = xs.shape
    nq = len(xq)
    io = np.empty((nq, k), np.int64)
    idp = None
    if ids is not None:
        ids = np.ascontiguousarray(ids, np.int64)
        idp = ids.ctypes.data
    if is64 or dim > 3:
        r2d = np.empty((nq, k), np.float64)
        fn = lib().orc_knn_brute_nd if method == "brute" else lib().orc_knn_kdtree_nd
        rc = fn(xs.ctypes.data, idp, ns, xq.ctypes.data, nq, dim, k, float(boxsize), int(is64), r2d, io)
        if rc not in (None, 0):
            raise ValueError(f"knn oracle: unsupported dim {dim}")
        return (r2d if is64 else r2d.astype(np.float32)), io
    r2o = np.empty((nq, k), np.float32)
    fn = {"brute": lib().orc_knn_brute_f32, "grid": lib().orc_knn_grid_f32,
          "kdtree": lib().orc_knn_kdtree_f32}[method]
    fn(xs, idp, ns, xq, nq, dim, k, float(boxsize), r2o, io)
    return r2o, io


def fof(x, rlink, boxsize=0.0, key=None, method="grid"):
    """FoF labels (min key per group) for the predicate canonical r2 <= float32(rlink*rlink).

    method: "grid" (cell list with cells >= rlink), "cells" (fine cells of side <= rlink/sqrt(dim),
    each one group from the start: fast on dense clumps), "brute" (all pairs, tiny n).
    r2link is computed as in fof.py:50/:165: np.float32(rlink*rlink) with rlink*rlink in f64."""
    x = np.ascontiguousarray(x, np.float32)
    n, dim = x.shape
    r2link = float(np.float32(rlink * rlink))
    out = np.empty(n, np.int64)
    kp = None
    if key is not None:
        key = np.ascontiguousarray(key, np.int64)
        kp = key.ctypes.data
    fn = {"brute": lib().orc_fof_brute_f32, "grid": lib().orc_fof_f32, "cells": lib().orc_fof_cells_f32}[method]
    fn(x, n, dim, r2link, float(boxsize), kp, out)
    return out


def fof_is_superset(igroup_sup, igroup):
    """fof.py:203-218 in numpy."""
    igroup_sup = np.asarray(igroup_sup)
    igroup = np.asarray(igroup)
    label_map = np.zeros(len(igroup), np.int64)
    label_map[igroup] = igroup_sup
    return bool(np.all(igroup_sup == label_map[igroup]))


def canonical_labels(labels):
    """Relabel groups to the minimum particle index of each group."""
    labels = np.asarray(labels, np.int64)
    n = len(labels)
    mn = np.full(n if n else 1, np.iinfo(np.int64).max, np.int64)
    uniq, inv = np.unique(labels, return_inverse=True)
    mins = np.full(len(uniq), np.iinfo(np.int64).max, np.int64)
    np.minimum.at(mins, inv, np.arange(n, dtype=np.int64))
    del mn
    return mins[inv]


# ------------------------------------------------------------------------------------------------
# FoF catalogue (fof.py:510-519 _fof_order, :650-766 _fof_catalogue_from_groups; one device)
# ------------------------------------------------------------------------------------------------

def fof_order(igroup, npart=None):
    """fof.py:510-519: (isort, counts in group order).  igr = igroup for i < npart, else n; isort is
    the stable argsort of igr; counts[i] = number of filled particles labelled isort[i]."""
    igroup = np.asarray(igroup, np.int64)
    n = len(igroup)
    npart = n if npart is None else int(npart)
    igr = np.where(np.arange(n) < npart, igroup, n)
    counts = np.zeros(n + 1, np.int64)
    np.add.at(counts, igr, 1)
    counts = counts[:n]  # the scatter at index n is dropped, as jnp's .at[].add drops it
    isort = np.argsort(igr, kind="stable")
    return isort, counts[isort]


def _py_mod(x, y):
    """jnp.remainder: fmod, then + y where the remainder's sign differs from y's."""
    r = np.fmod(x, y)
    fix = (r != 0) & ((r < 0) != (y < 0))
    return np.where(fix, r + y, r).astype(x.dtype)


def fof_catalogue(pos, counts, npart_min, boxsize=0.0, mass=None, vel=None, v_rad=None, zplus1=None,
                  size_cata=None):
    """fof.py:650-766 for one device, on particles in group order: a dict of catalogue arrays.

    Products are formed in the particle type as the reference forms them ((pos - pos0) wrapped,
    times mass), sums in float64.  Groups: every i with counts[i] >= npart_min, in order.  A 2D
    inertia radius counts the second component twice: the reference indexes dx[..., 2], which JAX
    clamps to the last column."""
    pos = np.asarray(pos)
    T = pos.dtype.type
    n, dim = pos.shape
    counts = np.asarray(counts, np.int64)
    if size_cata is None:
        size_cata = (n + npart_min - 1) // npart_min
    box = T(boxsize or 0.0)
    wrap_dx = (lambda d: _py_mod(d + T(0.5) * box, box) - T(0.5) * box) if box > 0 else (lambda d: d)
    wrap_pos = (lambda x: _py_mod(x, box)) if box > 0 else (lambda x: x)
    starts = np.nonzero(counts >= npart_min)[0]
    ng = len(starts)
    k = min(ng, size_cata)
    out = {"ngroups": ng}
    count = np.zeros(size_cata, np.int64)
    offset = np.full(size_cata, n, np.int64)
    count[:k] = counts[starts[:k]]
    offset[:k] = starts[:k]
    out["count"], out["offset"] = count, offset
    m_all = None if mass is None else np.broadcast_to(np.asarray(mass, pos.dtype), (n,))
    mass_array = mass is not None and np.size(mass) > 1
    nan = np.full(size_cata, np.nan, pos.dtype)
    gm = np.zeros(size_cata, pos.dtype)
    com = np.full((size_cata, dim), np.nan, pos.dtype)
    cvel = np.full((size_cata, dim), np.nan, pos.dtype)
    vr, sf, rad = nan.copy(), nan.copy(), nan.copy()
    for g in range(k):
        s, c = int(offset[g]), int(count[g])
        sl = slice(s, s + c)
        m = m_all[sl] if m_all is not None else np.ones(c, pos.dtype)
        gm[g] = T(np.sum(m.astype(np.float64))) if mass_array else T(count[g]) * (T(np.asarray(mass).reshape(-1)[0])
                                                                                    if mass is not None else T(1))
        dx = wrap_dx(pos[sl] - pos[s])
        com[g] = wrap_pos(np.sum((dx * m[:, None]).astype(np.float64), 0).astype(pos.dtype) / gm[g] + pos[s])
        if vel is not None:
            v = np.asarray(vel)[sl]
            cvel[g] = np.sum(((v - v[0]) * m[:, None]).astype(np.float64), 0).astype(pos.dtype) / gm[g] + v[0]
        if v_rad is not None:
            vr[g] = T(np.sum((np.asarray(v_rad)[sl] * m).astype(np.float64))) / gm[g]
        if zplus1 is not None:
            sf[g] = T(np.sum(((T(1) / np.asarray(zplus1)[sl]) * m).astype(np.float64))) / gm[g]
        d = wrap_dx(pos[sl] - com[g])
        r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, min(2, dim - 1)] * d[:, min(2, dim - 1)]
        rad[g] = np.sqrt(T(np.sum((r2 * m).astype(np.float64))) / gm[g])
    out["mass"] = gm if mass is not None else None
    out["com_pos"], out["com_inertia_radius"] = com, rad
    out["com_vel"] = cvel if vel is not None else None
    out["v_rad"] = vr if v_rad is not None else None
    out["scale_factor"] = sf if zplus1 is not None else None
    return out


# ------------------------------------------------------------------------------------------------
# CenterOfMass (tree.cuh:378-429) and the mass-centred planes (tree.py:736-754)
# ------------------------------------------------------------------------------------------------

def center_of_mass(ispl, pos, mass, kahan=True):
    """Mass centre and total mass of the points in every range [ispl[i], ispl[i+1]) (C restatement
    of tree.cuh:378-429; mass broadcast as tree.py:488).  Returns (centres (n, dim), masses (n,))."""
    pos = np.ascontiguousarray(pos)
    dt = np.float64 if pos.dtype == np.float64 else np.float32
    pos = np.ascontiguousarray(pos, dt)
    m = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, dt), pos.shape[:-1]))
    isp = np.ascontiguousarray(ispl, np.int32)
    n, dim = len(isp) - 1, pos.shape[1]
    out = np.empty((max(n, 0), dim + 1), dt)
    fn = lib().orc_center_of_mass_f64 if dt == np.float64 else lib().orc_center_of_mass_f32
    fn(isp, pos, m, n, dim, int(bool(kahan)), out)
    return out[:, :dim], out[:, dim]

