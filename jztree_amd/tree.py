"""Z-order sort and tree hierarchy (the API of src/jztree/tree.py, single GPU).

Every device computation of the reference's FFI targets runs in libjztree_amd.so (HIP, gfx950);
the reference's jnp glue between them (compactions, prefix sums, scatters into PackedArrays) is
written with torch GPU ops that never synchronise with the host.  Function names, arguments and
outputs follow the reference; each docstring cites the lines it mirrors.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Any, Tuple

import numpy as np
import torch

from . import _lib as L
from .checks import raise_if
from .config import TreeConfig, RegularizationConfig
from .data import (InteractionList, RankIdx, LevelInfo, PackedArray, Pos, TreeHierarchy, get_num, get_num_total, get_pos,
                   same_width_int, tree_map_by_len, verify_ilist)
from .tools import cumsum_starting_with_zero, div_ceil


# level of a boundary between two sets of a batch: above every Morton level
LVL_SET = 1 << 16


def _i32(x, device) -> torch.Tensor:
    """A (1,) int32 device tensor holding x (python int or device scalar); written by a fill
    kernel, never copied from the host, so that the pipeline can be captured as a HIP graph."""
    if torch.is_tensor(x):
        return x.reshape(1).to(device=device, dtype=torch.int32)
    return torch.full((1,), int(x), dtype=torch.int32, device=device)


def _i32_list(vals, device) -> torch.Tensor:
    """int32 device tensor of python ints, filled on the device (capture-safe)."""
    out = torch.empty(len(vals), dtype=torch.int32, device=device)
    for i, v in enumerate(vals):
        out[i:i + 1].fill_(int(v))
    return out


def _tmp(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


# ------------------------------------------------------------------------------------------------
# z-order sort (tree.py:75-141)
# ------------------------------------------------------------------------------------------------

def _zsort_impl(x: torch.Tensor, block_size: int = 64):
    """tree.py:75-93 -> PosZorderSort."""
    assert x.ndim == 2
    n, dim = x.shape
    assert 2 <= dim <= 10, "positions of 2 to 10 dimensions"
    x = x.contiguous()
    pos = torch.empty_like(x)
    ids = torch.empty(n, dtype=same_width_int(x.dtype), device=x.device)
    if n == 0:
        return pos, ids
    dt = L.sort_dtype_code(x.dtype)
    nb = L.LIB.jzt_zsort_tmp_bytes(n, dim, dt)
    tmp = _tmp(nb, x.device)
    L.call("jzt_pos_zorder_sort", L.stream(), L.ptr(x), L.ptr(pos), L.ptr(ids), L.ptr(tmp), nb, n, dim, dt)
    return pos, ids


class _ZSort(torch.autograd.Function):
    """Custom VJP of tree.py:111-138: gradients scatter back through the inverse permutation."""

    @staticmethod
    def forward(ctx, x):
        pos, idz = _zsort_impl(x.detach())
        ctx.save_for_backward(idz)
        ctx.mark_non_differentiable(idz)
        return pos, idz

    @staticmethod
    def backward(ctx, gpos, gidz):
        (idz,) = ctx.saved_tensors
        idinv = torch.zeros_like(idz)
        idinv[idz.long()] = torch.arange(len(idz), dtype=idz.dtype, device=idz.device)
        return gpos[idinv.long()]


def zsort(pos):
    """Brings position vectors into z-order (tree.py:95-141).

    Args:
        pos: (N, dim) tensor, or a particle dataclass with a .pos attribute; every tensor field
            with leading length N is sorted along with it.
    Returns:
        (posz, idz) with posz = pos[idz].
    """
    assert get_pos(pos).ndim == 2, "positions must have shape (N,3)"
    if torch.is_tensor(pos):
        return _ZSort.apply(pos)
    posz, idz = _ZSort.apply(pos.pos)
    out = tree_map_by_len(lambda v: v[idz.long()], pos, len(posz))
    try:
        out.pos = posz
    except AttributeError:
        pass
    return out, idz


zsort.jit = zsort


# ------------------------------------------------------------------------------------------------
# leaves and nodes (tree.py:160-329)
# ------------------------------------------------------------------------------------------------

def find_regularization_level(lvl: torch.Tensor, linfo: LevelInfo, weights: torch.Tensor | None = None,
                              max_vol_fac: float = 50., min_percentile: float = 90.) -> torch.Tensor:
    """tree.py:160-187 (with _level_histogram :55-68) -> device int32 (1,) level."""
    dev = lvl.device
    out = torch.empty(1, dtype=torch.int32, device=dev)
    nb = L.LIB.jzt_regularization_tmp_bytes(linfo.min_lvl(), linfo.max_lvl())
    tmp = _tmp(nb, dev)
    lvl = lvl.to(torch.int32).contiguous()
    w = None if weights is None else weights.to(torch.int32).contiguous()
    L.call("jzt_regularization_level", L.stream(), L.ptr(lvl), L.ptr(w), lvl.numel(), L.ptr(out),
           linfo.min_lvl(), linfo.max_lvl(), float(max_vol_fac), float(min_percentile), L.ptr(tmp), nb)
    return out


def _where_flags(flags: torch.Tensor, size: int, fill) -> Tuple[torch.Tensor, torch.Tensor]:
    """jnp.where(flags, size=size, fill_value=fill)[0] as int32, plus the device count."""
    dev = flags.device
    n = flags.numel()
    out = torch.empty(size, dtype=torch.int32, device=dev)
    cnt = torch.empty(1, dtype=torch.int32, device=dev)
    nb = L.LIB.jzt_where_tmp_bytes(n)
    tmp = _tmp(nb, dev)
    fill_dev = fill.reshape(1).to(torch.int32) if torch.is_tensor(fill) else None
    fill_host = 0 if torch.is_tensor(fill) else int(fill)
    L.call("jzt_where_flags", L.stream(), L.ptr(flags.contiguous()), n, L.ptr(out), size, fill_host,
           L.ptr(fill_dev), L.ptr(cnt), L.ptr(tmp), nb)
    return out, cnt[0]


def detect_leaf_boundaries(posz: torch.Tensor, ptype: torch.Tensor | None = None, num_types: int | None = None,
                           leaf_size: int = 32, npart=None, lvl_bound=None, block_size: int = 64,
                           alloc_size: int | None = None, cfg_reg: RegularizationConfig | None = None,
                           set_id: torch.Tensor | None = None) -> torch.Tensor:
    """Leaf splits of z-sorted positions (tree.py:189-262): leaves are the largest binary Morton
    cells holding no more than leaf_size particles of any type.

    set_id (extension): int32 set of every row of a batch of independent sets (non-decreasing);
    set boundaries are always splits and no count runs across them."""
    n, dim = posz.shape
    dev = posz.device
    if alloc_size is None:
        alloc_size = int(div_ceil(n, max(leaf_size // 2, 1))) + 1
    if npart is None:
        npart = n
    linfo = LevelInfo(dim, posz.dtype)
    if lvl_bound is None:
        lvl_bound = (linfo.max_lvl(), linfo.max_lvl())
    lvl_bound_t = (lvl_bound.to(device=dev, dtype=torch.int32).reshape(2) if torch.is_tensor(lvl_bound)
                   else _i32_list(lvl_bound, dev))
    npart_t = _i32(npart, dev)
    if ptype is None:
        num_types = 1
        ptype_t = None
    else:
        assert num_types is not None, "Please provide num_types if providing ptype"
        assert len(ptype) == len(posz)
        ptype_t = ptype.to(torch.uint8).contiguous()
    flags = torch.empty(n + 1, dtype=torch.int8, device=dev)
    lvl = torch.empty(n + 1, dtype=torch.int32, device=dev)
    if set_id is not None:
        assert num_types == 1 and cfg_reg is None, "batches of sets use one particle type and no regularization"
        assert set_id.dtype == torch.int32 and len(set_id) == n
        L.call("jzt_flag_leaf_boundaries_sets", L.stream(), L.ptr(posz.contiguous()), L.ptr(set_id.contiguous()),
               L.ptr(lvl_bound_t), L.ptr(npart_t), L.ptr(flags), L.ptr(lvl), n, dim, L.dtype_code(posz.dtype),
               int(leaf_size), int(leaf_size + 1))
    else:
        L.call("jzt_flag_leaf_boundaries", L.stream(), L.ptr(posz.contiguous()), L.ptr(ptype_t), L.ptr(lvl_bound_t),
               L.ptr(npart_t), L.ptr(flags), L.ptr(lvl), n, dim, L.dtype_code(posz.dtype), int(leaf_size),
               int((leaf_size + 1) * num_types), int(num_types))

    if cfg_reg is not None:
        splits, _ = _where_flags(flags, alloc_size, npart_t)
        sl = splits.long()
        leaf_npart = splits[1:] - splits[:-1]
        leaf_lvl = torch.minimum(lvl[sl[1:]], lvl[sl[:-1]]) - 1
        lvl_max = find_regularization_level(leaf_lvl, linfo, weights=leaf_npart,
                                            max_vol_fac=cfg_reg.max_volume_fac,
                                            min_percentile=cfg_reg.regularize_percentile)
        L.call("jzt_flags_or_level", L.stream(), L.ptr(flags), L.ptr(lvl), n + 1,
               L.ptr(lvl_max.to(torch.int32).reshape(1).contiguous()), L.ptr(npart_t.reshape(1).contiguous()))

    splits, nfilled = _where_flags(flags, alloc_size, npart_t)
    raise_if(nfilled > alloc_size,
             "Tree-Leaf allocation too small: filled={filled} size={size}.\n"
             "Hint: Increase alloc_fac_nodes or max_leaf_size", filled=nfilled, size=alloc_size)
    return splits


detect_leaf_boundaries.jit = detect_leaf_boundaries


def determine_znode_boundaries(posz: torch.Tensor, block_size: int = 64, nleaves=None,
                               set_id: torch.Tensor | None = None):
    """tree.py:264-284 -> FindNodeBoundaries.  Returns (lvl, lbound, rbound), each (len+1,).

    set_id (extension): set of every row of a batch; boundaries between sets get level LVL_SET."""
    size, dim = posz.shape
    dev = posz.device
    if nleaves is None:
        nleaves = size
    nl = torch.clamp(_i32(nleaves, dev), max=size)
    info = LevelInfo(dim, posz.dtype)
    pos_bound = torch.full((2, dim), float("nan"), dtype=posz.dtype, device=dev)
    lvl, lb, rb = (torch.empty(size + 1, dtype=torch.int32, device=dev) for _ in range(3))
    if set_id is not None:
        assert set_id.dtype == torch.int32 and len(set_id) == size
        L.call("jzt_find_node_boundaries_sets", L.stream(), L.ptr(posz.contiguous()), L.ptr(set_id.contiguous()),
               L.ptr(pos_bound), L.ptr(nl), L.ptr(lvl), L.ptr(lb), L.ptr(rb), size + 1, dim,
               L.dtype_code(posz.dtype), -2000)
    else:
        L.call("jzt_find_node_boundaries", L.stream(), L.ptr(posz.contiguous()), L.ptr(pos_bound), L.ptr(nl),
               L.ptr(lvl), L.ptr(lb), L.ptr(rb), size + 1, dim, L.dtype_code(posz.dtype), info.max_lvl(), -2000)
    return lvl, lb, rb


determine_znode_boundaries.jit = determine_znode_boundaries


def get_node_geometry(posz: torch.Tensor, lbound: torch.Tensor, rbound: torch.Tensor, num=None,
                      block_size: int = 64, result: str = "lvl_cent_ext", upper_extent: bool = False):
    """tree.py:286-329 -> GetNodeGeometry."""
    assert lbound.shape == rbound.shape
    assert lbound.dtype == rbound.dtype == torch.int32
    dev = posz.device
    n, dim = posz.shape
    size = lbound.shape[0]
    mode = (1 if "lvl" in result else 0) | (2 if "cent" in result else 0) | (4 if "ext" in result else 0)
    if num is None:
        num = size
    num_t = _i32(num, dev)
    lvl = torch.empty(size if mode & 1 else 1, dtype=torch.int32, device=dev)
    cent = torch.empty((size if mode & 2 else 1, dim), dtype=posz.dtype, device=dev)
    ext = torch.empty((size if mode & 4 else 1, dim), dtype=posz.dtype, device=dev)
    L.call("jzt_get_node_geometry", L.stream(), L.ptr(posz.contiguous()), L.ptr(lbound.contiguous()),
           L.ptr(rbound.contiguous()), L.ptr(num_t), L.ptr(lvl), L.ptr(cent), L.ptr(ext), size, n, dim,
           L.dtype_code(posz.dtype), -2000, mode, int(bool(upper_extent)))
    rdict = {"lvl": lvl, "cent": cent, "ext": ext}
    res = [rdict[k] for k in result.split("_")]
    return res[0] if len(res) == 1 else tuple(res)


get_node_geometry.jit = get_node_geometry


# ------------------------------------------------------------------------------------------------
# hierarchy (tree.py:600-829)
# ------------------------------------------------------------------------------------------------

def _estimate_node_number(npart, max_node_size, alloc_fac=1.):
    """tree.py:600-609"""
    return int(div_ceil(npart * alloc_fac, np.maximum(max_node_size // 2, 1))) + 1


def _define_tree_level_node_sizes(npart: int, cfg_tree: TreeConfig):
    """tree.py:611-619"""
    max_num_leaves = _estimate_node_number(npart, cfg_tree.max_leaf_size)
    nlevels = np.log(max_num_leaves / cfg_tree.stop_coarsen) / np.log(cfg_tree.coarse_fac)
    nlevels = np.maximum(int(np.ceil(nlevels)), 1)
    return [int(cfg_tree.max_leaf_size * (cfg_tree.coarse_fac ** i)) for i in range(0, nlevels)]


def _scatter_drop(size: int, idx: torch.Tensor, values: torch.Tensor, dtype) -> torch.Tensor:
    """zeros(size).at[idx].set(values) with out-of-range indices dropped (== size here)."""
    out = torch.zeros(size + 1, dtype=dtype, device=idx.device)
    out[torch.clamp(idx, 0, size).reshape(-1)] = values.reshape(-1).to(dtype)
    return out[:size]


def _delete_rows(x: torch.Tensor, idx: torch.Tensor) -> torch.Tensor:
    """jnp.delete(x, idx, axis=0, assume_unique_indices=True) with device indices."""
    n = x.shape[0]
    m = idx.numel()
    mark = torch.zeros(n, dtype=torch.int64, device=x.device)
    # in range even when an (already recorded) allocation overflow made the indices meaningless
    mark.index_fill_(0, torch.clamp(idx.long(), 0, n - 1).reshape(-1), 1)  # scalar as kernel argument
    dest = torch.arange(n, device=x.device) - (torch.cumsum(mark, 0) - mark)
    dest = torch.where(mark.bool(), torch.full_like(dest, n - m), dest)
    out = torch.empty((n - m + 1,) + tuple(x.shape[1:]), dtype=x.dtype, device=x.device)
    out[dest] = x
    return out[:n - m]


def _splits_per_type(spl: torch.Tensor, ptype: torch.Tensor, num_types: int, npart) -> torch.Tensor:
    """tree.py:621-626: per type t, the number of filled type-t particles before each split."""
    n = len(ptype)
    filled = torch.arange(n, device=ptype.device) < (npart if npart is not None else n)
    out = []
    for t in range(num_types):
        prefix = cumsum_starting_with_zero(((ptype == t) & filled).to(torch.int32))
        out.append(prefix[spl.long().clamp(0, n)])
    return torch.stack(out).to(torch.int32)


def _define_split_hierarchy(posz, node_sizes, alloc_size, npart=None, ptype=None, num_types=None, lvl_bound=None,
                            cfg_reg: RegularizationConfig | None = None, set_id: torch.Tensor | None = None):
    """tree.py:628-734: leaf splits, then node activity per plane packed into PackedArrays.

    set_id (extension): batches of sets; boundaries between sets split every plane."""
    n, dim = posz.shape
    dev = posz.device
    nlevels = len(node_sizes)
    linfo = LevelInfo(dim, posz.dtype)
    if lvl_bound is None:
        lvl_bound = (linfo.max_lvl(), linfo.max_lvl())

    ispl = detect_leaf_boundaries(posz, ptype=ptype, num_types=num_types, leaf_size=node_sizes[0], npart=npart,
                                  alloc_size=alloc_size, lvl_bound=lvl_bound, cfg_reg=cfg_reg, set_id=set_id)
    # first index of the largest split (= the leaf count); the splits never decrease, so a search
    # finds it without a multi-block reduction (which would record a memset in a captured graph)
    nleaves = torch.searchsorted(ispl, ispl[-1:]).reshape(()).to(torch.int32)
    first = torch.clamp(ispl[:-1].long(), max=n - 1)
    reps = posz[first]
    leaf_set = None if set_id is None else set_id[first].contiguous()
    lvl, lbound, rbound = determine_znode_boundaries(reps, nleaves=nleaves, set_id=leaf_set)
    lbl, rbl = lbound.long(), rbound.long()
    if ptype is not None:
        # several types: a node may not hold more than the node size of any one type (tree.py:661-664)
        ispl_per_type = _splits_per_type(ispl, ptype, num_types, npart)
        npins = torch.amax(ispl_per_type[:, rbl] - ispl_per_type[:, lbl], dim=0)
    else:
        npins = ispl[rbl] - ispl[lbl]
        ispl_per_type = ispl.reshape((1,) + ispl.shape)

    sizes = _i32_list(node_sizes, dev)
    isl = npins[None, :] > sizes[:, None]
    isl = isl | ((lbound <= 0) & (lvl > lvl_bound[0]))[None, :]
    isl = isl | ((rbound >= nleaves) & (lvl > lvl_bound[1]))[None, :]
    if set_id is not None:
        # set boundaries split every plane; so do the two ends of the array, which a single-set
        # tree only splits through the size or level tests (both can miss a small end set, whose
        # top-plane node would then never be closed)
        ends = torch.arange(len(lvl), device=dev)
        isl = isl | ((lvl >= LVL_SET) | (ends == 0) | (ends == nleaves))[None, :]

    if cfg_reg is not None:
        new = [(npins > 0) | isl[0]]
        lvl_up = torch.minimum(lvl[lbl], lvl[rbl]) - 1
        ar = torch.arange(len(lvl), device=dev)
        for i in range(1, nlevels):
            is_node = (~isl[i]) & isl[i, lbl] & isl[i, rbl]
            is_node = is_node & (ar < nleaves)
            lvl_max = find_regularization_level(lvl_up, linfo, weights=torch.where(is_node, npins, 0),
                                                max_vol_fac=cfg_reg.max_volume_fac,
                                                min_percentile=cfg_reg.regularize_percentile)
            new.append(isl[i] | ((lvl > lvl_max) & (ar <= nleaves)))
        isl = torch.stack(new)

    offsets = torch.cumsum(isl.reshape(-1).to(torch.int64), 0).reshape(isl.shape)
    level_spl = torch.cat([torch.zeros(1, dtype=torch.int64, device=dev), offsets[:, -1]]).to(torch.int32)
    nnodes_on_level = level_spl[1:] - level_spl[:-1] - 1
    raise_if(level_spl[-1] > alloc_size,
             "Tree-node allocation too small: filled={filled} size={size}.\n"
             "Hint: Increase alloc_fac_nodes, max_leaf_size or coarse_fac", filled=level_spl[-1], size=alloc_size)
    offsets = torch.where(isl, offsets - isl.to(torch.int64), torch.full_like(offsets, alloc_size))

    ispl_n2l = _scatter_drop(alloc_size, offsets, offsets[0:1, :].expand_as(offsets), torch.int32)
    ispl_n2l = PackedArray.from_data(ispl_n2l, level_spl, fill_values=nnodes_on_level[0].expand(nlevels))

    value = torch.empty_like(offsets)
    value[0] = ispl.to(torch.int64)
    for li in range(1, nlevels):
        value[li] = offsets[li - 1] - level_spl[li - 1]
    ispl_n2n = _scatter_drop(alloc_size, offsets, value, torch.int32)
    fill_val = torch.cat([ispl[-1:], nnodes_on_level[:-1]])
    ispl_n2n = PackedArray.from_data(ispl_n2n, level_spl, fill_values=fill_val)
    if set_id is not None:
        return ispl, ispl_per_type, ispl_n2l, ispl_n2n, leaf_set
    return ispl, ispl_per_type, ispl_n2l, ispl_n2n


def center_of_mass(ispl: torch.Tensor, pos: torch.Tensor, mass, kahan_summation: bool = True):
    """tree.py:479-499 -> CenterOfMass: mass-weighted centre and total mass of the points in every
    range [ispl[i], ispl[i+1]) (NaN for empty ranges).  mass: per point, or a scalar (broadcast as
    in the reference).  Returns (pos (n, dim), mass (n,))."""
    assert ispl.dtype == torch.int32 and pos.ndim == 2
    dev = pos.device
    n = ispl.numel() - 1
    m = torch.as_tensor(mass, dtype=pos.dtype, device=dev)
    m = torch.broadcast_to(m, pos.shape[:-1]).contiguous()
    out = torch.empty((max(n, 0), pos.shape[-1] + 1), dtype=pos.dtype, device=dev)
    L.call("jzt_center_of_mass", L.stream(), L.ptr(ispl.contiguous()), L.ptr(pos.contiguous()), L.ptr(m),
           L.ptr(out), n, int(bool(kahan_summation)), pos.shape[-1], L.dtype_code(pos.dtype))
    return out[:, :-1], out[:, -1]


def _get_tree_mass_centers(partz, ispl_n2n: PackedArray, prop_array_spl: torch.Tensor):
    """tree.py:736-754: node mass centres and masses of every plane; plane 0 from the particles,
    plane p from the mass centres and masses of plane p-1."""
    pos = get_pos(partz)
    if getattr(partz, "mass", None) is None:
        raise ValueError("To use mass centering, please provide PosMass input")
    cpos, cmass = center_of_mass(ispl_n2n.get(0).to(torch.int32), pos, partz.mass)
    node_cent = PackedArray.from_data(cpos, ispl=prop_array_spl, fill_values=float("nan"))
    node_mass = PackedArray.from_data(cmass, ispl=prop_array_spl, fill_values=float("nan"))
    for i in range(1, ispl_n2n.nlevels()):
        cpos, cmass = center_of_mass(ispl_n2n.get(i, size=len(cmass) + 1).to(torch.int32), cpos.contiguous(),
                                     cmass.contiguous())
        node_mass = node_mass.set(i, cmass, num=ispl_n2n.num(i) - 1)
        node_cent = node_cent.set(i, cpos, num=ispl_n2n.num(i) - 1)
    return node_cent, node_mass


def build_tree_hierarchy(partz, cfg_tree: TreeConfig, lvl_bound=None, ptype=None, num_types=None,
                         set_id: torch.Tensor | None = None, nsets: int = 1) -> TreeHierarchy:
    """Builds a tree hierarchy from z-order positions (tree.py:756-829).

    Plane 0 holds leaves of at most cfg_tree.max_leaf_size particles; plane p nodes hold at most
    max_leaf_size * coarse_fac**p particles.  With cfg_tree.mass_centered, partz carries masses
    (PosMass) and th.mass / th.mass_cent hold every node's mass and mass centre (CenterOfMass).

    set_id, nsets (extension): a batch of nsets independent sets, set_id (int32, non-decreasing)
    naming the set of every row; the result is a forest whose nodes never span two sets, and
    the top-plane interaction list pairs only nodes of one set.  Regularization is skipped for
    batches (it never changes kNN or FoF results, only the leaf shapes)."""
    posz = get_pos(partz)
    npart_tot = get_num_total(partz, default_to_length=True)
    np_per_dev = npart_tot
    npart = get_num(partz, default_to_nancount=True)
    node_sizes = _define_tree_level_node_sizes(npart_tot, cfg_tree)
    alloc_size = _estimate_node_number(np_per_dev, cfg_tree.max_leaf_size, cfg_tree.alloc_fac_nodes)

    if set_id is None:
        ispl_l2p, ispl_per_type, ispl_n2l, ispl_n2n = _define_split_hierarchy(
            posz, node_sizes, alloc_size, npart=npart, ptype=ptype, num_types=num_types, lvl_bound=lvl_bound,
            cfg_reg=cfg_tree.regularization)
        leaf_set = None
    else:
        ispl_l2p, ispl_per_type, ispl_n2l, ispl_n2n, leaf_set = _define_split_hierarchy(
            posz, node_sizes, alloc_size, npart=npart, ptype=ptype, num_types=num_types, lvl_bound=lvl_bound,
            cfg_reg=None, set_id=set_id)

    ispl_n2p = ispl_l2p[ispl_n2l.data.long()]
    lvl, geom_cent = get_node_geometry(posz, ispl_n2p[:-1], ispl_n2p[1:], num=ispl_n2l.nfilled() - 1,
                                       result="lvl_cent")
    seams = ispl_n2l.ispl[1:-1] - 1
    lvl = _delete_rows(lvl, seams)
    geom_cent = _delete_rows(geom_cent, seams)
    prop_array_spl = cumsum_starting_with_zero(ispl_n2n.ispl[1:] - ispl_n2n.ispl[:-1] - 1)
    lvl = PackedArray.from_data(lvl, prop_array_spl, fill_values=-2000)
    geom_cent = PackedArray.from_data(geom_cent, prop_array_spl, fill_values=float("nan"))
    nmass_cent, nmass = None, None
    if cfg_tree.mass_centered:
        nmass_cent, nmass = _get_tree_mass_centers(partz, ispl_n2n, prop_array_spl)
    size_leaves = _estimate_node_number(np_per_dev, cfg_tree.max_leaf_size)
    return TreeHierarchy(size_leaves, ispl_n2n, ispl_n2l, ispl_per_type, lvl, geom_cent, mass=nmass,
                         mass_cent=nmass_cent, leaf_set=leaf_set, nsets=int(nsets))


build_tree_hierarchy.jit = build_tree_hierarchy


def zsort_and_tree(part, cfg_tree: TreeConfig = TreeConfig(), data: Any | None = None, ptype=None,
                   num_types=None, shrink: bool = True):
    """tree.py:363-421 (single GPU): z-sort (carrying `data`) and build the hierarchy.

    Returns (partz, th), or (partz, dataz, th) when data is given."""
    partz, isort = zsort(part)
    dtz = tree_map_by_len(lambda v: v[isort.long()], (data, ptype), len(get_pos(partz)))
    th = build_tree_hierarchy(partz, cfg_tree, ptype=dtz[1], num_types=num_types)
    out = [partz]
    if (data is not None) or not shrink:
        out.append(dtz[0])
    if not shrink:
        out.append(dtz[1])
    return (*out, th)


zsort_and_tree.jit = zsort_and_tree


def zsort_and_tree_multi_type(part: Tuple, cfg_tree: TreeConfig = TreeConfig(), data: Any | None = None):
    """tree.py:423-477 (single GPU): one tree over several particle types (e.g. sources and
    queries); leaves hold at most max_leaf_size particles of each type and
    th.ispl_l2p_per_type[t] splits the z-ordered particles of type t.  Returns (partz, th) or
    (partz, dataz, th), tuples with one entry per type, each type's particles in z-order."""
    num_types = len(part)
    sizes = [len(get_pos(p)) for p in part]
    dev = get_pos(part[0]).device
    if data is None:
        data = tuple(None for _ in part)
    # filled rows of each type first (rows with a NaN position last), as concatenate_pytrees does
    valid = torch.cat([~torch.isnan(get_pos(p)[:, 0]) for p in part])
    compress = torch.argsort((~valid).to(torch.int8), stable=True)
    ptype = torch.repeat_interleave(torch.arange(num_types, dtype=torch.uint8, device=dev),
                                    torch.tensor(sizes, device=dev))[compress]

    def stack(items):
        first = items[0]
        if torch.is_tensor(first):
            return torch.cat(items)[compress]
        if isinstance(first, RankIdx):
            return RankIdx(rank=None, idx=torch.cat([it.idx for it in items])[compress])
        if dataclasses.is_dataclass(first):
            vals = {}
            for f in dataclasses.fields(first):
                vs = [getattr(it, f.name) for it in items]
                if torch.is_tensor(vs[0]) and vs[0].ndim > 0 and len(vs[0]) == len(get_pos(first)):
                    vals[f.name] = torch.cat(vs)[compress]
                else:
                    vals[f.name] = vs[0]
            if "num" in vals:
                vals["num"] = None
            if "num_total" in vals:
                vals["num_total"] = None
            return dataclasses.replace(first, **vals)
        raise TypeError(f"cannot stack {type(first)}")

    part_stacked = stack(list(part))
    data_stacked = stack(list(data)) if data[0] is not None else None
    partz_s, dataz_s, ptypez, th = zsort_and_tree(part_stacked, cfg_tree=cfg_tree, data=data_stacked, ptype=ptype,
                                                  num_types=num_types, shrink=False)
    # separate the types, each in z-order (a stable sort by type)
    order = torch.argsort(ptypez.to(torch.int8), stable=True)
    out_p, out_d = [], []
    off = 0
    for t, sz in enumerate(sizes):
        sel = order[off:off + sz]
        off += sz
        out_p.append(tree_map_by_len(lambda v: v[sel], partz_s, len(ptypez)))
        out_d.append(None if dataz_s is None else tree_map_by_len(lambda v: v[sel], dataz_s, len(ptypez)))
    if data[0] is not None:
        return tuple(out_p), tuple(out_d), th
    return tuple(out_p), th


zsort_and_tree_multi_type.jit = zsort_and_tree_multi_type


def search_sorted_z(posz: torch.Tensor, posz_query: torch.Tensor, block_size: int = 64, leaf_search: bool = False):
    """tree.py:143-158: indices where the points of posz_query would be inserted into the z-sorted
    posz to keep the z-order (posz[idx-1] < x <= posz[idx]); with leaf_search, posz holds one point
    per leaf and the result is the leaf each query point belongs to (SearchSortedZ)."""
    assert posz.dtype == posz_query.dtype
    assert posz.shape[-1] == posz_query.shape[-1]
    out = torch.empty(len(posz_query), dtype=torch.int32, device=posz.device)
    L.call("jzt_search_sorted_z", L.stream(), L.ptr(posz.contiguous()), L.ptr(posz_query.contiguous()), len(posz),
           len(posz_query), posz.shape[-1], L.sort_dtype_code(posz.dtype), int(bool(leaf_search)), L.ptr(out))
    return out


search_sorted_z.jit = search_sorted_z


def zsort_and_tree_origin(part, cfg_tree: TreeConfig = TreeConfig()):
    """zsort_and_tree(part, cfg_tree, data=RankIdx(rank=None, idx=arange(n) int32)), without the
    gather of the arange: its z-ordered copy is the sort permutation itself.
    Returns (partz, originz, th)."""
    partz, isort = zsort(part)
    th = build_tree_hierarchy(partz, cfg_tree)
    return partz, RankIdx(rank=None, idx=isort.to(torch.int32)), th


def f32_view(th: TreeHierarchy) -> TreeHierarchy:
    """The same hierarchy (splits, planes) with float32 node geometry for float32 distance kernels
    on the float32-rounded particles (extension: float64 keys, float32 distances).  Each node's
    Morton cell is widened to the smallest enclosing cell whose boundaries are float32 numbers
    (jzt_geometry_to_f32), which contains the rounded particles of the node, so every pruning
    decision of the float32 node kernels stays exact for the rounded points.  A float32 tree is
    returned unchanged."""
    if th.geom_cent.data.dtype == torch.float32:
        return th
    g = th.geom_cent.data.to(torch.float64).contiguous()
    lv = th.lvl.data.to(torch.int32).contiguous()
    n, dim = g.shape
    cent = torch.empty((n, dim), dtype=torch.float32, device=g.device)
    lvl = torch.empty(n, dtype=torch.int32, device=g.device)
    L.call("jzt_geometry_to_f32", L.stream(), L.ptr(g), L.ptr(lv), n, dim, -2000, L.ptr(cent), L.ptr(lvl))
    fill = th.geom_cent.fill_values
    geom = dataclasses.replace(th.geom_cent, data=cent, fill_values=None if fill is None else fill.float())
    lvlp = dataclasses.replace(th.lvl, data=lvl.to(th.lvl.data.dtype))
    return dataclasses.replace(th, geom_cent=geom, lvl=lvlp)


# ------------------------------------------------------------------------------------------------
# interaction-list initialisation (tree.py:868-930)
# ------------------------------------------------------------------------------------------------

def grouped_dense_interaction_list(nnodes, size_ilist: int, ngroup: int = 32, size_super: int | None = None,
                                   node_range=None, dtype=torch.int32, with_r2: bool = False,
                                   node_set: torch.Tensor | None = None, nsets: int = 1):
    """All-to-all interaction list over super-nodes of `ngroup` top-level nodes (tree.py:868-930),
    built on the device (jzt_dense_ilist); only the nsuper^2 live entries are written.

    node_set, nsets (extension): set of every top node of a batch; super-nodes then never span
    two sets and interact only within their set (jzt_dense_ilist_sets).

    Returns (spl_super, ilist, nsuper_nodes); ilist.rad2 = zeros when with_r2."""
    if node_range is not None:
        raise NotImplementedError("node_range is a multi-GPU feature")
    assert dtype == torch.int32
    dev = nnodes.device if torch.is_tensor(nnodes) else torch.device("cuda")
    nn = _i32(nnodes, dev)
    if size_super is None:
        size_super = int(np.ceil(np.sqrt(size_ilist))) + 2
    spl_super = torch.empty(size_super + 1, dtype=torch.int32, device=dev)
    ispl = torch.empty(size_super + 1, dtype=torch.int32, device=dev)
    ilist = torch.empty(size_ilist, dtype=torch.int32, device=dev)
    r2 = torch.empty(size_ilist, dtype=torch.float32, device=dev) if with_r2 else None
    nsuper = torch.empty(1, dtype=torch.int32, device=dev)
    if node_set is not None:
        nb = L.LIB.jzt_dense_ilist_sets_tmp_bytes(int(nsets))
        tmp = _tmp(nb, dev)
        L.call("jzt_dense_ilist_sets", L.stream(), L.ptr(nn), L.ptr(node_set.to(torch.int32).contiguous()),
               int(nsets), int(ngroup), L.ptr(spl_super), L.ptr(ispl), size_super, L.ptr(ilist), L.ptr(r2),
               size_ilist, L.ptr(nsuper), L.ptr(tmp), nb)
        # the list length itself: sum over sets of nsuper_s^2 (saturated at 2^31-1)
        nint = tmp.view(torch.int64)[3 * int(nsets) + 2]
    else:
        L.call("jzt_dense_ilist", L.stream(), L.ptr(nn), int(ngroup), L.ptr(spl_super), L.ptr(ispl), size_super,
               L.ptr(ilist), L.ptr(r2), size_ilist, L.ptr(nsuper))
        nint = nsuper[0].to(torch.int64) ** 2
    raise_if(nint > size_ilist,
             "Cannot fit {n} interactions into interaction list with size {size}\nHint: Increase alloc_fac_ilist",
             n=nint, size=size_ilist)
    return spl_super, verify_ilist(InteractionList(ispl=ispl, isrc=ilist, rad2=r2)), nsuper[0]


grouped_dense_interaction_list.jit = grouped_dense_interaction_list


def top_dense_interaction_list(th: TreeHierarchy, size_ilist: int, ngroup: int = 32, with_r2: bool = False):
    """grouped_dense_interaction_list over the top plane of `th` (per set for a batch forest)."""
    nplanes = th.num_planes()
    size = th.size()
    if th.leaf_set is None:
        return grouped_dense_interaction_list(th.lvl.num(nplanes - 1), size_ilist, ngroup=ngroup, size_super=size,
                                              with_r2=with_r2)
    first_leaf = torch.clamp(th.ispl_n2l.get(nplanes - 1, size + 1)[:-1].long(), max=len(th.leaf_set) - 1)
    return grouped_dense_interaction_list(th.lvl.num(nplanes - 1), size_ilist, ngroup=ngroup, size_super=size,
                                          with_r2=with_r2, node_set=th.leaf_set[first_leaf], nsets=th.nsets)
