"""Batches of independent point sets through one pipeline (extension).

The reference batches only through vmap with vmap_method="sequential" (tree.py:88, :223, :962),
which runs the sets one after another, and its kNN and FoF are not batched at all.  Here the sets
of a batch are concatenated and go through the single-set pipeline once:

  jzt_pos_zorder_sort_segmented   every set z-ordered inside its own rows, by its own bounds
  jzt_flag_leaf_boundaries_sets   leaves, never across a set boundary
  jzt_find_node_boundaries_sets   nodes, never across a set boundary (set boundaries split every plane)
  jzt_dense_ilist_sets            top-plane lists pairing only nodes of one set
  walks and leaf kernels          unchanged: one launch per stage for the whole batch, the leaf
                                  kernels striding over the leaves of all sets

so every launch covers all sets, and the lists a walk derives from the top-plane lists stay inside
one set.  Results are those of a single-set run on each set: neighbour and group indices are
set-local row indices and the values are bit-identical (kNN is exact with ties by index, FoF
labels are the first z-order member of each group, and both the z-order and the distances inside
a set are the single-set ones).

Sets are passed as one (N, dim) tensor plus per-set sizes (or the N+1 offsets).
"""
from __future__ import annotations

import dataclasses

import torch

from . import _lib as L
from . import checks
from .config import FofConfig, KNNConfig
from .fof import _fof_dual_walk, _fof_leaf2leaf
from .knn import _knn_dual_walk, _knn_leaf2leaf
from .tree import _tmp, build_tree_hierarchy

__all__ = ["set_offsets", "set_ids", "zsort_sets", "tree_sets", "knn_sets", "fof_sets"]


def set_offsets(sizes, device=None) -> torch.Tensor:
    """int64 (nsets+1,) offsets of sets of the given sizes (a device tensor of offsets passes)."""
    if torch.is_tensor(sizes) and sizes.dtype == torch.int64 and sizes.is_cuda:
        return sizes
    sz = torch.as_tensor(sizes, dtype=torch.int64)
    off = torch.zeros(len(sz) + 1, dtype=torch.int64)
    off[1:] = torch.cumsum(sz, 0)
    return off.to(device if device is not None else "cuda")


def set_ids(offsets: torch.Tensor, n: int) -> torch.Tensor:
    """int32 set of every row: s with offsets[s] <= row < offsets[s+1]."""
    rows = torch.arange(n, dtype=torch.int64, device=offsets.device)
    return torch.searchsorted(offsets[1:].contiguous(), rows, right=True).to(torch.int32)


def zsort_sets(pos: torch.Tensor, offsets: torch.Tensor):
    """Z-order every set inside its rows (jzt_pos_zorder_sort_segmented).

    Returns (posz, idz) with posz = pos[idz]; idz are rows of the batch (subtract the set's offset
    for set-local indices)."""
    n, dim = pos.shape
    nsets = len(offsets) - 1
    assert nsets >= 1
    dev = pos.device
    pos = pos.contiguous()
    posz = torch.empty_like(pos)
    ids = torch.empty(n, dtype=torch.int32 if pos.dtype == torch.float32 else torch.int64, device=dev)
    if n == 0:
        return posz, ids
    dt = L.dtype_code(pos.dtype)
    nb = L.LIB.jzt_zsort_segmented_tmp_bytes(n, nsets, dim, dt)
    tmp = _tmp(nb, dev)
    L.call("jzt_pos_zorder_sort_segmented", L.stream(), L.ptr(pos), L.ptr(posz), L.ptr(ids),
           L.ptr(offsets.contiguous()), nsets, L.ptr(tmp), nb, n, dim, dt)
    return posz, ids


def _batch_tree_cfg(cfg_tree):
    # regularization changes leaf shapes only, never results; a batch tree is built without it
    return dataclasses.replace(cfg_tree, regularization=None)


def tree_sets(pos: torch.Tensor, sizes, cfg_tree):
    """Z-sort a batch and build its forest.  Returns (posz, idz, seg, offsets, th)."""
    n = len(pos)
    off = set_offsets(sizes, pos.device)
    seg = set_ids(off, n)  # set ranges are the same before and after the segmented sort
    posz, idz = zsort_sets(pos, off)
    th = build_tree_hierarchy(posz, _batch_tree_cfg(cfg_tree), set_id=seg, nsets=len(off) - 1)
    return posz, idz, seg, off, th


def knn_sets(pos: torch.Tensor, sizes, k: int, boxsize: float | None = None, cfg: KNNConfig = KNNConfig()):
    """Exact kNN of every set of a batch against its own set.

    pos: (N, dim) rows of all sets; sizes: per-set sizes (or a device int64 offsets tensor).
    Returns (rnn, inn), each (N, k), rows in input order; inn are set-local row indices.  Row r of
    set s equals row r - offsets[s] of knn(pos[offsets[s]:offsets[s+1]], k, boxsize)."""
    try:
        posz, idz, seg, off, th = tree_sets(pos, sizes, cfg.tree)
        local = (idz.long() - off[seg.long()]).to(torch.int32)  # z row -> set-local input index
        ilist = _knn_dual_walk(th, k, boxsize=boxsize, alloc_fac_ilist=cfg.alloc_fac_ilist)
        spl = th.splits_leaf_to_part(ptype=0)
        rnn, inn = _knn_leaf2leaf(ilist, spl, posz, k=k, boxsize=boxsize, splQ=spl, xQ=posz, src_ids=local,
                                  dst_rows=idz.to(torch.int32))
        checks.flush()
    except BaseException:
        checks.clear()
        raise
    return rnn, inn


def fof_sets(pos: torch.Tensor, sizes, rlink: float, boxsize: float = 0., cfg: FofConfig = FofConfig(),
             output_order: str = "input"):
    """Friends-of-friends labels of every set of a batch, each set linked only within itself.

    Returns int32 labels (N,) with set-local values: with output_order="input" rows in input order
    and labels the set-local input index of the group's first z-order member (as
    fof_labels(..., output_order="input") on the set alone); with "z" rows in the batch's z-order
    and labels set-local z positions.  Also returns idz (z row -> batch input row) for "z"."""
    assert output_order in ("input", "z")
    try:
        posz, idz, seg, off, th = tree_sets(pos, sizes, cfg.tree)
        node_data, ilist = _fof_dual_walk(th, rlink=rlink, boxsize=boxsize, alloc_fac_ilist=cfg.alloc_fac_ilist)
        igz = _fof_leaf2leaf(node_data, ilist, posz, rlink=rlink, boxsize=boxsize)
        checks.flush()
    except BaseException:
        checks.clear()
        raise
    segl = seg.long()
    if output_order == "z":
        return (igz.long() - off[segl]).to(torch.int32), idz
    idl = idz.long()
    out = torch.empty_like(igz)
    # z row i holds input row idz[i]; its label is the input row of the group's first z member
    out[idl] = (idl[igz.long()] - off[segl]).to(igz.dtype)
    return out
