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

Sets are passed as one (N, dim) tensor plus per-set sizes (or the N+1 offsets); their rows are all
points (no NaN padding rows inside a batch).
"""
from __future__ import annotations

import dataclasses

import torch

from . import _lib as L
from . import checks
from .config import FofConfig, KNNConfig
from .data import Pos
from .fof import _fof_dual_walk, _fof_leaf2leaf
from .graph import Captured
from .knn import _knn_dual_walk, _knn_leaf2leaf
from .tree import _tmp, build_tree_hierarchy

__all__ = ["set_offsets", "set_ids", "zsort_sets", "tree_sets", "knn_sets", "fof_sets", "pipeline_sets", "CapturedSets"]


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
    # every row is a point (no NaN padding inside sets): the count is known without a reduction
    th = build_tree_hierarchy(Pos(pos=posz, num=n), _batch_tree_cfg(cfg_tree), set_id=seg, nsets=len(off) - 1)
    return posz, idz, seg, off, th


def _knn_core(posz, idz, seg, off, th, k, boxsize, alloc_fac_ilist):
    local = (idz.long() - off[seg.long()]).to(torch.int32)  # z row -> set-local input index
    ilist = _knn_dual_walk(th, k, boxsize=boxsize, alloc_fac_ilist=alloc_fac_ilist)
    spl = th.splits_leaf_to_part(ptype=0)
    return _knn_leaf2leaf(ilist, spl, posz, k=k, boxsize=boxsize, splQ=spl, xQ=posz, src_ids=local,
                          dst_rows=idz.to(torch.int32))


def _r2_per_set(rlink, nsets: int, device):
    """float32 r2link per set (device) from one or nsets linking lengths, formed as fof_labels forms
    its r2link: float32(rlink^2) with the square in float64."""
    rl = torch.as_tensor(rlink, dtype=torch.float64).reshape(-1)
    if len(rl) == 1:
        rl = rl.expand(nsets)
    assert len(rl) == nsets, "one linking length, or one per set"
    return (rl * rl).to(torch.float32).to(device)


def _fof_core(posz, idz, seg, off, th, r2set, boxsize, alloc_fac_ilist, output_order="input"):
    node_data, ilist = _fof_dual_walk(th, rlink=0., boxsize=boxsize, alloc_fac_ilist=alloc_fac_ilist,
                                      r2link_set=r2set)
    leaf_r2 = r2set[th.leaf_set.long()].contiguous()
    igz = _fof_leaf2leaf(node_data, ilist, posz, rlink=0., boxsize=boxsize, leaf_r2link=leaf_r2)
    segl = seg.long()
    if output_order == "z":
        return (igz.long() - off[segl]).to(torch.int32)
    idl = idz.long()
    out = torch.empty_like(igz)
    # z row i holds input row idz[i]; its label is the input row of the group's first z member
    out[idl] = (idl[igz.long()] - off[segl]).to(igz.dtype)
    return out


def knn_sets(pos: torch.Tensor, sizes, k: int, boxsize: float | None = None, cfg: KNNConfig = KNNConfig()):
    """Exact kNN of every set of a batch against its own set.

    pos: (N, dim) rows of all sets; sizes: per-set sizes (or a device int64 offsets tensor).
    Returns (rnn, inn), each (N, k), rows in input order; inn are set-local row indices.  Row r of
    set s equals row r - offsets[s] of knn(pos[offsets[s]:offsets[s+1]], k, boxsize)."""
    try:
        posz, idz, seg, off, th = tree_sets(pos, sizes, cfg.tree)
        rnn, inn = _knn_core(posz, idz, seg, off, th, k, boxsize, cfg.alloc_fac_ilist)
        checks.flush()
    except BaseException:
        checks.clear()
        raise
    return rnn, inn


def fof_sets(pos: torch.Tensor, sizes, rlink, boxsize: float = 0., cfg: FofConfig = FofConfig(),
             output_order: str = "input"):
    """Friends-of-friends labels of every set of a batch, each set linked only within itself.

    rlink: one linking length for all sets, or one per set (a sequence or CPU tensor of nsets
    values).  r2link = float32(rlink^2) with the square in float64, as fof_labels forms it.
    Returns int32 labels (N,) with set-local values: with output_order="input" rows in input order
    and labels the set-local input index of the group's first z-order member (as
    fof_labels(..., output_order="input") on the set alone); with "z", rows in the batch's z-order,
    labels set-local z positions, and idz (z row -> batch input row) returned as well."""
    assert output_order in ("input", "z")
    try:
        posz, idz, seg, off, th = tree_sets(pos, sizes, cfg.tree)
        r2set = _r2_per_set(rlink, th.nsets, pos.device)
        lab = _fof_core(posz, idz, seg, off, th, r2set, boxsize, cfg.alloc_fac_ilist, output_order)
        checks.flush()
    except BaseException:
        checks.clear()
        raise
    return (lab, idz) if output_order == "z" else lab


def pipeline_sets(pos: torch.Tensor, sizes, k: int, rlink, boxsize: float = 0., cfg_tree=None,
                  alloc_fac_ilist_knn: float = 256., alloc_fac_ilist_fof: float = 32.):
    """Sort, forest, kNN and FoF of a batch on one shared forest (results as knn_sets and fof_sets).

    Returns (rnn, inn, labels), rows in input order, set-local indices."""
    cfg_tree = cfg_tree if cfg_tree is not None else KNNConfig().tree
    try:
        out = _pipeline_core(pos, set_offsets(sizes, pos.device), k, rlink, boxsize, cfg_tree,
                             alloc_fac_ilist_knn, alloc_fac_ilist_fof)
        checks.flush()
    except BaseException:
        checks.clear()
        raise
    return out


def _pipeline_core(pos, off, k, r2set_or_rlink, boxsize, cfg_tree, af_knn, af_fof):
    posz, idz, seg, off, th = tree_sets(pos, off, cfg_tree)
    r2set = (r2set_or_rlink if torch.is_tensor(r2set_or_rlink) and r2set_or_rlink.is_cuda
             else _r2_per_set(r2set_or_rlink, th.nsets, pos.device))
    rnn, inn = _knn_core(posz, idz, seg, off, th, k, boxsize, af_knn)
    lab = _fof_core(posz, idz, seg, off, th, r2set, boxsize, af_fof)
    return rnn, inn, lab


class CapturedSets:
    """pipeline_sets captured once as a HIP graph and replayed (extension).

    The whole sort -> forest -> kNN -> FoF chain is sync-free: sizes are fixed by the batch's
    shape and device-side overflow conditions are only recorded, so one capture replays for any
    positions of the same set sizes.  Every small constant is written by a fill kernel (no host
    copies), and the library records no memset nodes (it clears with fill kernels); memset_nodes()
    reports what the capture holds.

        g = CapturedSets(pos, sizes, k=8, rlink=rl, boxsize=1.0)
        rnn, inn, lab = g.run(new_pos)   # replays; g.check() raises if an allocation overflowed

    The outputs are the graph's own buffers, overwritten by the next run."""

    def __init__(self, pos: torch.Tensor, sizes, k: int, rlink, boxsize: float = 0., cfg_tree=None,
                 alloc_fac_ilist_knn: float = 256., alloc_fac_ilist_fof: float = 32.):
        dev = pos.device
        self.pos = pos.contiguous().clone()
        self.off = set_offsets(sizes, dev).clone()
        nsets = len(self.off) - 1
        self.r2set = _r2_per_set(rlink, nsets, dev)
        cfg_tree = cfg_tree if cfg_tree is not None else KNNConfig().tree
        args = (k, self.r2set, boxsize, cfg_tree, alloc_fac_ilist_knn, alloc_fac_ilist_fof)
        self._step = lambda: _pipeline_core(self.pos, self.off, *args)
        self._g = Captured(self._step)
        self.graph = self._g.graph
        self.out = self._g.out
        self.node_counts = self._g.node_counts

    def memset_nodes(self) -> int:
        return self.node_counts["memset"]

    def run(self, pos: torch.Tensor | None = None):
        """Replay on `pos` (same shape; None = the positions last given).  Returns (rnn, inn, labels)."""
        if pos is not None:
            self.pos.copy_(pos)
        self.graph.replay()
        return self.out

    def check(self):
        """Raise if the last replay overflowed an allocation (synchronises)."""
        self._g.check()
