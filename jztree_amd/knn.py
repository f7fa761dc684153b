"""Exact k-nearest-neighbour search (the API of src/jztree/knn.py, single GPU).

Pipeline (knn.py:198-356): z-sort + tree (tree.py) -> dual tree walk that refines a dense
top-level interaction list plane by plane into leaf-leaf lists (KnnNode2Node) -> leaf-vs-leaf
top-k (KnnLeaf2Leaf).  Neighbours come out ascending by (canonical FP32 r2, reported index).
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass
from typing import Any, Callable, Tuple

import numpy as np
import torch

from . import _lib as L
from . import checks
from .checks import raise_if
from .config import KNNConfig
from .data import InteractionList, Pos, RankIdx, TreeHierarchy, get_num, get_num_total, get_pos, tree_map_by_len
from .tools import inverse_of_splits, masked_inverse
from .tree import (_tmp, f32_view, top_dense_interaction_list, zsort_and_tree, zsort_and_tree_multi_type,
                   zsort_and_tree_origin)


@dataclass
class PosLvlNum:
    """Node data for the kNN walk (data.py:214-239)."""
    pos: torch.Tensor
    lvl: torch.Tensor
    npart: torch.Tensor

    def pos_lvl(self) -> torch.Tensor:
        """(size, dim+1): positions with the level bit-cast into the last column."""
        if self.pos.dtype == torch.float32:
            lvl_t = self.lvl.to(torch.int32).view(torch.float32)
        else:
            lvl_t = self.lvl.to(torch.int64).view(torch.float64)
        return torch.cat([self.pos, lvl_t[:, None]], dim=-1).contiguous()


def _knn_leaf2leaf(ilist: InteractionList, splT, xT, splQ=None, xQ=None, k=32, boxsize=0., src_ids=None,
                   dst_rows=None, out_rows: int | None = None, init_r2=None, out=None, balance: bool = True):
    """knn.py:28-60 -> KnnLeaf2Leaf.  Returns (rnn (rows,k), inn (rows,k) int32), plus the retry
    flags (nq int32) when init_r2 is given.

    Extensions: src_ids = id reported for each source particle (default its z-index); dst_rows =
    output row of each query (default its z-index; rows = out_rows, default nq); init_r2 = starting
    bound per query leaf (k <= 64; the lists must hold every leaf within it), queries left with
    fewer than k neighbours inside it are flagged and their rows must be recomputed; leaves with a
    negative bound are skipped.  out = (rnn, inn[, retry]) writes into existing buffers.
    balance: hand the query leaves to the waves through the kernel's work queue (one atomic per
    leaf); False keeps the fixed order, for passes in which most leaves are skipped (negative
    bounds), where the per-leaf atomics would cost more than the pass itself."""
    if xQ is None:
        xQ = xT
    if splQ is None:
        splQ = splT
    assert ilist.rad2.shape == ilist.isrc.shape
    assert xT.dtype == xQ.dtype and xT.shape[-1] == xQ.shape[-1]
    assert splT.dtype == splQ.dtype == torch.int32
    assert len(ilist.ispl) == len(splQ)
    dev = xQ.device
    nq = xQ.shape[0]
    rows = nq if out_rows is None else int(out_rows)
    dt = L.dtype_code(xT.dtype)
    if out is not None:
        rnn, inn = out[0], out[1]
        assert rnn.shape == (rows, k) and inn.shape == (rows, k) and rnn.is_contiguous() and inn.is_contiguous()
        retry = out[2] if len(out) > 2 else None
    else:
        rnn = torch.empty((rows, k), dtype=xT.dtype, device=dev)
        inn = torch.empty((rows, k), dtype=torch.int32, device=dev)
        retry = None
    nb = L.LIB.jzt_knn_leaf2leaf_tmp_bytes(nq, k, dt)
    if not balance:  # only the multi-pass state: the kernel then runs the leaves in a fixed order
        nb = L.LIB.jzt_knn_leaf2leaf_tmp_bytes(nq, k, dt) - 256
    tmp = _tmp(nb, dev)
    sid = None if src_ids is None else src_ids.to(torch.int32).contiguous()
    drow = None if dst_rows is None else dst_rows.to(torch.int32).contiguous()
    ir2 = None if init_r2 is None else init_r2.to(torch.float32).contiguous()
    if init_r2 is not None and retry is None:
        retry = torch.empty(nq, dtype=torch.int32, device=dev)
    assert retry is None or (retry.shape == (nq,) and retry.dtype == torch.int32)
    L.call("jzt_knn_leaf2leaf_ex", L.stream(), L.ptr(ilist.ispl), L.ptr(ilist.isrc), L.ptr(ilist.rad2),
           ilist.size(), L.ptr(splT), L.ptr(xT.contiguous()), L.ptr(splQ), L.ptr(xQ.contiguous()), L.ptr(sid),
           L.ptr(drow), L.ptr(ir2), L.ptr(retry), L.ptr(rnn), L.ptr(inn), L.ptr(tmp), nb, len(splQ) - 1, nq, rows,
           int(k), float(boxsize or 0.), xT.shape[-1], dt)
    if init_r2 is not None:
        return rnn, inn, retry
    return rnn, inn


_knn_leaf2leaf.jit = _knn_leaf2leaf


def _knn_node2node_ilist(ilist: InteractionList, spl_parent, node_data: PosLvlNum, k: int = 32,
                         boxsize: float = 0., with_radii: bool = False, rmax2_in=None, overflow: list | None = None):
    """knn.py:62-110 -> KnnNode2Node.

    rmax2_in (extension): search radius^2 per node instead of the guaranteed k-neighbour radius;
    each list then holds every node within rmax2_in * (1 + 2^-16).  overflow (extension): instead of
    raising on a too small list, append the device flag "list overflowed" to this list."""
    boxsize = 0. if boxsize is None else boxsize
    assert ilist.ispl.shape[0] == spl_parent.shape[0]
    assert ilist.isrc.shape == ilist.rad2.shape
    dev = spl_parent.device
    size = len(node_data.pos)
    poslvl = node_data.pos_lvl()
    radii = torch.empty(size, dtype=torch.float32, device=dev)
    ispl = torch.empty(size + 1, dtype=torch.int32, device=dev)
    il = torch.empty(ilist.size(), dtype=torch.int32, device=dev)
    ilr2 = torch.empty(ilist.size(), dtype=torch.float32, device=dev)
    npart = node_data.npart.to(torch.int32).contiguous()
    r2in = None if rmax2_in is None else rmax2_in.to(torch.float32).contiguous()
    assert r2in is None or r2in.shape[0] >= size
    L.call("jzt_knn_node2node_ex", L.stream(), L.ptr(ilist.ispl), L.ptr(ilist.isrc), L.ptr(ilist.rad2),
           L.ptr(spl_parent.contiguous()), L.ptr(poslvl), L.ptr(npart), L.ptr(r2in), L.ptr(radii), L.ptr(ispl),
           L.ptr(il), L.ptr(ilr2), spl_parent.shape[0] - 1, size, ilist.size(), ilist.size(), int(k),
           float(boxsize), node_data.pos.shape[-1], L.dtype_code(node_data.pos.dtype))
    if overflow is not None:  # the caller checks (and handles) overflow itself
        overflow.append(ispl[-1] > il.numel())
    else:
        raise_if(ispl[-1] > il.numel(),
                 "The interaction list allocation is too small. (need: {n1}, have: {n2})\n"
                 "Hint: increase alloc_fac_ilist", n1=ispl[-1], n2=il.numel())
    if with_radii:
        return InteractionList(ispl, il, rad2=ilr2), radii
    return InteractionList(ispl, il, rad2=ilr2)


def _knn_dual_walk(th: TreeHierarchy, k: int, boxsize: float | None = None,
                   alloc_fac_ilist: float = 32., stop_level: int = 0, return_parents: bool = False):
    """knn.py:132-193 (single GPU): from the dense top-plane list down to the lists of plane
    `stop_level` (0 = leaf-leaf lists, as in the reference).

    With return_parents, also returns (rmax2 of the plane's nodes or None at the top, the
    node-to-child splits of that plane)."""
    nlevels = th.num_planes()
    size = th.size()
    spl, ilist, nsup = top_dense_interaction_list(th, int(th.size_leaves * alloc_fac_ilist), with_r2=True)
    spl_n2n = th.ispl_n2n.append(spl, nsup + 1, fill_value=spl[-1], resize=True)
    raise_if(spl_n2n.nfilled() > spl_n2n.size(), "No space for top-level splits\nHint: Increase alloc_fac_nodes.")
    rmax2 = None
    for i in range(nlevels - stop_level):
        level = nlevels - i - 1
        parent_spl = spl_n2n.get(level + 1, size + 1)
        node_data = PosLvlNum(pos=th.geom_cent.get(level, size), lvl=th.lvl.get(level, size),
                              npart=th.npart(level, size=size, ptype=0))
        ilist, rmax2 = _knn_node2node_ilist(ilist, parent_spl, node_data, k=k, boxsize=boxsize, with_radii=True)
    if return_parents:
        return ilist, rmax2, spl_n2n.get(stop_level, size + 1)
    return ilist


def _leaf_parent(n2c: torch.Tensor, size: int) -> torch.Tensor:
    """Parent slot of every leaf slot.  The padded tail of the splits piles up at the leaf count,
    so empty leaf slots would map to `size`; they are clamped into range (their kernels exit
    before reading it)."""
    from .tools import inverse_of_splits
    return torch.clamp(inverse_of_splits(n2c, size), 0, size - 1)


def _leaf_plane_args(th: TreeHierarchy, n2c: torch.Tensor):
    """(leaf -> plane-1 parent, leaf centre+level rows, leaf -> particle splits) for the fused
    leaf-plane kernels."""
    size = th.size()
    leaf_parent = _leaf_parent(n2c, size)
    leaf_nodes = PosLvlNum(pos=th.geom_cent.get(0, size), lvl=th.lvl.get(0, size), npart=leaf_parent).pos_lvl()
    return leaf_parent, leaf_nodes, th.splits_leaf_to_part().contiguous()


def _knn_fused(th: TreeHierarchy, posz, k: int, boxsize: float | None = None, alloc_fac_ilist: float = 256.,
               src_ids=None, dst_rows=None, out_rows: int | None = None, init_r2=None):
    """Leaf plane as one fused kernel (extension; see csrc/knn_fused.hip): walks the plane-1 lists
    and evaluates child leaves on the fly instead of materialising leaf-leaf lists.

    init_r2: per-leaf starting bound (default: the parent's guaranteed radius).  Queries that hold
    fewer than k neighbours within it are redone with an unbounded search inside the call.
    Results are identical to _knn_dual_walk + _knn_leaf2leaf with the same src_ids."""
    dev = posz.device
    n = len(posz)
    size = th.size()
    ilist_p, rmax_p, n2c = _knn_dual_walk(th, k, boxsize=boxsize, alloc_fac_ilist=alloc_fac_ilist, stop_level=1,
                                          return_parents=True)
    leaf_parent, leaf_nodes, l2p = _leaf_plane_args(th, n2c)
    if init_r2 is None and rmax_p is not None:
        init_r2 = rmax_p[leaf_parent.long()]
    rows = n if out_rows is None else out_rows
    rnn = torch.empty((rows, k), dtype=posz.dtype, device=dev)
    inn = torch.empty((rows, k), dtype=torch.int32, device=dev)
    retry = torch.empty(n, dtype=torch.int32, device=dev) if init_r2 is not None else None
    sid = None if src_ids is None else src_ids.to(torch.int32).contiguous()
    drow = None if dst_rows is None else dst_rows.to(torch.int32).contiguous()
    ir2 = None if init_r2 is None else init_r2.to(torch.float32).contiguous()
    L.call("jzt_knn_fused", L.stream(), L.ptr(ilist_p.ispl), L.ptr(ilist_p.isrc), L.ptr(ilist_p.rad2),
           ilist_p.size(), L.ptr(n2c.contiguous()), L.ptr(leaf_parent), L.ptr(leaf_nodes), L.ptr(l2p),
           L.ptr(posz.contiguous()), L.ptr(sid), L.ptr(drow), L.ptr(ir2), L.ptr(retry), L.ptr(rnn), L.ptr(inn),
           size, n, int(k), float(boxsize or 0.), posz.shape[-1], L.dtype_code(posz.dtype))
    return rnn, inn, retry


def _guaranteed_leaf_r2(th: TreeHierarchy, k: int) -> torch.Tensor:
    """Per leaf slot, a radius^2 within which every query of the leaf has its k nearest (itself
    included): the squared cell diagonal of the smallest enclosing node (the leaf or an ancestor)
    that holds >= k particles (all of them lie in that cell, and the periodic wrap only shortens
    differences); +inf when no node holds k particles.  One jzt_plane_cap launch per plane, from
    the top plane down (a node takes its parent's value unless it holds k particles itself)."""
    size = th.size()
    dim = th.geom_cent.data.shape[-1]
    dev = th.lvl.data.device
    spl = th.splits_leaf_to_part().to(torch.int32).contiguous()
    cap_up = n2c_up = None
    for p in reversed(range(th.num_planes())):
        n2l = th.ispl_n2l.get(p, size + 1).to(torch.int32).contiguous()
        lvl = th.lvl.get(p, size).to(torch.int32).contiguous()
        cap = torch.empty(size, dtype=torch.float64, device=dev)
        L.call("jzt_plane_cap", L.stream(), L.ptr(n2l), L.ptr(spl), spl.numel(), L.ptr(lvl), L.ptr(n2c_up),
               L.ptr(cap_up), size, int(k), int(dim), L.ptr(cap))
        cap_up = cap
        n2c_up = th.ispl_n2n.get(p, size + 1).to(torch.int32).contiguous() if p > 0 else None
    return cap_up


def _plane_radii(th: TreeHierarchy, leaf_r2: torch.Tensor, upper_r2: torch.Tensor | None = None) -> list:
    """Search radius^2 for every plane: leaf_r2 for the leaves (negative = no search), and for a
    node the largest radius among its children.  A node's list then holds every node that a leaf
    inside it may reach: mindist(child, T') >= mindist(node, T) for children of T.

    upper_r2 (per leaf, >= leaf_r2): the planes above the leaves take their radii from these
    instead (e.g. guaranteed radii, so that the plane-1 lists also serve an unbounded search)."""
    size = th.size()
    dev = leaf_r2.device
    ar = torch.arange(size, device=dev)
    neg = torch.full((size,), -1.0, dtype=torch.float32, device=dev)
    radii = [torch.where(ar < th.num(0), leaf_r2.float(), neg).contiguous()]
    chain = radii[0] if upper_r2 is None else torch.where(ar < th.num(0), upper_r2.float(), neg).contiguous()
    for p in range(1, th.num_planes()):
        n2c = th.ispl_n2n.get(p, size + 1).to(torch.int32).contiguous()
        r = torch.empty(size, dtype=torch.float32, device=dev)
        L.call("jzt_segment_max_f32", L.stream(), L.ptr(n2c), L.ptr(chain), size, L.ptr(r), size, -1.0)
        chain = torch.where(ar < th.num(p), r, neg).contiguous()
        radii.append(chain)
    return radii


def _knn_radius_walk(th: TreeHierarchy, k: int, radii: list, boxsize: float = 0., alloc_fac_ilist: float = 256.,
                     overflow: list | None = None, keep_plane1: bool = False):
    """Dual walk with given search radii on every plane (no guaranteed-radius search): from the dense
    top-plane list down to leaf lists holding every leaf within each leaf's radius.

    keep_plane1: also return (the plane-1 lists, plane-1 -> leaf splits) -- with a one-plane tree
    the dense top-level list and its super-node splits."""
    nlevels = th.num_planes()
    size = th.size()
    spl, ilist, nsup = top_dense_interaction_list(th, int(th.size_leaves * alloc_fac_ilist), with_r2=True)
    spl_n2n = th.ispl_n2n.append(spl, nsup + 1, fill_value=spl[-1], resize=True)
    raise_if(spl_n2n.nfilled() > spl_n2n.size(), "No space for top-level splits\nHint: Increase alloc_fac_nodes.")
    plane1 = None
    for i in range(nlevels):
        level = nlevels - i - 1
        parent_spl = spl_n2n.get(level + 1, size + 1)
        if level == 0:
            plane1 = (ilist, parent_spl)
        node_data = PosLvlNum(pos=th.geom_cent.get(level, size), lvl=th.lvl.get(level, size),
                              npart=th.npart(level, size=size, ptype=0))
        ilist = _knn_node2node_ilist(ilist, parent_spl, node_data, k=k, boxsize=boxsize, rmax2_in=radii[level],
                                     overflow=overflow)
    if keep_plane1:
        return ilist, plane1
    return ilist


def _knn_bounded(th: TreeHierarchy, posz, k: int, leaf_r2: torch.Tensor, boxsize: float | None = None,
                 alloc_fac_ilist: float = 256., src_ids=None, dst_rows=None, out_rows: int | None = None,
                 grow: float | None = None, stats: dict | None = None, direct_retry: bool = True):
    """Exact kNN from a proposed search radius^2 per leaf (extension; k <= 64).

    Round 1 walks the tree with the proposed radii at the leaf plane (each clipped to the guaranteed
    one, _guaranteed_leaf_r2) and runs the leaf kernel with each leaf's radius as starting bound.
    Queries that find fewer than k neighbours inside it are flagged.  With direct_retry (default)
    the planes above the leaves search the guaranteed radii (a node as far as its farthest-reaching
    child), so round 1's plane-1 lists already hold every node a flagged query can need, and the
    flagged queries go straight to the fused retry pass over those lists.  Without it every plane
    searches the proposed radii, and the next round redoes the flagged queries' leaves with the
    guaranteed radius while every other leaf sits out (with `grow`, a round with the radius^2 grown
    by that factor, still clipped, comes first); each such round walks the whole tree again.
    Whatever is left -- or everything, should a round's lists overflow their allocation -- is
    recomputed without any proposal: by the fused retry pass (after its own walk to plane 1 when
    round 1 did not keep its plane-1 lists), or by the plain walk when round 1 itself overflowed.
    The result is therefore independent of the proposal: identical to knn() for any leaf_r2.
    Returns (rnn, inn, retry flags of round 1)."""
    assert k <= 64, "a start bound needs k <= 64"
    dev = posz.device
    n = len(posz)
    size = th.size()
    boxsize = 0. if boxsize is None else boxsize
    rows = n if out_rows is None else out_rows
    spl = th.splits_leaf_to_part()
    rnn = torch.empty((rows, k), dtype=posz.dtype, device=dev)
    inn = torch.empty((rows, k), dtype=torch.int32, device=dev)
    retry = torch.zeros(n, dtype=torch.int32, device=dev)
    st = stats if stats is not None else {}
    # never search beyond the guaranteed radius (a wild proposal would only inflate the lists); the
    # 2^-10 widening keeps rounding of the canonical r2 from causing avoidable retries
    cap = (_guaranteed_leaf_r2(th, k) * (1.0 + 2.0 ** -10)).to(torch.float32)
    # a NaN or negative proposal (a diverged network) would make the leaf kernel skip the leaf
    # (that is how later rounds leave leaves untouched): such leaves search the guaranteed radius
    lr2 = leaf_r2.to(torch.float32)
    r2 = torch.where(torch.isnan(lr2) | (lr2 < 0), cap, torch.minimum(lr2, cap)).contiguous()

    plane1 = []

    def run(r2, upper=None, balance=True):
        ovf = []
        il, p1 = _knn_radius_walk(th, k, _plane_radii(th, r2, upper), boxsize=boxsize,
                                  alloc_fac_ilist=alloc_fac_ilist, overflow=ovf, keep_plane1=True)
        if upper is not None:
            plane1.append(p1)
        _knn_leaf2leaf(il, spl, posz, k=k, boxsize=boxsize, src_ids=src_ids, dst_rows=dst_rows, out_rows=rows,
                       init_r2=r2, out=(rnn, inn, retry), balance=balance)
        return torch.stack(ovf).any()

    # with direct_retry the planes above the leaves search the guaranteed radii: their plane-1 lists
    # then serve the fused retry pass directly (no second walk)
    ovf1 = run(r2, cap if direct_retry else None)
    # the fused retry pass only reads the flags; later rounds (direct_retry off) rewrite them
    first = retry if direct_retry else retry.clone()
    if checks.is_deferred():
        # under graph capture nothing is read back: an overflow of round 1's lists is reported after
        # the replay (as the reference reports list overflows), and the fused retry pass always runs
        # (it returns at once when no query is flagged)
        assert direct_retry, "graph capture of the bounded search needs direct_retry"
        raise_if(ovf1, "Interaction list allocation too small for the proposed radii\n"
                       "Hint: increase alloc_fac_ilist")
        h = None
    else:
        # one read-back per round: (lists overflowed, queries left)
        h = torch.stack([ovf1.to(torch.int64), first.sum().to(torch.int64)]).cpu()
    if h is None:
        st["fallback"] = "fused retry (unconditional)"
        plane1_lists, n2c = plane1[0]
        _fused_retry(th, plane1_lists, n2c, posz, k, boxsize, src_ids, dst_rows, cap, retry, rnn, inn)
        return rnn, inn, first
    st["short_round1"] = int(h[1])
    if bool(h[0]):
        st["fallback"] = "plain (round 1 lists overflowed)"
        il = _knn_dual_walk(th, k, boxsize=boxsize, alloc_fac_ilist=alloc_fac_ilist)
        _knn_leaf2leaf(il, spl, posz, k=k, boxsize=boxsize, src_ids=src_ids, dst_rows=dst_rows, out_rows=rows,
                       out=(rnn, inn))
        return rnn, inn, first
    if int(h[1]) == 0:
        return rnn, inn, first
    # Later rounds: the leaves holding a flagged query (segment sums of the flags over the leaf
    # splits), optionally with the radius grown by `grow`, then with the guaranteed radius; each
    # round's lists and flags are read back once.
    sl = spl.long().clamp(0, n)
    later = ((lambda: torch.minimum(r2 * grow, cap)),) if grow else ()
    later = () if direct_retry else later + ((lambda: cap),)
    for rnd, r2_next in enumerate(later, start=2):
        cs = torch.cat([torch.zeros(1, dtype=torch.int64, device=dev), torch.cumsum(retry, 0)])
        short = (cs[sl[1:]] - cs[sl[:-1]]) > 0
        ovf = run(torch.where(short, r2_next(), torch.full_like(r2, -1.0)).contiguous(), balance=False)
        h = torch.stack([ovf.to(torch.int64), retry.sum().to(torch.int64), short.sum().to(torch.int64)]).cpu()
        st[f"short_leaves_round{rnd}"] = int(h[2])
        st[f"short_round{rnd}"] = int(h[1])
        if bool(h[0]):  # this round's lists overflowed: every row it rewrote (all short leaves) is suspect
            leafof = inverse_of_splits(spl, n).long().clamp_(0, size - 1)
            retry.copy_(short[leafof].to(torch.int32))
            break
        if int(h[1]) == 0:
            return rnn, inn, first
    st["fallback"] = "fused retry"
    if plane1:
        ilist_p, n2c = plane1[0]
    else:
        ilist_p, _, n2c = _knn_dual_walk(th, k, boxsize=boxsize, alloc_fac_ilist=alloc_fac_ilist, stop_level=1,
                                         return_parents=True)
    _fused_retry(th, ilist_p, n2c, posz, k, boxsize, src_ids, dst_rows, cap, retry, rnn, inn)
    return rnn, inn, first


def _fused_retry(th: TreeHierarchy, ilist_p: InteractionList, n2c, posz, k: int, boxsize: float, src_ids, dst_rows,
                 cap, retry, rnn, inn):
    """The flagged queries (retry != 0) redone over the plane-1 lists (jzt_knn_fused_retry), each
    starting from its leaf's guaranteed radius (every query has its k nearest within it)."""
    leaf_parent, leaf_nodes, l2p = _leaf_plane_args(th, n2c)
    sid = None if src_ids is None else src_ids.to(torch.int32).contiguous()
    drow = None if dst_rows is None else dst_rows.to(torch.int32).contiguous()
    L.call("jzt_knn_fused_retry", L.stream(), L.ptr(ilist_p.ispl), L.ptr(ilist_p.isrc), L.ptr(ilist_p.rad2),
           ilist_p.size(), L.ptr(n2c.contiguous()), L.ptr(leaf_parent), L.ptr(leaf_nodes), L.ptr(l2p),
           L.ptr(posz.contiguous()), L.ptr(sid), L.ptr(drow), L.ptr(cap.contiguous()), L.ptr(retry), L.ptr(rnn),
           L.ptr(inn), th.size(), len(posz), int(k), float(boxsize), posz.shape[-1], L.dtype_code(posz.dtype))


def _knn_leaf_start_bound(ilist: InteractionList, spl, posz, k: int, boxsize, leaf_r2, src_ids=None,
                          dst_rows=None, splQ=None, xQ=None):
    """Leaf kernel over the guaranteed-radius lists with each leaf's guaranteed k-th radius^2
    (from KnnNode2Node, widened by 2^-10) as the start bound of its queries (extension, k <= 64):
    candidates beyond it are never buffered, so the top-k fills with fewer flushes.  The lists hold
    every leaf within that radius, so a query can only come up short if the float bound rounds
    below its k-th neighbour; such leaves are redone without a bound.  Results equal the
    unbounded kernel's.  splQ/xQ: separate queries (default: the sources); the guaranteed radius of
    a leaf holds for any point of its cell."""
    splQ = spl if splQ is None else splQ
    xQ = posz if xQ is None else xQ
    nq = len(xQ)
    size = len(splQ) - 1
    init = (leaf_r2.to(torch.float64) * (1.0 + 2.0 ** -10)).to(torch.float32).contiguous()
    rnn, inn, retry = _knn_leaf2leaf(ilist, spl, posz, k=k, boxsize=boxsize, splQ=splQ, xQ=xQ, src_ids=src_ids,
                                     dst_rows=dst_rows, init_r2=init)
    # eagerly the redo runs only when a query came up short (one read-back); under graph capture
    # it always runs, and every leaf without a short query skips it (negative radius)
    if checks.is_deferred() or bool((retry != 0).any()):
        # per leaf: does it hold a flagged query (a segment maximum over the query splits; a
        # scatter of every flag onto its leaf costs an atomic per query)
        short = torch.empty(size, dtype=torch.float32, device=xQ.device)
        L.call("jzt_segment_max_f32", L.stream(), L.ptr(splQ.contiguous()), L.ptr(retry.to(torch.float32)), nq,
               L.ptr(short), size, 0.0)
        redo = torch.where(short > 0, torch.full_like(init, float("inf")), torch.full_like(init, -1.0))
        _knn_leaf2leaf(ilist, spl, posz, k=k, boxsize=boxsize, splQ=splQ, xQ=xQ, src_ids=src_ids,
                       dst_rows=dst_rows, init_r2=redo.contiguous(), out=(rnn, inn, retry), balance=False)
    return rnn, inn


def _distance_view(th: TreeHierarchy, posz: torch.Tensor, dist_dtype):
    """(tree, z-ordered positions) for the distance kernels: unchanged, or for float32 distances on
    float64 positions the float32 node geometry (tree.f32_view) and the rounded positions."""
    if dist_dtype is None or dist_dtype == posz.dtype:
        return th, posz
    if not (dist_dtype == torch.float32 and posz.dtype == torch.float64):
        raise ValueError("dist_dtype: only float32 distances on float64 positions are supported")
    return f32_view(th), posz.to(torch.float32).contiguous()


def knn(part, k: int, boxsize: float | None = None, th: TreeHierarchy | None = None, part_query=None,
        result: str = "rad_globalidx", reduce_func: Callable | None = None, output_order: str = "input",
        cfg: KNNConfig = KNNConfig(), fused: bool = False, dist_dtype: torch.dtype | None = None):
    """Exact k nearest neighbours (knn.py:198-347).

    Returns (rnn, inn) by default: radii and indices of the k nearest neighbours of every particle,
    ascending by (distance, index).  result: "_"-separated keys among rad, drad, rankidx,
    globalidx, part, reduce, or a particle attribute name.  output_order: "input" or "z".
    fused (extension, k <= 64, same results): True = the fused leaf-plane kernel (one wave per query
    leaf over its parent's list; float32 positions in 2-3 dims, ignored otherwise); False runs the
    reference's KnnNode2Node + KnnLeaf2Leaf decomposition (the default: fastest with the leaf sizes
    the autotuner picks).  Positions: float32 or float64, 2 to 10 dims (the reference's
    instances); radii come in the positions' type.
    dist_dtype (extension): torch.float32 with float64 positions sorts and builds the tree in
    float64 (63-bit keys) and searches the float32-rounded positions with the float32 kernels; the
    result is the exact kNN of the rounded points (radii float32).
    """
    assert output_order in ("z", "input")
    pos = get_pos(part)
    dev = pos.device
    size = len(pos)
    try:
        num_origin = get_num(part, default_to_length=True)

        if part_query is not None:
            # one tree over sources (type 0) and queries (type 1), knn.py:259-264
            assert th is None, "Querying with pre-sorted particles is not supported yet"
            size_q = len(get_pos(part_query))
            num_origin_q = get_num(part_query, default_to_length=True)
            ar = lambda m: RankIdx(rank=None, idx=torch.arange(m, dtype=torch.int32, device=dev))
            (partz, partz_q), (originz, originz_q), th = zsort_and_tree_multi_type(
                (part, part_query), cfg.tree, data=(ar(size), ar(size_q)))
        else:
            if th is not None:
                partz = part
                originz = RankIdx(rank=None, idx=torch.arange(size, dtype=torch.int32, device=dev))
            else:
                partz, originz, th = zsort_and_tree_origin(part, cfg.tree)
            partz_q, originz_q, size_q, num_origin_q = partz, originz, size, num_origin
        if output_order == "z":
            originz = RankIdx(rank=None, idx=torch.arange(size, dtype=torch.int32, device=dev))
            num_origin = len(get_pos(partz))
            originz_q = RankIdx(rank=None, idx=torch.arange(size_q, dtype=torch.int32, device=dev))
            num_origin_q = len(get_pos(partz_q))

        spl = th.splits_leaf_to_part(ptype=0)
        spl_q = th.splits_leaf_to_part(ptype=0 if part_query is None else 1)
        posz = get_pos(partz)
        th, posz = _distance_view(th, posz, dist_dtype)
        posz_q = posz if part_query is None else get_pos(partz_q).to(posz.dtype).contiguous()
        keys = result.split("_")
        # rows can be written straight to input order when only radii/indices are requested
        direct = (output_order == "input" and set(keys) <= {"rad", "globalidx"}
                  and isinstance(num_origin_q, int) and num_origin_q == size_q)
        drows = originz_q.idx if direct else None
        # the start-bound leaf kernels take up to 64 slots (32 for float64 positions) in one pass;
        # the fused leaf-plane kernel is float32 in 2-3 dims
        kb = 64 if posz.dtype == torch.float32 else 32
        fused = fused and posz.dtype == torch.float32 and posz.shape[-1] <= 3
        if part_query is not None and k <= kb:
            ilist, leaf_r2, _ = _knn_dual_walk(th, k, boxsize=boxsize, alloc_fac_ilist=cfg.alloc_fac_ilist,
                                               return_parents=True)
            rnnz, inn_g = _knn_leaf_start_bound(ilist, spl, posz, k, boxsize, leaf_r2, originz.idx, drows,
                                                splQ=spl_q, xQ=posz_q)
        elif part_query is not None:
            ilist = _knn_dual_walk(th, k, boxsize=boxsize, alloc_fac_ilist=cfg.alloc_fac_ilist)
            rnnz, inn_g = _knn_leaf2leaf(ilist, spl, posz, k=k, boxsize=boxsize, splQ=spl_q, xQ=posz_q,
                                         src_ids=originz.idx, dst_rows=drows)
        elif fused and k <= 64:
            # neighbours are reported (and tie-broken) by their origin index directly
            rnnz, inn_g, _ = _knn_fused(th, posz, k, boxsize=boxsize, alloc_fac_ilist=cfg.alloc_fac_ilist,
                                        src_ids=originz.idx, dst_rows=drows)
        elif k <= kb:
            ilist, leaf_r2, _ = _knn_dual_walk(th, k, boxsize=boxsize, alloc_fac_ilist=cfg.alloc_fac_ilist,
                                               return_parents=True)
            rnnz, inn_g = _knn_leaf_start_bound(ilist, spl, posz, k, boxsize, leaf_r2, originz.idx, drows)
        else:
            ilist = _knn_dual_walk(th, k, boxsize=boxsize, alloc_fac_ilist=cfg.alloc_fac_ilist)
            rnnz, inn_g = _knn_leaf2leaf(ilist, spl, posz, k=k, boxsize=boxsize, splQ=spl, xQ=posz,
                                         src_ids=originz.idx, dst_rows=drows)

        res = []
        zidx_cache = {}

        def innz():
            if "z" not in zidx_cache:
                inv = torch.zeros(size + 1, dtype=torch.int64, device=dev)
                inv[originz.idx.long()] = torch.arange(size, device=dev)
                zidx_cache["z"] = inv[inn_g.long().clamp(min=0, max=size)]
            return zidx_cache["z"]

        for key in keys:
            if key == "rad":
                res.append(rnnz)
            elif key == "drad":
                res.append(torch.linalg.norm(posz_q[:, None] - posz[innz()], dim=-1))
            elif key == "rankidx":
                res.append(RankIdx(rank=None, idx=inn_g))
            elif key == "globalidx":
                if get_num_total(part, default_to_length=True) >= 2 ** 31:
                    raise ValueError("I have > 2**31 particles, globalidx will overflow. Use rankidx instead!")
                res.append(inn_g)
            elif key == "part":
                res.append(tree_map_by_len(lambda v: v[innz()], partz, size))
            elif key == "reduce":
                assert reduce_func is not None, "Please provide reduce_func if using 'reduce'"
                x = reduce_func(part=partz, rnn=rnnz, inn=innz(), origin=originz)
                res.append(x)
            elif hasattr(partz, key):
                arr = getattr(partz, key)
                res.append(tree_map_by_len(lambda v: v[innz()], arr, size))
            else:
                raise ValueError(f"Invalid result key {key}")
        res = tuple(res) if len(res) > 1 else res[0]

        if output_order == "input" and not direct:
            idx = originz_q.idx
            inverse = masked_inverse(idx, mask=torch.arange(len(idx), device=dev) < num_origin_q)
            res = tree_map_by_len(lambda v: v[inverse.long().clamp(max=size_q - 1)], res, size_q)
        checks.flush()
    except BaseException:
        checks.clear()
        raise
    return res


knn.jit = knn
