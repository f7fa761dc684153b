"""Friends-of-friends labels and catalogues (the API of src/jztree/fof.py, single GPU).

Pipeline (fof.py:772-816): z-sort + tree -> dual walk over the planes (NodeToChildLabel,
FofNode2Node: guaranteed node links + lists of uncertain node pairs) -> particle level
(NodeToChildLabel, FofLeaf2Leaf).  Labels in z-order are the minimum z-index of each group.
Catalogue (fof.py:510-519, :650-766, :872-907): particles brought into group order (jzt_fof_order),
then per-group sums (jzt_fof_catalogue).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Tuple

import numpy as np
import torch

from . import _lib as L
from . import checks
from .checks import raise_if
from .config import FofCatalogueConfig, FofConfig
from .data import (FofCatalogue, InteractionList, PackedArray, TreeHierarchy, count_not_nan, get_pos,
                   tree_map_by_len, verify_ilist)
from .tree import _tmp, build_tree_hierarchy, top_dense_interaction_list, zsort


@dataclass
class PosLvl:
    """Node positions and levels (data.py:199-212)."""
    pos: torch.Tensor
    lvl: torch.Tensor

    def pos_lvl(self) -> torch.Tensor:
        if self.pos.dtype == torch.float32:
            lvl_t = self.lvl.to(torch.int32).view(torch.float32)
        else:
            lvl_t = self.lvl.to(torch.int64).view(torch.float64)
        return torch.cat([self.pos, lvl_t[:, None]], dim=-1).contiguous()


@dataclass
class FofNodeData:
    """data.py:566-570.  cent (extension): the nodes' cell centres, when known (the leaf kernel
    then drops list entries whose cell is out of reach of a query chunk)."""
    lvl: torch.Tensor
    igroup: torch.Tensor
    spl: torch.Tensor
    cent: torch.Tensor | None = None


def _r2link(rlink: float) -> float:
    """fof.py:50, :101, :165: np.float32(rlink*rlink), the product formed in float64."""
    return float(np.float32(rlink * rlink))


def _r2_items(r2: torch.Tensor, n: int) -> torch.Tensor:
    """A float32 r2link per item, at least n entries (batches of sets pass these instead of rlink)."""
    assert r2.dtype == torch.float32 and r2.numel() >= n
    return r2.contiguous()


def _fof_node2node(ilist: InteractionList, spl_parent, node_data: PosLvl, node_igroup, rlink: float,
                   boxsize: float = 0., block_size: int = 32):
    """fof.py:31-70 -> FofNode2Node.  Returns (node_igroup, node_ilist)."""
    assert ilist.ispl.shape[0] == spl_parent.shape[0]
    assert len(node_data.lvl) == len(node_igroup)
    dev = spl_parent.device
    size = len(node_data.pos)
    ig = node_igroup.to(torch.int32).contiguous()
    il_spl = torch.empty(size + 1, dtype=torch.int32, device=dev)
    il = torch.empty(ilist.size(), dtype=torch.int32, device=dev)
    if torch.is_tensor(rlink):  # r2link per parent node (batches of sets)
        L.call("jzt_fof_node2node_sets", L.stream(), L.ptr(ilist.ispl), L.ptr(ilist.isrc),
               L.ptr(spl_parent.contiguous()), L.ptr(_r2_items(rlink, spl_parent.shape[0] - 1)),
               L.ptr(node_data.pos_lvl()), L.ptr(ig), L.ptr(ig), L.ptr(il_spl), L.ptr(il), spl_parent.shape[0] - 1,
               size, ilist.size(), ilist.size(), float(boxsize or 0.), node_data.pos.shape[-1],
               L.dtype_code(node_data.pos.dtype))
    else:
        L.call("jzt_fof_node2node", L.stream(), L.ptr(ilist.ispl), L.ptr(ilist.isrc), L.ptr(spl_parent.contiguous()),
               L.ptr(node_data.pos_lvl()), L.ptr(ig), L.ptr(ig), L.ptr(il_spl), L.ptr(il), spl_parent.shape[0] - 1,
               size, ilist.size(), ilist.size(), _r2link(rlink), float(boxsize or 0.), node_data.pos.shape[-1],
               L.dtype_code(node_data.pos.dtype))
    node_ilist = verify_ilist(InteractionList(il_spl, il))
    raise_if(node_ilist.nfilled() > il.numel(),
             "The interaction list allocation is too small. (need: {n1} have: {n2})\n"
             "Hint: Increase alloc_fac_ilist", n1=node_ilist.nfilled(), n2=il.numel())
    return ig, node_ilist


def _node_to_child_label(igroup, lvl, spl, rlink: float, dim: int, num=None, size_child: int | None = None,
                         flag_local=None, block_size: int = 64, dtype=torch.float32):
    """fof.py:85-109 -> NodeToChildLabel."""
    assert len(spl) == len(igroup) + 1 == len(lvl) + 1
    dev = igroup.device
    if size_child is None:
        size_child = igroup.numel()
    out = torch.empty(size_child, dtype=torch.int32, device=dev)
    loc = None if flag_local is None else flag_local.to(torch.uint8).contiguous()
    if torch.is_tensor(rlink):  # r2link per parent node (batches of sets)
        assert flag_local is None
        L.call("jzt_node_to_child_label_sets", L.stream(), L.ptr(igroup.to(torch.int32).contiguous()),
               L.ptr(lvl.to(torch.int32).contiguous()), L.ptr(spl.to(torch.int32).contiguous()),
               L.ptr(_r2_items(rlink, igroup.numel())), L.ptr(out), igroup.numel(), size_child, int(dim),
               L.dtype_code(dtype))
    else:
        L.call("jzt_node_to_child_label", L.stream(), L.ptr(igroup.to(torch.int32).contiguous()), L.ptr(loc),
               L.ptr(lvl.to(torch.int32).contiguous()), L.ptr(spl.to(torch.int32).contiguous()), L.ptr(out),
               igroup.numel(), size_child, _r2link(rlink), int(dim), L.dtype_code(dtype))
    if num is not None:
        out = torch.where(torch.arange(size_child, device=dev) < num, out, torch.zeros_like(out))
    return out


def _plane_node_sets(th: TreeHierarchy, spl_super: torch.Tensor) -> list:
    """Set of every node slot on each plane of a batch forest, plus the top super-nodes last."""
    size = th.size()
    nl = len(th.leaf_set)
    sets = []
    for p in range(th.num_planes()):
        first_leaf = torch.clamp(th.ispl_n2l.get(p, size + 1)[:-1].long(), max=nl - 1)
        sets.append(th.leaf_set[first_leaf])
    top = sets[-1]
    sets.append(top[torch.clamp(spl_super[:-1].long(), max=len(top) - 1)])
    return sets


def _fof_dual_walk(th: TreeHierarchy, rlink: float, boxsize: float = 0., alloc_fac_ilist: float = 128,
                   r2link_set: torch.Tensor | None = None):
    """fof.py:111-151: plane-by-plane node linking and interaction lists.

    r2link_set (extension, batch forests): float32 r2link of every set, used in place of rlink."""
    nplanes = th.num_planes()
    size = th.size()
    spl, ilist, nsup = top_dense_interaction_list(th, int(th.size_leaves * alloc_fac_ilist))
    if r2link_set is not None:
        assert th.leaf_set is not None
        r2_plane = [r2link_set[s.long()].contiguous() for s in _plane_node_sets(th, spl)]
    dev = spl.device
    igroup = torch.arange(size, dtype=torch.int32, device=dev)
    spl_n2n = th.ispl_n2n.append(spl, nsup + 1, fill_value=spl[-1], resize=True)
    raise_if(spl_n2n.nfilled() > spl_n2n.size(), "No space for top-level splits\nHint: Increase alloc_fac_nodes.")
    max_lvl = th.info().max_lvl()
    nlvl = th.lvl.append(torch.full_like(igroup, max_lvl), nsup, max_lvl, resize=True)
    dim = th.geom_cent.data.shape[-1]
    for i in range(nplanes):
        level = nplanes - 1 - i
        parent_spl = spl_n2n.get(level + 1, size + 1)
        rl = rlink if r2link_set is None else r2_plane[level + 1]
        igroup = _node_to_child_label(igroup, nlvl.get(level + 1, size), parent_spl, num=th.num(level), rlink=rl,
                                      dim=dim, dtype=th.geom_cent.data.dtype)
        node_data = PosLvl(pos=th.geom_cent.get(level, size), lvl=nlvl.get(level, size))
        igroup, ilist = _fof_node2node(ilist, parent_spl, node_data, igroup, rlink=rl, boxsize=boxsize)
    node_data = FofNodeData(nlvl.get(0, size), igroup, spl_n2n.get(0, size + 1), cent=th.geom_cent.get(0, size))
    return node_data, ilist


def _fof_leaf2leaf(leaf_data: FofNodeData, ilist: InteractionList, posz, rlink: float, boxsize: float = 0.,
                   block_size=32, leaf_r2link: torch.Tensor | None = None):
    """fof.py:153-169 -> NodeToChildLabel (leaf -> particle) + FofLeaf2Leaf.

    leaf_r2link (extension, batch forests): float32 r2link of every leaf slot, used in place of rlink."""
    assert len(leaf_data.spl) == len(ilist.ispl) == len(leaf_data.igroup) + 1 == len(leaf_data.lvl) + 1
    dev = posz.device
    rl = rlink if leaf_r2link is None else leaf_r2link
    # (num=spl[-1] of fof.py:160 masks nothing here: NodeToChildLabel writes only children below
    # the parents' last split and zero-fills the rest)
    part_igroup = _node_to_child_label(leaf_data.igroup, leaf_data.lvl, leaf_data.spl, rlink=rl,
                                       size_child=len(posz), dim=posz.shape[-1], dtype=posz.dtype)
    out = part_igroup  # linked in place
    if leaf_r2link is not None:
        L.call("jzt_fof_leaf2leaf_sets", L.stream(), L.ptr(ilist.ispl), L.ptr(ilist.isrc), ilist.size(),
               L.ptr(leaf_data.spl.contiguous()), L.ptr(_r2_items(leaf_r2link, len(leaf_data.spl) - 1)),
               L.ptr(posz.contiguous()), L.ptr(part_igroup), L.ptr(out), len(leaf_data.spl) - 1, len(posz),
               float(boxsize or 0.), posz.shape[-1], L.dtype_code(posz.dtype))
    elif leaf_data.cent is not None and leaf_data.cent.dtype == posz.dtype:
        cells = PosLvl(pos=leaf_data.cent, lvl=leaf_data.lvl).pos_lvl()
        L.call("jzt_fof_leaf2leaf_cells", L.stream(), L.ptr(ilist.ispl), L.ptr(ilist.isrc), ilist.size(),
               L.ptr(leaf_data.spl.contiguous()), L.ptr(posz.contiguous()), L.ptr(cells), L.ptr(part_igroup),
               L.ptr(out), len(leaf_data.spl) - 1, len(posz), _r2link(rlink), float(boxsize or 0.), posz.shape[-1],
               L.dtype_code(posz.dtype))
    else:
        L.call("jzt_fof_leaf2leaf", L.stream(), L.ptr(ilist.ispl), L.ptr(ilist.isrc), ilist.size(),
               L.ptr(leaf_data.spl.contiguous()),
               L.ptr(posz.contiguous()), L.ptr(part_igroup), L.ptr(out), len(leaf_data.spl) - 1, len(posz),
               _r2link(rlink), float(boxsize or 0.), posz.shape[-1], L.dtype_code(posz.dtype))
    return out


def fof_is_superset(igroup_sup, igroup, mask=None):
    """fof.py:203-218: every group of igroup_sup is a union of groups of igroup."""
    label_map = torch.zeros(len(igroup), dtype=torch.int64, device=igroup.device)
    label_map[igroup.long()] = igroup_sup.long()
    ok = igroup_sup.long() == label_map[igroup.long()]
    if mask is not None:
        ok = ok | ~mask
    return bool(torch.all(ok))


def fof_labels(part, rlink: float, boxsize: float = 0., cfg: FofConfig = FofConfig(), th: TreeHierarchy | None = None,
               output_order: str = "z", dist_dtype: torch.dtype | None = None):
    """Friends-of-friends groups on one GPU (fof.py:772-816).

    Returns (part, igroup): with output_order="z" the z-ordered particles and labels pointing to
    the first z-order particle of each group; with "input" the input particles and labels in
    input order (the input index of the group's first z-order member).  dist_dtype (extension):
    torch.float32 with float64 positions links the float32-rounded positions with the float32
    kernels over the float64 tree (z-order from the float64 keys)."""
    try:
        if th is None:
            partz, idz = zsort(part)
            th = build_tree_hierarchy(partz, cfg_tree=cfg.tree)
        else:
            partz, idz = part, None
        from .knn import _distance_view
        th, posz = _distance_view(th, get_pos(partz), dist_dtype)
        node_data, ilist = _fof_dual_walk(th, rlink=rlink, boxsize=boxsize, alloc_fac_ilist=cfg.alloc_fac_ilist)
        igroupz = _fof_leaf2leaf(node_data, ilist, posz, rlink=rlink, boxsize=boxsize)
        checks.flush()
    except BaseException:
        checks.clear()
        raise
    if output_order == "input" and idz is not None:
        idl = idz.long()
        igroup = torch.empty_like(igroupz)
        igroup[idl] = idz[igroupz.long()].to(igroupz.dtype)
        return part, igroup
    return partz, igroupz


fof_labels.jit = fof_labels


# ------------------------------------------------------------------------------------------------
# group order and catalogue (fof.py:510-519, :650-766, :872-907)
# ------------------------------------------------------------------------------------------------

def _fof_order(igroup: torch.Tensor, part, npart=None):
    """fof.py:510-519: particles in group order -- every group a contiguous run starting at its
    first z-order member, groups ordered by that member, members in z-order -- and the group size
    at each group's first member (0 elsewhere).  Returns (partf, counts)."""
    pos = get_pos(part)
    n = len(igroup)
    dev = igroup.device
    if npart is None:
        npart = count_not_nan(pos)
    if not torch.is_tensor(npart):
        npart = torch.full((), int(npart), dtype=torch.int32, device=dev)
    npart = npart.to(torch.int32).reshape(()).contiguous()
    isort = torch.empty(n, dtype=torch.int32, device=dev)
    counts = torch.empty(n, dtype=torch.int32, device=dev)
    nb = L.LIB.jzt_fof_order_tmp_bytes(n)
    tmp = _tmp(nb, dev)
    L.call("jzt_fof_order", L.stream(), L.ptr(igroup.to(torch.int32).contiguous()), L.ptr(npart), n, L.ptr(isort),
           L.ptr(counts), L.ptr(tmp), nb)
    il = isort.long()
    partf = tree_map_by_len(lambda x: x[il], part, n)
    return partf, counts


def _fof_catalogue_from_groups(part, group_counts: torch.Tensor, cfg_cata: FofCatalogueConfig = FofCatalogueConfig(),
                               boxsize: float = 0., size_cata: int | None = None) -> FofCatalogue:
    """fof.py:650-766 on one GPU: the catalogue of the groups with at least cfg_cata.npart_min
    members, from particles in group order (_fof_order).  Properties the particles lack (mass,
    vel, v_rad, zplus1) are None; without masses every particle weighs 1."""
    pos = get_pos(part).contiguous()
    n = len(group_counts)
    dev = pos.device
    dtype = pos.dtype
    dim = pos.shape[-1]
    npart_min = int(cfg_cata.npart_min)
    if size_cata is None:
        size_cata = (n + npart_min - 1) // npart_min  # worst case
    mass = getattr(part, "mass", None)
    mass_arr, mass_scalar = None, 1.0
    if mass is not None:
        if torch.is_tensor(mass) and mass.numel() > 1:
            mass_arr = mass.to(dtype).contiguous()
        else:
            mass_scalar = float(mass)
    vel = getattr(part, "vel", None)
    vrad = getattr(part, "v_rad", None)
    zp1 = getattr(part, "zplus1", None)
    prep = lambda t: None if t is None else t.to(dtype).contiguous()
    vel, vrad, zp1 = prep(vel), prep(vrad), prep(zp1)
    f = lambda *shape: torch.empty(shape, dtype=dtype, device=dev)
    ngroups = torch.empty((), dtype=torch.int32, device=dev)
    count = torch.empty(size_cata, dtype=torch.int32, device=dev)
    offset = torch.empty(size_cata, dtype=torch.int32, device=dev)
    mass_out = f(size_cata) if mass is not None else None
    com_pos, radius = f(size_cata, dim), f(size_cata)
    com_vel = f(size_cata, dim) if vel is not None else None
    vrad_out = f(size_cata) if vrad is not None else None
    a_out = f(size_cata) if zp1 is not None else None
    nb = L.LIB.jzt_fof_catalogue_tmp_bytes(n, size_cata, dim, L.dtype_code(dtype))
    tmp = _tmp(nb, dev)
    L.call("jzt_fof_catalogue", L.stream(), L.ptr(pos), L.ptr(mass_arr), mass_scalar, L.ptr(vel), L.ptr(vrad),
           L.ptr(zp1), L.ptr(group_counts.to(torch.int32).contiguous()), n, size_cata, npart_min,
           float(boxsize or 0.), dim, L.dtype_code(dtype), L.ptr(ngroups), L.ptr(count), L.ptr(offset),
           L.ptr(mass_out), L.ptr(com_pos), L.ptr(radius), L.ptr(com_vel), L.ptr(vrad_out), L.ptr(a_out),
           L.ptr(tmp), nb)
    return FofCatalogue(ngroups=ngroups, mass=mass_out, count=count, offset=offset, com_pos=com_pos,
                        com_vel=com_vel, com_inertia_radius=radius, scale_factor=a_out, v_rad=vrad_out)


def fof_and_catalogue(part, rlink: float, boxsize: float = 0., cfg: FofConfig = FofConfig(),
                      th: TreeHierarchy | None = None):
    """Particles in FoF group order and the group catalogue (fof.py:872-907, one GPU).

    part: positions (N, dim) or a particle dataclass (pos and optionally mass -- an array or a
    constant --, vel, v_rad, zplus1); every per-particle field is reordered with the positions.
    Returns (partf, catalogue): each group a contiguous run of partf starting at offset, groups with
    at least cfg.catalogue.npart_min members catalogued."""
    try:
        partz, igroup = fof_labels(part, rlink=rlink, th=th, boxsize=boxsize, cfg=cfg)
        partf, counts = _fof_order(igroup, partz)
        cata = _fof_catalogue_from_groups(partf, counts, cfg.catalogue, boxsize=boxsize)
        checks.flush()
    except BaseException:
        checks.clear()
        raise
    return partf, cata


fof_and_catalogue.jit = fof_and_catalogue
