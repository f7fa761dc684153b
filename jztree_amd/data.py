"""Particle and tree data structures (torch tensors on the GPU).

Mirrors src/jztree/data.py: the particle dataclasses (:25-63), the accessors (:69-94),
PackedArray (:251-345), LevelInfo (:352-379), TreeHierarchy (:381-452), InteractionList
(:458-514) and RankIdx (:681-686).  Array shapes and contents are those of the reference; the
containers hold torch tensors instead of jax arrays.  Sizes that JAX keeps static are Python
ints here; counts that JAX keeps traced stay device tensors (no host synchronisation).
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from typing import Any, List

import numpy as np
import torch


def same_width_int(dtype: torch.dtype) -> torch.dtype:
    """data.py:13-19"""
    size = torch.empty((), dtype=dtype).element_size()
    if size == 4:
        return torch.int32
    if size == 8:
        return torch.int64
    raise ValueError(f"unsupported type {dtype}")


# ------------------------------------------------------------------------------------------------
# user-facing particle containers
# ------------------------------------------------------------------------------------------------

@dataclass(kw_only=True)
class Pos:
    """Positions (size, dim); num/num_total only matter for padded or distributed data."""
    pos: torch.Tensor
    num: Any = None
    num_total: int | None = None


@dataclass(kw_only=True)
class PosMass:
    pos: torch.Tensor
    mass: torch.Tensor
    num: Any = None
    num_total: int | None = None


@dataclass(kw_only=True)
class ParticleData:
    pos: torch.Tensor
    mass: Any = None
    vel: torch.Tensor | None = None
    id: torch.Tensor | None = None
    num: Any = None
    num_total: int | None = None


@dataclass
class RankIdx:
    """Rank and index of a particle's origin (data.py:681-686)."""
    rank: torch.Tensor | None
    idx: torch.Tensor


def get_pos(part) -> torch.Tensor:
    if torch.is_tensor(part):
        assert part.shape[-1] <= 10 and part.ndim == 2
        return part
    if hasattr(part, "pos"):
        assert part.pos.ndim == 2
        return part.pos
    raise ValueError("Invalid input particles")


def get_num(part, default_to_length=False, default_to_nancount=True):
    if getattr(part, "num", None) is None:
        if default_to_length:
            return len(get_pos(part))
        if default_to_nancount:
            return count_not_nan(get_pos(part))
        raise ValueError("Need .num attribute in particle structure for distributed mode")
    num = part.num
    return torch.sum(num) if torch.is_tensor(num) else num


def count_not_nan(pos: torch.Tensor) -> torch.Tensor:
    """Number of rows whose first coordinate is not NaN, as a device int32 scalar: one library
    kernel (a torch reduction over many blocks would clear its scratch with a memset, which a
    captured HIP graph should not hold)."""
    from . import _lib as L
    pos = pos.contiguous()
    out = torch.empty((), dtype=torch.int32, device=pos.device)
    L.call("jzt_count_not_nan", L.stream(), L.ptr(pos), pos.shape[0], pos.shape[-1], L.dtype_code(pos.dtype),
           L.ptr(out))
    return out


def get_num_total(part, default_to_length=False):
    num = getattr(part, "num_total", None)
    if num is not None:
        return num
    if default_to_length:
        return len(get_pos(part))
    raise ValueError("Need num_total attribute on particle data structure (for distributed mode)")


def tree_map_by_len(fn, tree, length: int):
    """Apply fn to every tensor leaf whose leading dimension equals `length` (jax_ext.py:58-71)."""
    if tree is None:
        return None
    if torch.is_tensor(tree):
        return fn(tree) if (tree.ndim > 0 and tree.shape[0] == length) else tree
    if isinstance(tree, (tuple, list)):
        return type(tree)(tree_map_by_len(fn, t, length) for t in tree)
    if isinstance(tree, dict):
        return {k: tree_map_by_len(fn, v, length) for k, v in tree.items()}
    if dataclasses.is_dataclass(tree):
        vals = {f.name: tree_map_by_len(fn, getattr(tree, f.name), length) for f in dataclasses.fields(tree)}
        return dataclasses.replace(tree, **vals)
    return tree


# ------------------------------------------------------------------------------------------------
# PackedArray
# ------------------------------------------------------------------------------------------------

@dataclass
class PackedArray:
    """Several per-plane arrays stacked in one buffer (data.py:251-345).

    data: the buffer, ispl: (levels+1,) device start offsets, fill_values: (levels,) device values
    returned for out-of-range reads."""
    data: torch.Tensor
    ispl: torch.Tensor
    fill_values: torch.Tensor | None = None
    levels_filled: torch.Tensor | None = None

    @classmethod
    def from_data(cls, data, ispl, fill_values=None, levels_filled=None):
        if levels_filled is None:
            levels_filled = torch.full((1,), len(ispl) - 1, dtype=torch.int32, device=data.device)
        if fill_values is not None and not torch.is_tensor(fill_values):
            fill_values = torch.full((len(ispl) - 1,), fill_values, dtype=data.dtype, device=data.device)
        elif fill_values is not None:
            fill_values = fill_values.to(data.dtype)
        return cls(data, ispl, fill_values, levels_filled)

    def get(self, level, size=None, fill_value=None):
        if size is None:
            size = self.size()
        fast = self._get_device(level, size, fill_value)
        if fast is not None:
            return fast
        idx = torch.arange(size, dtype=torch.int64, device=self.data.device) + self.ispl[level]
        valid = idx < self.ispl[level + 1]
        vals = self.data[torch.clamp(idx, max=len(self.data) - 1)]
        if fill_value is None:
            fill = self.fill_values[level]
        else:
            fill = (fill_value.to(device=self.data.device, dtype=self.data.dtype) if torch.is_tensor(fill_value)
                    else torch.full((), fill_value, dtype=self.data.dtype, device=self.data.device))
        if self.data.ndim > 1:
            valid = valid.reshape((-1,) + (1,) * (self.data.ndim - 1))
        return torch.where(valid, vals, fill)

    _GET_DTYPES = {torch.int32: 4, torch.float32: 4, torch.int64: 8, torch.float64: 8}

    def _get_device(self, level, size, fill_value):
        """get() as one kernel (jzt_packed_get) for device data of 4- or 8-byte elements and a
        scalar fill; None where that does not apply (the caller then uses the torch ops)."""
        d = self.data
        esize = self._GET_DTYPES.get(d.dtype)
        if not isinstance(level, int) or not isinstance(size, int) or not 0 <= level < len(self.ispl) - 1:
            return None
        if not d.is_cuda or esize is None or self.ispl.dtype not in (torch.int32, torch.int64) or not self.ispl.is_cuda:
            return None
        if fill_value is None:
            if (self.fill_values is None or not self.fill_values.is_cuda or self.fill_values.dtype != d.dtype
                    or level >= len(self.fill_values)):
                return None
            fill = self.fill_values.contiguous()[level:level + 1]
        elif torch.is_tensor(fill_value):
            if fill_value.numel() != 1:
                return None
            fill = fill_value.to(device=d.device, dtype=d.dtype).reshape(1).contiguous()
        else:
            fill = torch.full((1,), fill_value, dtype=d.dtype, device=d.device)
        from . import _lib as L
        d = d.contiguous()
        row = int(np.prod(d.shape[1:])) if d.ndim > 1 else 1
        out = torch.empty((size,) + tuple(d.shape[1:]), dtype=d.dtype, device=d.device)
        L.call("jzt_packed_get", L.stream(), L.ptr(d), d.shape[0], L.ptr(self.ispl.contiguous()),
               int(self.ispl.dtype == torch.int64), int(level), L.ptr(fill), esize, max(row, 1), L.ptr(out), int(size))
        return out

    def set(self, level, values, num=None, fill_value=None):
        """data.py:300-307 (values written at ispl[level], later levels start after them)."""
        if num is None:
            num = values.shape[0]
        lvl_idx = torch.arange(len(self.ispl), device=self.ispl.device)
        start = self.ispl[level]
        new_spl = torch.where(lvl_idx <= level, self.ispl, start + num).to(self.ispl.dtype)
        new_data = self.data.clone()
        n = values.shape[0]
        pos = start + torch.arange(n, device=self.data.device)
        keep = torch.arange(n, device=self.data.device) < num
        dump = torch.cat([new_data, new_data[:1]])  # one spare slot for masked writes
        tgt = torch.where(keep & (pos < len(self.data)), pos, torch.full_like(pos, len(self.data)))
        dump[tgt] = values.to(self.data.dtype)
        new_data = dump[:len(self.data)]
        fv = self.fill_values
        if fill_value is not None:
            fv = fv.clone()
            if torch.is_tensor(fill_value):
                fv[level] = fill_value.to(device=fv.device, dtype=fv.dtype)
            else:
                fv[level:level + 1].fill_(fill_value)  # a fill kernel, no host copy
        lf = torch.full((1,), level + 1, dtype=torch.int32, device=self.data.device)
        return PackedArray(new_data, new_spl, fv, lf)

    def append(self, values, num=None, fill_value=None, resize=False):
        arr = self.resize_levels(self.nlevels() + 1) if resize else self
        return arr.set(self.nlevels(), values, num, fill_value)

    def resize_levels(self, levels):
        ispl_new = torch.full((levels + 1,), 0, dtype=self.ispl.dtype, device=self.ispl.device)
        ispl_new[:] = self.ispl[-1]
        ispl_new[:len(self.ispl)] = self.ispl
        fv = self.fill_values
        if fv is not None:
            reps = int(np.ceil(levels / max(len(fv), 1)))
            fv = fv.repeat(reps)[:levels]  # jnp.resize semantics
        return PackedArray(self.data, ispl_new, fv, self.levels_filled)

    def size(self):
        return len(self.data)

    def num(self, level):
        return self.ispl[level + 1] - self.ispl[level]

    def nfilled(self):
        return self.ispl[-1]

    def nlevels(self):
        return len(self.ispl) - 1


# ------------------------------------------------------------------------------------------------
# tree structure
# ------------------------------------------------------------------------------------------------

def _min_max_msb_diff(dtype):
    if dtype == torch.float32:
        return -150, 128
    if dtype == torch.float64:
        return -1075, 1024
    if dtype == torch.int32:
        return -1, 32
    if dtype == torch.int64:
        return -1, 64
    raise ValueError(dtype)


@dataclass(frozen=True)
class LevelInfo():
    """Minimum and maximum Morton level of a tree (data.py:352-379)."""
    dim: int
    dtype: torch.dtype

    def min_lvl(self) -> int:
        return self.dim * _min_max_msb_diff(self.dtype)[0]

    def max_lvl(self) -> int:
        return self.dim * (_min_max_msb_diff(self.dtype)[1] + 1)


@dataclass
class TreeHierarchy():
    """The tree-plane hierarchy (data.py:381-452); see the reference for the field meanings."""
    size_leaves: int
    ispl_n2n: PackedArray
    ispl_n2l: PackedArray
    ispl_l2p_per_type: torch.Tensor
    lvl: PackedArray
    geom_cent: PackedArray
    mass: PackedArray | None = None
    mass_cent: PackedArray | None = None
    # batches of independent sets (extension): set id of every leaf slot and the number of sets
    leaf_set: torch.Tensor | None = None
    nsets: int = 1

    def splits_leaf_to_part(self, ptype: int = 0, size: int | None = None) -> torch.Tensor:
        if size is None:
            size = self.size() + 1
        return self.ispl_l2p_per_type[ptype][:size]

    def npart(self, level: int, ptype: int = 0, size=None) -> torch.Tensor:
        if size is None:
            size = self.size()
        spl = self.splits_leaf_to_part(ptype)
        n2l = self.ispl_n2l.get(level, size + 1).long()
        n2p = spl[torch.clamp(n2l, max=len(spl) - 1)]
        return n2p[1:] - n2p[:-1]

    def center(self) -> PackedArray:
        return self.mass_cent if self.mass_cent is not None else self.geom_cent

    def num_planes(self) -> int:
        return len(self.ispl_n2l.ispl) - 1

    def num(self, level):
        return self.lvl.num(level)

    def size(self) -> int:
        return self.ispl_n2n.size() - 1

    def info(self) -> LevelInfo:
        return LevelInfo(self.geom_cent.data.shape[-1], self.geom_cent.data.dtype)


@dataclass
class InteractionList:
    """Receiver i interacts with sources isrc[ispl[i]:ispl[i+1]] (data.py:458-507)."""
    ispl: torch.Tensor
    isrc: torch.Tensor
    rad2: torch.Tensor | None = None
    ids: torch.Tensor | None = None
    dev_spl: torch.Tensor | None = None

    def nfilled(self):
        return self.ispl[-1]

    def size(self):
        return self.isrc.numel()

    def dtype(self):
        return self.isrc.dtype


def verify_ilist(ilist: InteractionList) -> InteractionList:
    """data.py:509-514: int32 interaction lists must stay below 2**31 entries."""
    if ilist.isrc.numel() >= 2 ** 31:
        raise ValueError("Intercation list allocation too large and may produce integer overflows",
                         "Hint: Use more GPUs or decrease alloc_fac_ilist (if possible)")
    return ilist


# ------------------------------------------------------------------------------------------------
# FoF catalogue (data.py:573-680 of the reference)
# ------------------------------------------------------------------------------------------------

@dataclass
class FofCatalogue:
    """Per-group properties of a FoF catalogue, arrays of the catalogue's allocated size with
    `ngroups` (device int32 scalar) filled; optional properties are None when the particles lack
    the data.  count/offset: members and first member in the group-ordered particle array;
    com_pos/com_vel: centre of mass and mean velocity; com_inertia_radius: sqrt of the mean squared
    distance to the centre of mass; scale_factor/v_rad: mass-weighted means for light-cone data;
    offset_rank: only set on catalogues squeezed from several devices."""
    ngroups: torch.Tensor
    mass: torch.Tensor | None = None
    count: torch.Tensor | None = None
    offset: torch.Tensor | None = None
    com_pos: torch.Tensor | None = None
    com_vel: torch.Tensor | None = None
    com_inertia_radius: torch.Tensor | None = None
    scale_factor: torch.Tensor | None = None
    v_rad: torch.Tensor | None = None
    offset_rank: torch.Tensor | None = None


def squeeze_catalogue(cata: FofCatalogue, size_out: int | None = None, offset_mode: str = "rank",
                      nparts=None) -> FofCatalogue:
    """The filled part of a catalogue (data.py:619-663).  On one GPU this only drops the slots past
    ngroups (size_out defaults to ngroups, read back from the device)."""
    assert offset_mode in ("rank", "global")
    if cata.count.ndim != 1:
        raise NotImplementedError("catalogues of several devices are not part of this one-GPU build")
    if size_out is None:
        size_out = int(cata.ngroups)
    out = tree_map_by_len(lambda x: x[:size_out], cata, cata.count.shape[0])
    return out


def sort_catalogue(cata: FofCatalogue, by: str = "count", descending: bool = True) -> FofCatalogue:
    """Catalogue entries ordered by one property (data.py:665-673), stable: equal keys keep their
    order (as jnp.argsort's stable descending sort)."""
    key = getattr(cata, by, None)
    assert key is not None, f"catalogue doesn't have {by}"
    assert cata.count.ndim == 1, "Please squeeze catalogue before sorting"
    isort = torch.argsort(key, descending=descending, stable=True)
    return tree_map_by_len(lambda x: x[isort], cata, len(isort))


def catalogues_equal(c1: FofCatalogue, c2: FofCatalogue) -> bool:
    """Every property equal (data.py:675-679)."""
    for f in dataclasses.fields(FofCatalogue):
        a, b = getattr(c1, f.name), getattr(c2, f.name)
        if (a is None) != (b is None):
            return False
        if a is not None and not bool(torch.equal(a, b)):
            return False
    return True
