"""Small array helpers (src/jztree/tools.py:102-133), torch on the GPU."""
from __future__ import annotations

import torch


def div_ceil(a, b):
    return (a + b - 1) // b


def cumsum_starting_with_zero(x: torch.Tensor) -> torch.Tensor:
    """tools.py:105-106: pad(cumsum(x), (1, 0))."""
    cs = torch.cumsum(x, 0).to(x.dtype)
    return torch.cat([torch.zeros(1, dtype=x.dtype, device=x.device), cs])


def inverse_of_splits(ispl: torch.Tensor, size: int) -> torch.Tensor:
    """tools.py:122-125: given [0, 4, 7] returns [0,0,0,0,1,1,1] for size=7.

    Element i belongs to the last segment j with ispl[j] <= i (a binary search per element; the
    reference's scatter-add of ones onto the split positions is equivalent, but padded splits pile
    up on one position, which serialises device atomics)."""
    ar = torch.arange(size, dtype=ispl.dtype, device=ispl.device)
    return (torch.searchsorted(ispl.contiguous(), ar, right=True) - 1).to(torch.int32)


def inverse_indices(iargsort: torch.Tensor) -> torch.Tensor:
    """tools.py:122-126"""
    inv = torch.zeros_like(iargsort)
    inv[iargsort.long()] = torch.arange(len(iargsort), dtype=iargsort.dtype, device=iargsort.device)
    return inv


def masked_inverse(indices: torch.Tensor, mask: torch.Tensor) -> torch.Tensor:
    """tools.py:128-133: inverse permutation of the masked entries, len(indices) elsewhere."""
    n = len(indices)
    inverse = torch.full((n + 1,), n, dtype=indices.dtype, device=indices.device)
    imask = torch.where(mask, indices.long(), torch.full_like(indices, n, dtype=torch.int64))
    inverse[imask] = torch.arange(n, dtype=indices.dtype, device=indices.device)
    return inverse[:n]
