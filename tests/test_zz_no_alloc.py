"""jzt_pos_zorder_sort_noalloc (fixed device scratch, no workspace argument) equals
jzt_pos_zorder_sort with a caller workspace, bit for bit, on two calls in a row (the second with a
smaller input, so stale scratch contents of the first call must not leak into it), and refuses an
input larger than its scratch with a message."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _sort_with_tmp(L, x):
    n, dim = x.shape
    dt = L.dtype_code(x.dtype)
    out = torch.empty_like(x)
    ids = torch.empty(n, dtype=torch.int32 if x.dtype == torch.float32 else torch.int64, device=x.device)
    nb = L.LIB.jzt_zsort_tmp_bytes(n, dim, dt)
    tmp = torch.empty(nb, dtype=torch.uint8, device=x.device)
    L.call("jzt_pos_zorder_sort", L.stream(), L.ptr(x), L.ptr(out), L.ptr(ids), L.ptr(tmp), nb, n, dim, dt)
    return out, ids


def _sort_noalloc(L, x):
    n, dim = x.shape
    out = torch.empty_like(x)
    ids = torch.empty(n, dtype=torch.int32 if x.dtype == torch.float32 else torch.int64, device=x.device)
    L.call("jzt_pos_zorder_sort_noalloc", L.stream(), L.ptr(x), L.ptr(out), L.ptr(ids), n, dim,
           L.dtype_code(x.dtype))
    return out, ids


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_noalloc_sort_equals_workspace_sort(dev, dtype):
    import jztree_amd._lib as L
    rng = np.random.default_rng(11)
    for n in (1_100_000, 70_001):  # second call smaller than the first
        x = rng.random((n, 3)).astype(np.float32 if dtype == torch.float32 else np.float64)
        x[::97] = x[1::97]  # duplicated rows: equal keys go through the run fix-up
        x[5] = -x[5]  # a negative row: the general key layout
        t = torch.from_numpy(x).to(dev)
        print(f"[noalloc] dtype={dtype} n={n}", flush=True)
        p0, i0 = _sort_with_tmp(L, t)
        p1, i1 = _sort_noalloc(L, t)
        torch.cuda.synchronize()
        assert torch.equal(i0, i1), "sort permutations differ"
        assert torch.equal(p0, p1), "sorted positions differ"
        assert torch.equal(t[i1.long()], p1)


def test_noalloc_sort_rejects_too_large(dev):
    import jztree_amd._lib as L
    n = 1_500_000  # needs more than the 64 MiB of fixed scratch (f32 3D)
    x = torch.zeros((n, 3), dtype=torch.float32, device=dev)
    out = torch.empty_like(x)
    ids = torch.empty(n, dtype=torch.int32, device=dev)
    rc = L.LIB.jzt_pos_zorder_sort_noalloc(L.stream(), L.ptr(x), L.ptr(out), L.ptr(ids), n, 3, L.JZT_F32)
    assert rc != 0
    assert b"fixed" in L.LIB.jzt_last_error()
