"""Exact kNN for float64 positions (dims 2-10) and float32 positions in 4-10 dims (the reference
instantiates KnnLeaf2Leaf / KnnNode2Node for float and double over dims 2-10,
_generate_ffi.py:9-21): neighbour ids bit-exact against the oracle (ties by index), r2 bit-exact in
the positions' type, radii = correctly rounded sqrt of r2."""
import numpy as np
import pytest
import torch

from gen import clustered, uniform

pytestmark = pytest.mark.gpu


def _knn(x, k, dev, **kw):
    import jztree_amd as jz
    print(f"[knn-nd] n={len(x)} dim={x.shape[1]} dtype={x.dtype} k={k} {kw}", flush=True)
    t = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    rnn, inn = jz.knn.knn(t, k=k, **kw)
    torch.cuda.synchronize()
    return rnn.cpu().numpy(), inn.cpu().numpy()


def _check(x, k, rnn, inn, oracle, boxsize=0.0, xq=None):
    r2o, io = oracle.knn(x, k, xq=xq, boxsize=boxsize)
    assert rnn.dtype == x.dtype
    assert np.array_equal(inn.astype(np.int64), io), f"{np.sum(inn != io)} ids differ"
    assert np.array_equal(rnn, np.sqrt(r2o))


def _cfg(dim):
    """Interaction lists grow quickly with the dimension (a leaf's search ball touches ever more
    leaves): above 4 dims the list allocation per leaf is raised from 256 to 2048 entries."""
    import dataclasses
    from jztree_amd.config import KNNConfig
    return dataclasses.replace(KNNConfig(), alloc_fac_ilist=256. if dim <= 4 else 2048.)


def _f64(n, dim, seed):
    return np.random.default_rng(seed).random((n, dim))


@pytest.mark.parametrize("k", [1, 8, 20, 32, 45, 70])
@pytest.mark.parametrize("dim", [2, 3])
def test_knn_f64_k_sweep(dev, oracle, dim, k):
    """float64 in 2 and 3 dims, k from 1 to 70: up to 32 one pass with the start bound, above it
    passes of 32 without."""
    x = _f64(1 << 14, dim, seed=k + dim)
    rnn, inn = _knn(x, k, dev)
    _check(x, k, rnn, inn, oracle)


@pytest.mark.parametrize("boxsize", [1.0, 0.5])
def test_knn_f64_periodic(dev, oracle, boxsize):
    """Periodic float64 box: the wrap uses the float half box like the reference."""
    x = _f64(1 << 14, 3, seed=7) * boxsize
    rnn, inn = _knn(x, 16, dev, boxsize=boxsize)
    _check(x, 16, rnn, inn, oracle, boxsize=boxsize)


def test_knn_f64_clustered_duplicates(dev, oracle):
    """Clustered float64 points (float32 halos plus sub-float32 jitter) with exact duplicates: ties
    broken by index."""
    rng = np.random.default_rng(3)
    x = clustered(1 << 15, nhalos=64, seed=3).astype(np.float64) + 1e-9 * rng.random((1 << 15, 3))
    x = np.concatenate([x, x[:2000]])
    import dataclasses
    from jztree_amd.config import KNNConfig
    cfg = KNNConfig()
    cfg = dataclasses.replace(cfg, tree=dataclasses.replace(cfg.tree, alloc_fac_nodes=2.0))  # strong clustering
    rnn, inn = _knn(x, 24, dev, cfg=cfg)
    _check(x, 24, rnn, inn, oracle)


@pytest.mark.parametrize("k", [8, 32, 70])
@pytest.mark.parametrize("dim", [4, 5, 7, 10])
def test_knn_f32_high_dims(dev, oracle, dim, k):
    """float32 positions in 4-10 dims (the any-dim kernel): k = 70 runs two passes of 64."""
    n = 1 << 14 if dim <= 5 else 1 << 13
    x = uniform(n, dim=dim, seed=10 * dim + k)
    rnn, inn = _knn(x, k, dev, cfg=_cfg(dim))
    _check(x, k, rnn, inn, oracle)


@pytest.mark.parametrize("dim", [5, 10])
def test_knn_f64_high_dims(dev, oracle, dim):
    """float64 positions in 5 and 10 dims, open and periodic."""
    x = _f64(1 << 13, dim, seed=dim)
    for box in (0.0, 1.0):
        rnn, inn = _knn(x, 16, dev, boxsize=box or None, cfg=_cfg(dim))
        _check(x, 16, rnn, inn, oracle, boxsize=box)


def test_knn_f64_part_query(dev, oracle):
    """Separate float64 query points (one tree over sources and queries)."""
    import jztree_amd as jz
    x = _f64(1 << 14, 3, seed=11)
    q = _f64(5000, 3, seed=12)
    print("[knn-nd] part_query f64", flush=True)
    rnn, inn = jz.knn.knn(torch.from_numpy(x).to(dev), 12, part_query=torch.from_numpy(q).to(dev))
    _check(x, 12, rnn.cpu().numpy(), inn.cpu().numpy(), oracle, xq=q)


@pytest.mark.parametrize("dim,dtype", [(3, torch.float32), (10, torch.float32), (3, torch.float64),
                                       (5, torch.float64), (10, torch.float64)])
def test_regularization_level_wide_ranges(dev, dim, dtype):
    """find_regularization_level (tree.py:160-187) over the full level range of the position type:
    float64 from 4 dims on has more levels than the LDS histogram holds and counts in global
    memory.  The closed forms do not depend on the range: all weight at level -60 gives -55 (mean
    volume 1, first volume >= 20 at +5); half the weight at -40 gives -36."""
    import jztree_amd as jz
    from jztree_amd.data import LevelInfo
    linfo = LevelInfo(dim, dtype)
    print(f"[reg] dim={dim} dtype={dtype} levels {linfo.min_lvl()}..{linfo.max_lvl()}", flush=True)
    lvl = torch.full((1000,), -60, dtype=torch.int32, device=dev)
    w = torch.ones(1000, dtype=torch.int32, device=dev)
    assert int(jz.tree.find_regularization_level(lvl, linfo, w, max_vol_fac=20., min_percentile=90.)[0]) == -55
    lvl2 = torch.cat([lvl, torch.full((5,), -40, dtype=torch.int32, device=dev)])
    w2 = torch.cat([w, torch.full((5,), 200, dtype=torch.int32, device=dev)])
    assert int(jz.tree.find_regularization_level(lvl2, linfo, w2, max_vol_fac=20., min_percentile=90.)[0]) == -36


@pytest.mark.parametrize("dim", [4, 7])
def test_knn_f64_keys_f32_distances_high_dims(dev, oracle, dim):
    """float64 keys and tree, float32 distance kernels (dist_dtype) in 4 and 7 dims: the exact kNN
    of the float32-rounded points, including a clump below float32 resolution."""
    rng = np.random.default_rng(60 + dim)
    x = np.concatenate([rng.random((6000, dim)), 0.3 + 1e-9 * rng.random((500, dim))])
    x32 = x.astype(np.float32)
    rnn, inn = _knn(x, 12, dev, cfg=_cfg(dim), dist_dtype=torch.float32)
    assert rnn.dtype == np.float32
    r2o, io = oracle.knn(x32, 12)
    assert np.array_equal(inn.astype(np.int64), io), f"{np.sum(inn != io)} ids differ"
    assert np.array_equal(rnn, np.sqrt(r2o))


@pytest.mark.parametrize("dim,dtype", [(5, np.float32), (4, np.float64)])
def test_knn_part_query_high_dims(dev, oracle, dim, dtype):
    """Separate query points in 4-5 dims (one two-type tree over sources and queries)."""
    import jztree_amd as jz
    rng = np.random.default_rng(70 + dim)
    x = rng.random((12000, dim)).astype(dtype)
    q = rng.random((3000, dim)).astype(dtype)
    print(f"[knn-nd] part_query dim={dim} {np.dtype(dtype).name}", flush=True)
    rnn, inn = jz.knn.knn(torch.from_numpy(x).to(dev), 10, part_query=torch.from_numpy(q).to(dev), cfg=_cfg(dim))
    _check(x, 10, rnn.cpu().numpy(), inn.cpu().numpy(), oracle, xq=q)


def test_knn_double_like_reference(dev, oracle):
    """checks/tests/test_knn.py:76-80: 2^18 float64 points in [0, 10), k = 13."""
    x = np.random.default_rng(76).uniform(0.0, 10.0, (1 << 18, 3))
    rnn, inn = _knn(x, 13, dev)
    _check(x, 13, rnn, inn, oracle)


@pytest.mark.parametrize("boxsize", [0.03, 170.0])
def test_knn_f64_periodic_box_sizes(dev, oracle, boxsize):
    """checks/tests/test_knn.py:83-88 on float64 positions."""
    x = _f64(1 << 14, 3, seed=88) * boxsize
    x[x >= boxsize] = 0.0
    rnn, inn = _knn(x, 16, dev, boxsize=boxsize)
    _check(x, 16, rnn, inn, oracle, boxsize=boxsize)
