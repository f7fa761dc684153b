"""Tree kernels and hierarchy on the GPU vs the oracle's restatement (bit-exact)."""
import numpy as np
import pytest
import torch

from gen import clustered, uniform

pytestmark = pytest.mark.gpu


def _t(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


@pytest.mark.parametrize("leaf", [16, 32, 48])
def test_flag_leaf_boundaries(dev, oracle, leaf):
    import jztree_amd as jz
    x = clustered(1 << 16, nhalos=64, seed=leaf)
    pz, _ = oracle.zsort(x)
    print(f"[flag_leaf] leaf={leaf}", flush=True)
    spl = jz.tree.detect_leaf_boundaries(_t(pz, dev), leaf_size=leaf)
    torch.cuda.synchronize()
    flags, lvl = oracle.flag_leaf_boundaries(pz, leaf_size=leaf)
    ref = np.nonzero(flags)[0]
    got = spl.cpu().numpy()
    assert np.array_equal(got[:len(ref)], ref)
    assert np.all(got[len(ref):] == len(pz))


def test_node_boundaries_and_geometry(dev, oracle):
    import jztree_amd as jz
    x = uniform(1 << 15, seed=3)
    pz, _ = oracle.zsort(x)
    flags, _ = oracle.flag_leaf_boundaries(pz, leaf_size=32)
    spl = np.nonzero(flags)[0]
    reps = pz[spl[:-1]]
    print("[find_node_boundaries]", flush=True)
    lv, lb, rb = jz.tree.determine_znode_boundaries(_t(reps, dev), nleaves=len(spl) - 1)
    olv, olb, orb = oracle.find_node_boundaries(reps, len(spl) - 1)
    assert np.array_equal(lv.cpu().numpy(), olv)
    assert np.array_equal(lb.cpu().numpy(), olb)
    assert np.array_equal(rb.cpu().numpy(), orb)
    ispl = np.sort(np.random.default_rng(42).integers(0, len(pz), 300)).astype(np.int32)
    for upper in (False, True):
        print(f"[get_node_geometry] upper={upper}", flush=True)
        g = jz.tree.get_node_geometry(_t(pz, dev), _t(ispl[:-1], dev), _t(ispl[1:], dev), upper_extent=upper)
        o = oracle.get_node_geometry(pz, ispl[:-1], ispl[1:], upper_extent=upper)
        for a, b in zip(g, o):
            assert np.array_equal(a.cpu().numpy(), b, equal_nan=True)


@pytest.mark.parametrize("cfg_name", ["knn", "fof", "default"])
def test_build_tree_hierarchy(dev, oracle, cfg_name):
    import jztree_amd as jz
    from jztree_amd.config import FofConfig, KNNConfig, TreeConfig
    import dataclasses
    cfg = {"knn": KNNConfig().tree, "fof": FofConfig().tree, "default": TreeConfig()}[cfg_name]
    cfg = dataclasses.replace(cfg, alloc_fac_nodes=1.3)  # strongly clustered input needs more leaves
    x = clustered(1 << 18, nhalos=512, seed=7)
    pz, _ = oracle.zsort(x)
    reg = None if cfg.regularization is None else (cfg.regularization.max_volume_fac,
                                                   cfg.regularization.regularize_percentile)
    print(f"[build_tree_hierarchy] {cfg_name}", flush=True)
    th = jz.tree.build_tree_hierarchy(_t(pz, dev), cfg)
    jz.checks.flush()
    oth = oracle.build_tree_hierarchy(pz, cfg.max_leaf_size, cfg.coarse_fac, cfg.stop_coarsen,
                                      cfg.alloc_fac_nodes, regularize=reg)
    assert th.size_leaves == oth.size_leaves
    for name in ("ispl_n2n", "ispl_n2l", "lvl", "geom_cent"):
        a, b = getattr(th, name), getattr(oth, name)
        assert np.array_equal(a.ispl.cpu().numpy(), b.ispl), name
        assert np.array_equal(a.data.cpu().numpy(), b.data, equal_nan=True), name
        assert np.array_equal(a.fill_values.cpu().numpy(), b.fill_values, equal_nan=True), name
    assert np.array_equal(th.ispl_l2p_per_type.cpu().numpy(), oth.ispl_l2p_per_type)


def test_tree_nan_padding_invariance(dev):
    """checks/tests/test_tree.py:135-153 (geometric tree): NaN padding leaves the hierarchy unchanged."""
    import jztree_amd as jz
    from jztree_amd.config import TreeConfig
    from jztree_amd.data import Pos
    x = clustered(1 << 16, nhalos=64, seed=2)
    pz, _ = jz.tree.zsort(_t(x, dev))
    pad = torch.cat([pz, torch.full((1356, 3), float("nan"), device=dev)])
    th1 = jz.tree.build_tree_hierarchy(pz, TreeConfig())
    th2 = jz.tree.build_tree_hierarchy(Pos(pos=pad, num_total=len(pz)), TreeConfig())
    for name in ("ispl_n2n", "ispl_n2l", "lvl", "geom_cent"):
        a, b = getattr(th1, name), getattr(th2, name)
        assert torch.equal(a.ispl, b.ispl), name
        assert torch.equal(torch.nan_to_num(a.data, nan=-7.0), torch.nan_to_num(b.data, nan=-7.0)), name


def test_leaf_allocation_overflow_matches_oracle(dev, oracle):
    """tree.py:251-255 error path: the regularised kNN tree of a strongly clustered set overflows the
    default leaf allocation; the GPU path raises with the same counts as the oracle."""
    import jztree_amd as jz
    from jztree_amd.config import KNNConfig
    x = clustered(1 << 18, nhalos=512, seed=7)
    pz, _ = oracle.zsort(x)
    with pytest.raises(RuntimeError) as e_ref:
        oracle.build_tree_hierarchy(pz, 48, 8.0, 2048, 1.0, regularize=(20.0, 90.0))
    with pytest.raises(RuntimeError) as e_gpu:
        jz.tree.build_tree_hierarchy(_t(pz, dev), KNNConfig().tree)
        jz.checks.flush()
    assert str(e_ref.value).split(".")[0] in str(e_gpu.value)


@pytest.mark.parametrize("dtype,row", [(torch.int32, ()), (torch.float32, (3,)), (torch.float64, (3,)),
                                       (torch.int64, ()), (torch.float32, ())])
@pytest.mark.parametrize("ispl_dtype", [torch.int32, torch.int64])
def test_packed_get_kernel_matches_torch_ops(dev, dtype, row, ispl_dtype):
    """PackedArray.get on the device (one jzt_packed_get launch) equals the index + select torch
    ops that the same call runs on CPU tensors, for every plane, sizes past the plane's end, and
    fills from the array, a scalar and a one-element tensor."""
    from jztree_amd.data import PackedArray
    g = torch.Generator().manual_seed(3)
    n = 1000
    data = (torch.randn((n,) + row, generator=g) * 100).to(dtype)
    ispl = torch.tensor([0, 10, 10, 400, 997], dtype=ispl_dtype)
    fills = torch.tensor([-1, 7, -2000, 5], dtype=torch.float64).to(dtype)
    cpu = PackedArray(data, ispl, fills)
    gpu = PackedArray(data.to(dev), ispl.to(dev), fills.to(dev))
    for level in range(4):
        for size in (0, 1, 5, 390, 1000, 1500):
            for fv in (None, 3, torch.tensor(9, dtype=torch.int32)):
                a = cpu.get(level, size, fv)
                b = gpu.get(level, size, fv)
                assert b.dtype == a.dtype and b.shape == a.shape
                assert torch.equal(b.cpu(), a), (level, size, fv)


def test_zsort_and_tree_origin_equals_gathered_arange(dev):
    """The z-ordered origin indices taken from the sort permutation equal the gathered arange of
    zsort_and_tree(data=...), and the trees are the same."""
    import dataclasses
    import jztree_amd as jz
    x = torch.from_numpy(clustered(1 << 15, nhalos=32, seed=4)).to(dev)
    cfg = dataclasses.replace(jz.config.KNNConfig().tree, alloc_fac_nodes=2.0)
    ridx = jz.data.RankIdx(rank=None, idx=torch.arange(len(x), dtype=torch.int32, device=dev))
    pz1, o1, th1 = jz.tree.zsort_and_tree(x, cfg, data=ridx)
    pz2, o2, th2 = jz.tree.zsort_and_tree_origin(x, cfg)
    assert o2.idx.dtype == o1.idx.dtype and torch.equal(o1.idx, o2.idx)
    assert torch.equal(pz1, pz2)
    assert torch.equal(th1.ispl_n2n.data, th2.ispl_n2n.data) and torch.equal(th1.lvl.data, th2.lvl.data)


@pytest.mark.parametrize("dim", [4, 5, 6, 7, 8, 9, 10])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_zsort_high_dims(dev, oracle, dim, dtype):
    """PosZorderSort in 4-10 dimensions (the reference's instances, _generate_ffi.py:9-21):
    positions and permutation equal the oracle's exact Morton order, ties by index -- uniform
    points, a tight clump (most coordinates share their leading bits), duplicates and negative
    coordinates."""
    import jztree_amd as jz
    rng = np.random.default_rng(100 + dim)
    x = np.concatenate([rng.random((30000, dim)), 0.5 + 1e-5 * rng.random((3000, dim)),
                        np.repeat(rng.random((50, dim)), 4, axis=0), rng.normal(size=(2000, dim))]).astype(dtype)
    x = x[rng.permutation(len(x))]
    print(f"[zsort] dim={dim} dtype={np.dtype(dtype).name} n={len(x)}", flush=True)
    pz, idz = jz.tree.zsort(torch.from_numpy(x).to(dev))
    opz, oidz = oracle.zsort(x)
    assert np.array_equal(idz.cpu().numpy().astype(np.int64), oidz)
    assert np.array_equal(pz.cpu().numpy(), opz)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("dim", [2, 3, 5, 10])
def test_zsort_int_positions(dev, oracle, dim, dtype):
    """PosZorderSort and SearchSortedZ on integer positions (the reference's int32_t/int64_t sort
    instances, _generate_ffi.py:9-21): permutation and positions equal the oracle's exact
    Morton order of the signed integers -- the full range with negatives, a small-range clump,
    duplicates -- and insertion at the search_sorted_z indices keeps the z-order."""
    import jztree_amd as jz
    rng = np.random.default_rng(300 + dim)
    info = np.iinfo(dtype)
    x = np.concatenate([rng.integers(info.min, info.max, (20000, dim), dtype=dtype, endpoint=True),
                        rng.integers(-40, 40, (5000, dim)).astype(dtype),
                        (1000 + rng.integers(0, 8, (3000, dim))).astype(dtype)])
    x = np.concatenate([x, x[:500]])
    x = x[rng.permutation(len(x))]
    print(f"[zsort-int] dim={dim} dtype={np.dtype(dtype).name} n={len(x)}", flush=True)
    pz, idz = jz.tree.zsort(torch.from_numpy(x).to(dev))
    opz, oidz = oracle.zsort(x)
    assert np.array_equal(idz.cpu().numpy().astype(np.int64), oidz)
    assert np.array_equal(pz.cpu().numpy(), opz)
    q = np.concatenate([rng.integers(-60, 60, (1500, dim)), rng.integers(info.min, info.max, (1500, dim),
                                                                          dtype=dtype)]).astype(dtype)
    qz, _ = oracle.zsort(q)
    ins = jz.tree.search_sorted_z(_t(opz, dev), _t(qz, dev)).cpu().numpy()
    merged = np.insert(opz, ins, qz, axis=0)
    mz, _ = oracle.zsort(merged)
    assert np.array_equal(merged, mz), "insertion at the returned indices breaks the z-order"
    assert np.all(np.diff(ins) >= 0)


@pytest.mark.parametrize("dim", [4, 7, 10])
def test_build_tree_hierarchy_high_dims(dev, oracle, dim):
    """Leaves, node planes, levels and geometric centres in 4, 7 and 10 dimensions bit-exact
    against the oracle's assembly (single-set tree kernels are instantiated for dims 2-10), and
    search_sorted_z keeps the z-order in those dimensions."""
    import jztree_amd as jz
    from jztree_amd.config import TreeConfig
    import dataclasses
    rng = np.random.default_rng(200 + dim)
    x = np.concatenate([rng.random((40000, dim)), 0.25 + 0.01 * rng.random((20000, dim))]).astype(np.float32)
    pz, _ = oracle.zsort(x)
    cfg = dataclasses.replace(TreeConfig(), alloc_fac_nodes=1.5, regularization=None)
    print(f"[tree] dim={dim}", flush=True)
    th = jz.tree.build_tree_hierarchy(_t(pz, dev), cfg)
    jz.checks.flush()
    oth = oracle.build_tree_hierarchy(pz, cfg.max_leaf_size, cfg.coarse_fac, cfg.stop_coarsen,
                                      cfg.alloc_fac_nodes, regularize=None)
    assert th.size_leaves == oth.size_leaves
    for name in ("ispl_n2n", "ispl_n2l", "lvl", "geom_cent"):
        a, b = getattr(th, name), getattr(oth, name)
        assert np.array_equal(a.ispl.cpu().numpy(), b.ispl), name
        assert np.array_equal(a.data.cpu().numpy(), b.data, equal_nan=True), name
    q = rng.random((3000, dim)).astype(np.float32)
    qz, _ = oracle.zsort(q)
    ins = jz.tree.search_sorted_z(_t(pz, dev), _t(qz, dev)).cpu().numpy()
    merged = np.insert(pz, ins, qz, axis=0)
    mz, _ = oracle.zsort(merged)
    assert np.array_equal(merged, mz), "insertion at the returned indices breaks the z-order"
