"""Separate query sets through one tree over two particle types, and SearchSortedZ.

- zsort_and_tree_multi_type: the reference's invariants (checks/tests/test_tree.py:52-76: per-type
  splits monotone with the right totals, every particle inside its leaf's cell) and the leaf splits
  bit-exact against the oracle's FlagLeafBoundaries with two types;
- knn(part, k, part_query=q): ids and radii bit-exact against the oracle (the reference's
  checks/tests/test_knn.py:96-100 case, clustered/periodic, k > 64, z output order);
- search_sorted_z: checks/tests/test_tree.py:110-133 (insertion keeps z-order, leaf search gives
  the leaf splits).
"""
import numpy as np
import pytest
import torch

from gen import clustered, uniform

pytestmark = pytest.mark.gpu


def test_multi_type_tree(dev, oracle):
    import jztree_amd as jz
    n1, n2 = 100_000, 200_000
    p1 = torch.from_numpy(uniform(n1, seed=41)).to(dev)
    p2 = torch.from_numpy(uniform(n2, seed=42)).to(dev)
    (p1z, p2z), th = jz.tree.zsort_and_tree_multi_type((p1, p2), cfg_tree=jz.config.KNNConfig().tree)
    torch.cuda.synchronize()
    ispl_t = th.ispl_l2p_per_type
    assert bool(torch.all(ispl_t[0, 1:] >= ispl_t[0, :-1])) and bool(torch.all(ispl_t[1, 1:] >= ispl_t[1, :-1]))
    assert int(ispl_t[0, -1]) == n1 and int(ispl_t[1, -1]) == n2
    nleaves = int(th.num(0))
    cent = th.geom_cent.get(0, nleaves).double()
    lvl = th.lvl.get(0, nleaves).cpu().numpy()
    ext = torch.from_numpy(np.stack([oracle_ext(l) for l in lvl])).to(dev)
    for pz, t, n in ((p1z, 0, n1), (p2z, 1, n2)):
        leaf = jz.tools.inverse_of_splits(ispl_t[t][:nleaves + 1].contiguous(), n).long()
        assert bool(torch.all(pz.double() >= cent[leaf] - 0.5 * ext[leaf]))
        assert bool(torch.all(pz.double() <= cent[leaf] + 0.5 * ext[leaf]))
    # without regularization the leaf splits equal the oracle's two-type FlagLeafBoundaries on the
    # stacked z-order (per-type counts at every split).  (Plane 0 may then group several leaves:
    # its nodes split where the largest per-type count of a cell exceeds the leaf size, while the
    # leaf flags attribute each scanned boundary to the type of the particle left of it, as
    # tree.cuh:67-86 does; with regularization plane 0 keeps every leaf, tree.py:676.)
    import dataclasses
    cfg0 = dataclasses.replace(jz.config.KNNConfig().tree, regularization=None)
    _, th0 = jz.tree.zsort_and_tree_multi_type((p1, p2), cfg_tree=cfg0)
    xs = np.concatenate([p1.cpu().numpy(), p2.cpu().numpy()])
    pz, idz = oracle.zsort(xs)
    ptype = (idz >= n1).astype(np.uint8)
    flags, _ = oracle.flag_leaf_boundaries(pz, leaf_size=cfg0.max_leaf_size, ptype=ptype, num_types=2)
    spl_stacked = np.nonzero(flags)[0]
    for t in range(2):
        pre = np.concatenate([[0], np.cumsum(ptype == t)])
        got = th0.ispl_l2p_per_type[t].cpu().numpy()
        assert np.array_equal(got[:len(spl_stacked)], pre[spl_stacked]), f"type {t}"
        assert np.all(got[len(spl_stacked):] == pre[-1])


def oracle_ext(level, dim=3):
    """Per-dim full extent of a node of binary level `level` (math.cuh:583-594)."""
    o = (level + 2000 * dim) // dim - 2000
    r = level - o * dim
    return np.array([2.0 ** (o + (1 if r >= dim - i else 0)) for i in range(dim)])


@pytest.mark.parametrize("case", ["reference", "clustered_periodic", "k80", "z_order"])
def test_knn_separate_queries(dev, oracle, case):
    import jztree_amd as jz
    box = 0.0
    k = 16
    if case == "reference":  # checks/tests/test_knn.py:96-100
        x, xq = uniform(100_000, seed=43, lo=-1.0, hi=1.0), uniform(130_000, seed=44, lo=-0.5, hi=0.5)
    elif case == "clustered_periodic":
        x, xq, box, k = clustered(1 << 17, nhalos=128, seed=45), clustered(50_000, nhalos=64, seed=46), 1.0, 32
    elif case == "k80":
        x, xq, k = uniform(40_000, seed=47), uniform(20_000, seed=48), 80
    else:
        x, xq = uniform(60_000, seed=49), uniform(30_000, seed=50)
    print(f"[query] {case} n={len(x)} nq={len(xq)} k={k}", flush=True)
    t, tq = torch.from_numpy(x).to(dev), torch.from_numpy(xq).to(dev)
    order = "z" if case == "z_order" else "input"
    rnn, inn = jz.knn.knn(t, k, boxsize=box or None, part_query=tq, output_order=order)
    torch.cuda.synchronize()
    if order == "z":
        xz, _ = oracle.zsort(x)
        xqz, _ = oracle.zsort(xq)
        r2o, io = oracle.knn(xz, k, xq=xqz, boxsize=box)
    else:
        r2o, io = oracle.knn(x, k, xq=xq, boxsize=box)
    got = inn.cpu().numpy().astype(np.int64)
    assert np.array_equal(got, io), f"{int(np.sum(got != io))} ids differ"
    assert np.array_equal(rnn.cpu().numpy(), np.sqrt(r2o).astype(np.float32))


def test_search_sorted_z(dev, oracle):
    import jztree_amd as jz
    a = torch.from_numpy(uniform(1387, seed=0, lo=0.1, hi=0.4)).to(dev)
    b = torch.from_numpy(uniform(2222, seed=2, lo=0.1, hi=0.4)).to(dev)
    posz, _ = jz.tree.zsort(a)
    posz2, _ = jz.tree.zsort(b)
    iself = jz.tree.search_sorted_z(posz, posz)
    assert torch.equal(iself, torch.arange(len(posz), dtype=torch.int32, device=dev))
    i2 = jz.tree.search_sorted_z(posz, posz2).cpu().numpy()
    ins = np.insert(posz.cpu().numpy(), i2, posz2.cpu().numpy(), axis=0)
    ref, _ = oracle.zsort(ins)
    assert np.array_equal(ins, ref), "insertion at the returned indices breaks the z-order"


def test_leaf_search(dev):
    """The leaf of every particle, found from the leaf centres alone, gives the leaf splits."""
    import jztree_amd as jz
    x = torch.from_numpy(clustered(1 << 16, nhalos=64, seed=51)).to(dev)
    cfg = jz.config.KNNConfig().tree
    import dataclasses
    cfg = dataclasses.replace(cfg, alloc_fac_nodes=2.0, regularization=None)
    posz, th = jz.tree.zsort_and_tree(x, cfg)
    nleaves = int(th.num(0))
    xleaf = th.geom_cent.get(0, nleaves).contiguous()
    spl = th.splits_leaf_to_part(size=nleaves + 1)
    ileaf = jz.tree.search_sorted_z(xleaf, posz, leaf_search=True)
    spl2 = torch.searchsorted(ileaf.contiguous(), torch.arange(nleaves + 1, dtype=torch.int32, device=dev),
                              side="left")
    assert torch.equal(spl.long(), spl2.long()), "leaf ranges differ"
