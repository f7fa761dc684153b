"""Batches of independent sets: every set's result equals a single-set run on that set, bit for bit
(sort order and ids, leaves inside sets, kNN radii and ids, FoF labels)."""
import numpy as np
import pytest
import torch

from gen import clustered, uniform

pytestmark = pytest.mark.gpu


def _sizes(nsets, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return np.floor(2.0 ** rng.uniform(lo, hi, nsets)).astype(np.int64)


def _batch(sizes, seed, kind="uniform", box=1.0):
    parts = []
    for i, s in enumerate(sizes):
        if kind == "mixed" and i % 3 == 1:
            x = uniform(int(s), seed=seed + i, lo=-0.5, hi=0.25) if s else np.zeros((0, 3), np.float32)
        elif kind == "mixed" and i % 3 == 2 and s >= 64:
            x = clustered(int(s), nhalos=max(int(s) // 200, 1), seed=seed + i)
        else:
            x = uniform(int(s), seed=seed + i, hi=box) if s else np.zeros((0, 3), np.float32)
        if box and kind == "uniform":
            x[x >= np.float32(box)] = 0.0
        parts.append(np.ascontiguousarray(x, dtype=np.float32))
    return parts


@pytest.mark.parametrize("case", ["few", "many_sets", "mixed_domains"])
def test_zsort_sets_equals_single(dev, case):
    import jztree_amd as jz
    if case == "few":
        sizes = np.array([1500, 0, 4096, 7, 1, 2500])
        parts = _batch(sizes, 10)
    elif case == "many_sets":  # > 256 sets: two stable passes on the set id
        sizes = _sizes(300, 3, 9, 11)
        parts = _batch(sizes, 20)
    else:
        sizes = _sizes(12, 10, 13, 12)
        parts = _batch(sizes, 30, kind="mixed")
        parts[3][:100] = parts[3][100:200]  # duplicated points: equal-key runs inside a set
    print(f"[zsort_sets] case={case} nsets={len(sizes)} n={int(sizes.sum())}", flush=True)
    x = torch.from_numpy(np.concatenate(parts)).to(dev)
    off = jz.batch.set_offsets(sizes, dev)
    posz, idz = jz.batch.zsort_sets(x, off)
    torch.cuda.synchronize()
    offh = off.cpu().numpy()
    for s, p in enumerate(parts):
        if len(p) == 0:
            continue
        pz1, id1 = jz.tree.zsort(torch.from_numpy(p).to(dev))
        a, b = offh[s], offh[s + 1]
        assert torch.equal(idz[a:b].long() - a, id1.long()), f"set {s}: order differs"
        assert torch.equal(posz[a:b], pz1), f"set {s}: positions differ"


def test_forest_never_spans_sets(dev):
    import jztree_amd as jz
    sizes = _sizes(40, 6, 12, 3)
    parts = _batch(sizes, 40)
    x = torch.from_numpy(np.concatenate(parts)).to(dev)
    cfg = jz.config.KNNConfig()
    print(f"[forest] nsets={len(sizes)} n={int(sizes.sum())}", flush=True)
    posz, idz, seg, off, th = jz.batch.tree_sets(x, sizes, cfg.tree)
    torch.cuda.synchronize()
    segh = seg.cpu().numpy()
    l2p = th.splits_leaf_to_part().cpu().numpy()
    nl = int(th.num(0))
    st, en = l2p[:nl], l2p[1:nl + 1]
    assert np.all(en > st)
    assert np.all(segh[st] == segh[en - 1]), "a leaf spans two sets"
    # the set boundaries are leaf boundaries
    assert set(off.cpu().numpy()[1:-1].tolist()) <= set(l2p[:nl + 1].tolist())
    for p in range(1, th.num_planes()):
        n2l = th.ispl_n2l.get(p, th.size() + 1).cpu().numpy()
        m = int(th.num(p))
        a, b = l2p[n2l[:m]], l2p[n2l[1:m + 1]]
        assert np.all(segh[a] == segh[b - 1]), f"a plane-{p} node spans two sets"


@pytest.mark.parametrize("k,box,kind", [(8, 1.0, "uniform"), (16, 0.0, "uniform"), (8, 0.0, "mixed")])
def test_knn_sets_equals_single(dev, k, box, kind):
    import jztree_amd as jz
    sizes = _sizes(24, 10, 14, 5)
    parts = _batch(sizes, 50, kind=kind, box=box if box else 1.0)
    print(f"[knn_sets] k={k} box={box} kind={kind} nsets={len(sizes)} n={int(sizes.sum())}", flush=True)
    x = torch.from_numpy(np.concatenate(parts)).to(dev)
    cfg = jz.config.KNNConfig()
    if kind == "mixed":
        import dataclasses
        cfg = dataclasses.replace(cfg, tree=dataclasses.replace(cfg.tree, alloc_fac_nodes=2.0))
    rnn, inn = jz.batch.knn_sets(x, sizes, k, boxsize=box, cfg=cfg)
    torch.cuda.synchronize()
    off = np.concatenate([[0], np.cumsum(sizes)])
    for s, p in enumerate(parts):
        r1, i1 = jz.knn.knn(torch.from_numpy(p).to(dev), k=k, boxsize=box, cfg=cfg)
        a, b = off[s], off[s + 1]
        assert torch.equal(inn[a:b], i1), f"set {s}: ids differ"
        assert torch.equal(rnn[a:b], r1), f"set {s}: radii differ"


@pytest.mark.parametrize("box", [1.0, 0.0])
def test_fof_sets_equals_single(dev, box):
    import jztree_amd as jz
    sizes = _sizes(24, 10, 14, 6)
    parts = _batch(sizes, 60)
    rlink = 0.2 * float(np.mean(sizes)) ** (-1.0 / 3.0)
    print(f"[fof_sets] box={box} nsets={len(sizes)} n={int(sizes.sum())} rlink={rlink}", flush=True)
    x = torch.from_numpy(np.concatenate(parts)).to(dev)
    lab = jz.batch.fof_sets(x, sizes, rlink, boxsize=box)
    torch.cuda.synchronize()
    off = np.concatenate([[0], np.cumsum(sizes)])
    ngroups = 0
    for s, p in enumerate(parts):
        _, l1 = jz.fof.fof_labels(torch.from_numpy(p).to(dev), rlink, boxsize=box, output_order="input")
        a, b = off[s], off[s + 1]
        assert torch.equal(lab[a:b], l1), f"set {s}: labels differ"
        ngroups += int(torch.unique(l1).numel())
    assert ngroups < int(sizes.sum())  # some linking happened
