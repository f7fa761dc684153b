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


def test_small_sets_at_both_ends(dev):
    """Sets at the two ends of the batch smaller than a top-plane node (whose size follows the
    total count) still get their own nodes on every plane: regression test."""
    import jztree_amd as jz
    cfg = jz.config.KNNConfig()
    sizes = np.array([1100] + [1 << 19] * 7 + [1500])
    top = jz.tree._define_tree_level_node_sizes(int(sizes.sum()), cfg.tree)[-1]
    assert sizes[0] < top and sizes[-1] < top
    parts = _batch(sizes, 70)
    print(f"[ends] sizes={sizes.tolist()} top node size {top}", flush=True)
    x = torch.from_numpy(np.concatenate(parts)).to(dev)
    rl = [0.2 * float(s) ** (-1.0 / 3.0) for s in sizes]
    rnn, inn, lab = jz.batch.pipeline_sets(x, sizes, 8, rl, boxsize=1.0)
    posz, idz, seg, off, th = jz.batch.tree_sets(x, sizes, cfg.tree)
    torch.cuda.synchronize()
    ls = th.leaf_set.cpu().numpy()
    for p in range(th.num_planes()):
        m = int(th.num(p))
        n2l = th.ispl_n2l.get(p, th.size() + 1).cpu().numpy()
        assert n2l[0] == 0 and n2l[m] == int(th.num(0)), f"plane {p} does not cover every leaf"
        assert len(np.unique(ls[n2l[:m]])) == len(sizes), f"plane {p} lost a set"
    off = np.concatenate([[0], np.cumsum(sizes)])
    for s in (0, len(sizes) - 1, 3):
        a, b = off[s], off[s + 1]
        xs = torch.from_numpy(parts[s]).to(dev)
        r1, i1 = jz.knn.knn(xs, k=8, boxsize=1.0)
        _, l1 = jz.fof.fof_labels(xs, rl[s], boxsize=1.0, output_order="input")
        assert torch.equal(inn[a:b], i1) and torch.equal(rnn[a:b], r1), f"set {s}: kNN differs"
        assert torch.equal(lab[a:b], l1), f"set {s}: labels differ"


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


@pytest.mark.parametrize("box,per_set", [(1.0, False), (0.0, False), (1.0, True), (0.0, True)])
def test_fof_sets_equals_single(dev, box, per_set):
    """One linking length for the batch, or b = 0.2 of each set's own mean spacing."""
    import jztree_amd as jz
    sizes = _sizes(24, 10, 14, 6)
    parts = _batch(sizes, 60)
    if per_set:
        rlink = [0.2 * float(s) ** (-1.0 / 3.0) for s in sizes]
    else:
        rlink = 0.2 * float(np.mean(sizes)) ** (-1.0 / 3.0)
    print(f"[fof_sets] box={box} per_set={per_set} nsets={len(sizes)} n={int(sizes.sum())}", flush=True)
    x = torch.from_numpy(np.concatenate(parts)).to(dev)
    lab = jz.batch.fof_sets(x, sizes, rlink, boxsize=box)
    torch.cuda.synchronize()
    off = np.concatenate([[0], np.cumsum(sizes)])
    ngroups = 0
    for s, p in enumerate(parts):
        rl = rlink[s] if per_set else rlink
        _, l1 = jz.fof.fof_labels(torch.from_numpy(p).to(dev), rl, boxsize=box, output_order="input")
        a, b = off[s], off[s + 1]
        assert torch.equal(lab[a:b], l1), f"set {s}: labels differ"
        ngroups += int(torch.unique(l1).numel())
    assert ngroups < int(sizes.sum())  # some linking happened


@pytest.mark.parametrize("big", [False, True])
def test_pipeline_and_captured_graph(dev, big):
    """One forest for kNN and FoF equals knn_sets/fof_sets; the HIP graph of the whole pipeline
    records no memset node and every replay (new positions, eager work in between) equals the
    eager result.  The big case has sets up to 2^19 points (multi-block paths everywhere)."""
    import jztree_amd as jz
    sizes = _sizes(8, 16, 19.5, 9) if big else _sizes(30, 10, 13, 8)
    k = 8
    rl = [0.2 * float(s) ** (-1.0 / 3.0) for s in sizes]
    xs = [torch.from_numpy(np.concatenate(_batch(sizes, 100 * t))).to(dev) for t in range(3)]
    print(f"[pipeline] nsets={len(sizes)} n={int(sizes.sum())}", flush=True)
    rnn, inn, lab = jz.batch.pipeline_sets(xs[0], sizes, k, rl, boxsize=1.0)
    r1, i1 = jz.batch.knn_sets(xs[0], sizes, k, boxsize=1.0)
    l1 = jz.batch.fof_sets(xs[0], sizes, rl, boxsize=1.0)
    assert torch.equal(inn, i1) and torch.equal(rnn, r1) and torch.equal(lab, l1)
    print("[pipeline] capture", flush=True)
    g = jz.batch.CapturedSets(xs[0], sizes, k, rl, boxsize=1.0)
    print(f"[pipeline] graph nodes {g.node_counts}", flush=True)
    assert g.memset_nodes() == 0
    for rep in range(4):
        x = xs[rep % 3]
        out = g.run(x)
        torch.cuda.synchronize()
        g.check()
        got = [t.clone() for t in out]
        ref = jz.batch.pipeline_sets(x, sizes, k, rl, boxsize=1.0)
        torch.cuda.synchronize()
        print(f"[pipeline] replay {rep}", flush=True)
        for a, b in zip(got, ref):
            assert torch.equal(a, b), f"replay {rep} differs from the eager pipeline"
