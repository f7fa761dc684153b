"""Full configuration sizes, bit-exact against the oracle on the host cores.

- 2^24 clustered points (4096 Plummer halos, 70% in halos), periodic unit box:
  * kNN k=32: ids and radii of 8192 sampled queries equal the oracle's exact k-d tree result
    (neighbours ascending by (canonical FP32 r2, input index); radius = sqrt of that r2);
  * FoF b=0.2: the groups of all 2^24 points equal the oracle's (cell FoF on the same FP32
    predicate), compared as labels canonicalised to the minimum input index of each group.
- 256 uniform sets with sizes floor(2^U(10,20)) (seed 7, about 3.8e7 points), kNN k=8 + FoF b=0.2
  of each set's own mean spacing: the one-launch-per-stage batch pipeline equals the same pipeline
  run one set at a time and every replay of its captured HIP graph; three sets are checked
  against the oracle directly.
- 2^26 clustered points (65536 halos), kNN k=64 with the learned radius proposer trained on the
  plain kNN of 4 other seeds: bit-identical to the plain kNN; fallback rate and log2-radius RMSE
  are printed, and the first 4096 queries are checked against the oracle.
"""
import dataclasses
import multiprocessing as mp
import os
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np
import pytest
import torch

from gen import clustered

pytestmark = pytest.mark.gpu

N = 1 << 24
BOX = 1.0


@pytest.fixture(scope="module")
def points(dev):
    x = clustered(N, nhalos=4096, seed=5)
    return x, torch.from_numpy(x).to(dev)


def _canon_min_index(lab: torch.Tensor) -> torch.Tensor:
    """Group labels -> the minimum index of each group (same partition, canonical names)."""
    n = lab.numel()
    lab = lab.long()
    mins = torch.full((n,), n, dtype=torch.int64, device=lab.device)
    mins.scatter_reduce_(0, lab, torch.arange(n, device=lab.device), "amin", include_self=True)
    return mins[lab]


def test_knn_2e24_bit_exact(dev, oracle, points):
    import jztree_amd as jz
    x, t = points
    k = 32
    cfg = jz.config.KNNConfig()
    cfg = dataclasses.replace(cfg, tree=dataclasses.replace(cfg.tree, alloc_fac_nodes=2.0, max_leaf_size=128))
    print(f"[full] knn n={N} k={k}", flush=True)
    rnn, inn = jz.knn.knn(t, k=k, boxsize=BOX, cfg=cfg)
    torch.cuda.synchronize()
    q = np.sort(np.random.default_rng(0).choice(N, 8192, replace=False))
    t0 = time.time()
    r2o, io = oracle.knn(x, k, xq=x[q], boxsize=BOX)
    print(f"[full] oracle knn of {len(q)} queries: {time.time() - t0:.1f}s", flush=True)
    qi = torch.from_numpy(q).to(dev)
    got_i = inn[qi].cpu().numpy().astype(np.int64)
    got_r = rnn[qi].cpu().numpy()
    assert np.array_equal(got_i, io), f"{int(np.sum(got_i != io))} ids differ from the oracle"
    assert np.array_equal(got_r, np.sqrt(r2o).astype(np.float32)), "radii differ from the oracle"


def test_fof_2e24_bit_exact(dev, oracle, points):
    import jztree_amd as jz
    x, t = points
    rl = 0.2 * N ** (-1.0 / 3.0)
    cfg = jz.config.FofConfig()
    cfg = dataclasses.replace(cfg, tree=dataclasses.replace(cfg.tree, alloc_fac_nodes=2.0))
    print(f"[full] fof n={N} rlink={rl}", flush=True)
    _, lab = jz.fof.fof_labels(t, rl, boxsize=BOX, cfg=cfg, output_order="input")
    torch.cuda.synchronize()
    lab = lab.long()
    assert torch.all(lab[lab] == lab), "a label does not name a member of its own group"
    got = _canon_min_index(lab).cpu().numpy()
    t0 = time.time()
    ref = oracle.fof(x, rl, boxsize=BOX, key=np.arange(N), method="cells")
    print(f"[full] oracle fof: {time.time() - t0:.1f}s, {len(np.unique(ref))} groups", flush=True)
    assert np.array_equal(got, ref), f"{int(np.sum(got != ref))} particles are grouped differently"
    # the minimum group size 20 of the configuration is a catalogue cut: count the groups it keeps
    sizes = np.bincount(ref)
    print(f"[full] groups with >= 20 members: {int(np.sum(sizes >= 20))}", flush=True)


def _batch_256(seed=7, nsets=256):
    rng = np.random.default_rng(seed)
    sizes = np.floor(2.0 ** rng.uniform(10, 20, nsets)).astype(np.int64)
    x = rng.random((int(sizes.sum()), 3), dtype=np.float32)
    return sizes, x


def test_batch_256_sets(dev, oracle):
    import jztree_amd as jz
    sizes, xh = _batch_256()
    n = int(sizes.sum())
    k = 8
    rl = [0.2 * float(s) ** (-1.0 / 3.0) for s in sizes]
    off = np.concatenate([[0], np.cumsum(sizes)])
    x = torch.from_numpy(xh).to(dev)
    print(f"[c3] {len(sizes)} sets, {n} points", flush=True)
    ref = [a.clone() for a in jz.batch.pipeline_sets(x, sizes, k, rl, boxsize=BOX)]
    torch.cuda.synchronize()
    print("[c3] one set at a time", flush=True)
    for s in range(len(sizes)):
        a, b = int(off[s]), int(off[s + 1])
        one = jz.batch.pipeline_sets(x[a:b], sizes[s:s + 1], k, rl[s:s + 1], boxsize=BOX)
        for i, name in enumerate(("radii", "ids", "labels")):
            assert torch.equal(one[i], ref[i][a:b]), f"set {s}: {name} differ from the batch"
    print("[c3] oracle on three sets", flush=True)
    for s in (int(np.argmin(sizes)), 17, int(np.argmax(sizes))):
        a, b = int(off[s]), int(off[s + 1])
        r2o, io = oracle.knn(xh[a:b], k, boxsize=BOX)
        assert np.array_equal(ref[1][a:b].cpu().numpy().astype(np.int64), io), f"set {s}: kNN ids"
        assert np.array_equal(ref[0][a:b].cpu().numpy(), np.sqrt(r2o).astype(np.float32)), f"set {s}: radii"
        lab = ref[2][a:b].long()  # set-local indices
        fo = oracle.fof(xh[a:b], rl[s], boxsize=BOX, key=np.arange(b - a), method="cells")
        assert np.array_equal(_canon_min_index(lab).cpu().numpy(), fo), f"set {s}: FoF groups"
    print("[c3] captured graph", flush=True)
    g = jz.batch.CapturedSets(x, sizes, k, rl, boxsize=BOX)
    assert g.memset_nodes() == 0
    for rep in range(3):
        out = g.run()
        torch.cuda.synchronize()
        g.check()
        for i in range(3):
            assert torch.equal(out[i], ref[i]), f"replay {rep} differs from the eager batch"
        _ = torch.cumsum(x[:, 0], 0)  # eager work between replays
        torch.cuda.synchronize()


def _gen_clustered(args):
    n, halos, seed = args
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from gen import clustered as cl
    return cl(n, nhalos=halos, seed=seed)


def test_learned_k64_2e26_identical(dev, oracle):
    """2^26 clustered points, 65536 halos, k=64: the learned search equals the plain search bit
    for bit; the proposer is trained on the plain kNN radii of 4 other seeds."""
    import jztree_amd as jz
    n, k, halos = 1 << 26, 64, 65536
    t0 = time.time()
    # spawned workers: this process has initialised the GPU, forked children must not inherit it
    with ProcessPoolExecutor(max_workers=5, mp_context=mp.get_context("spawn")) as ex:
        sets = list(ex.map(_gen_clustered, [(n, halos, s) for s in (0, 2001, 2002, 2003, 2004)]))
    print(f"[c4] generated 5 sets of {n} points in {time.time() - t0:.1f}s", flush=True)
    cfg = jz.config.KNNConfig()
    cfg = dataclasses.replace(cfg, tree=dataclasses.replace(cfg.tree, alloc_fac_nodes=2.0, max_leaf_size=128))
    feats, targets = [], []
    for xs in sets[1:]:
        f, tt = jz.proposer.training_data([torch.from_numpy(xs).to(dev)], k, cfg, boxsize=BOX)
        feats.append(f)
        targets.append(tt)
    feats, targets = torch.cat(feats), torch.cat(targets)
    prop = jz.proposer.RadiusProposer(k, device=dev, seed=0)
    rep = prop.train(feats, targets, steps=2000, batch=8192, lr=3e-3, lr_final=1e-4)
    prop.calibrate(feats, targets, quantile=0.999)
    print(f"[c4] trained on {len(targets)} leaves: loss {rep.loss_first:.4f} -> {rep.loss_last:.4f}, "
          f"margin {prop.margin:.3f}", flush=True)
    assert rep.loss_last < rep.loss_first
    del feats, targets
    x = torch.from_numpy(sets[0]).to(dev)
    xh = sets[0]
    del sets
    r0, i0 = jz.knn.knn(x, k=k, boxsize=BOX, cfg=cfg)
    r1, i1, retry = jz.proposer.knn_learned(x, k, prop, boxsize=BOX, cfg=cfg, return_retry=True)
    torch.cuda.synchronize()
    assert torch.equal(i0, i1) and torch.equal(r0, r1), "learned kNN differs from the plain kNN"
    fallback = float(retry.float().mean())
    partz, originz, th = jz.tree.zsort_and_tree(x, cfg.tree, data=jz.data.RankIdx(
        rank=None, idx=torch.arange(n, dtype=torch.int32, device=dev)))
    nl = int(th.num(0))
    target = jz.proposer.leaf_log2_radius(th, r0[originz.idx.long()], k)[:nl]
    pred = prop.predict_log2r(jz.proposer.leaf_features(th)[:nl])
    rmse = float(torch.sqrt(torch.mean((pred.double() - target.double()) ** 2)))
    print(f"[c4] identical; fallback rate {fallback:.2e}, log2-radius RMSE {rmse:.3f}", flush=True)
    assert rmse < 1.0
    q = np.arange(4096)
    r2o, io = oracle.knn(xh, k, xq=xh[q], boxsize=BOX)
    assert np.array_equal(i1[:4096].cpu().numpy().astype(np.int64), io), "ids differ from the oracle"
    assert np.array_equal(r1[:4096].cpu().numpy(), np.sqrt(r2o).astype(np.float32)), "radii differ"
