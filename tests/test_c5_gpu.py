"""Float64 positions for the keys and the tree, float32 positions for the distance kernels
(dist_dtype=torch.float32): results are the exact kNN / FoF of the float32-rounded points.

- Points whose differences lie below float32 resolution (tiny clumps that round onto few float32
  values): kNN and FoF bit-exact against the oracle on the rounded points.
- 2^22 points of the 60%-halo mix: kNN ids and radii of sampled queries and all FoF labels
  bit-exact against the oracle on the rounded points; the learned search equals the plain one.
- 2^28 points of the same mix (generated on the GPU): learned kNN k=32 equals the plain kNN,
  sampled queries equal the oracle's cell-list kNN on the rounded points, and every FoF group that the
  oracle finds well inside a sub-box equals the GPU group of its members.
"""
import dataclasses
import time

import numpy as np
import pytest
import torch

from gen import clustered, clustered_torch

pytestmark = pytest.mark.gpu


def _cfgs(leaf=128, ilist=256.0):
    import jztree_amd as jz
    k = jz.config.KNNConfig()
    k = dataclasses.replace(k, alloc_fac_ilist=ilist,
                            tree=dataclasses.replace(k.tree, alloc_fac_nodes=2.0, max_leaf_size=leaf))
    f = jz.config.FofConfig()
    f = dataclasses.replace(f, tree=dataclasses.replace(f.tree, alloc_fac_nodes=2.0))
    return k, f


def _canon(lab):
    """labels -> minimum index of each group (numpy)"""
    lab = np.asarray(lab, np.int64)
    mins = np.full(len(lab), len(lab), np.int64)
    np.minimum.at(mins, lab, np.arange(len(lab)))
    return mins[lab]


def _check_small(dev, oracle, x64, k, box, rl):
    import jztree_amd as jz
    kcfg, fcfg = _cfgs(leaf=48)
    t = torch.from_numpy(x64).to(dev)
    x32 = x64.astype(np.float32)
    rnn, inn = jz.knn.knn(t, k, boxsize=box or None, cfg=kcfg, dist_dtype=torch.float32)
    _, lab = jz.fof.fof_labels(t, rl, boxsize=box, cfg=fcfg, output_order="input", dist_dtype=torch.float32)
    torch.cuda.synchronize()
    assert rnn.dtype == torch.float32
    r2o, io = oracle.knn(x32, k, boxsize=box)
    got_i = inn.cpu().numpy().astype(np.int64)
    assert np.array_equal(got_i, io), f"{int(np.sum(got_i != io))} ids differ"
    assert np.array_equal(rnn.cpu().numpy(), np.sqrt(r2o).astype(np.float32))
    ref = oracle.fof(x32, rl, boxsize=box, key=np.arange(len(x32)), method="cells")
    assert np.array_equal(_canon(lab.cpu().numpy()), ref)


@pytest.mark.parametrize("box", [0.0, 1.0])
def test_sub_float32_clumps(dev, oracle, box):
    """Clumps of points a few float64 ulps to 1e-9 apart: after rounding they share float32 values
    (exact duplicates), their float64 cells are far below float32 resolution."""
    rng = np.random.default_rng(21)
    base = rng.random((3000, 3))
    spread = 10.0 ** rng.uniform(-15, -8, 3000)
    x = np.repeat(base, 8, axis=0) + rng.normal(size=(24000, 3)) * np.repeat(spread, 8)[:, None]
    x = np.concatenate([x, rng.random((20000, 3))]) % 1.0
    x = x[rng.permutation(len(x))]
    print(f"[c5] clumps n={len(x)} box={box}", flush=True)
    _check_small(dev, oracle, x, 16, box, 0.7 * len(x) ** (-1 / 3))


def test_mixed_precision_2e22(dev, oracle):
    import jztree_amd as jz
    n, k = 1 << 22, 32
    x64 = clustered(n, nhalos=4096, halo_frac=0.6, mmax=2 ** 18, seed=31).astype(np.float64)
    # float64 coordinates that are not float32 numbers
    x64 = (x64 + np.random.default_rng(31).uniform(-2e-9, 2e-9, x64.shape)) % 1.0
    x32 = x64.astype(np.float32)
    kcfg, fcfg = _cfgs()
    t = torch.from_numpy(x64).to(dev)
    print(f"[c5] n={n} mixed precision", flush=True)
    rnn, inn = jz.knn.knn(t, k, boxsize=1.0, cfg=kcfg, dist_dtype=torch.float32)
    # a proposer trained on one float32 set of the same mix
    xs = torch.from_numpy(clustered(n, nhalos=4096, halo_frac=0.6, mmax=2 ** 18, seed=32)).to(dev)
    f, tt = jz.proposer.training_data([xs], k, kcfg, boxsize=1.0)
    prop = jz.proposer.RadiusProposer(k, device=dev, seed=0)
    prop.train(f, tt, steps=300, batch=8192, lr=3e-3, lr_final=3e-4)
    prop.calibrate(f, tt, quantile=0.99)
    r1, i1 = jz.proposer.knn_learned(t, k, prop, boxsize=1.0, cfg=kcfg, dist_dtype=torch.float32)
    torch.cuda.synchronize()
    assert torch.equal(inn, i1) and torch.equal(rnn, r1), "learned differs from plain"
    q = np.sort(np.random.default_rng(3).choice(n, 4096, replace=False))
    r2o, io = oracle.knn(x32, k, xq=x32[q], boxsize=1.0)
    qi = torch.from_numpy(q).to(dev)
    assert np.array_equal(inn[qi].cpu().numpy().astype(np.int64), io)
    assert np.array_equal(rnn[qi].cpu().numpy(), np.sqrt(r2o).astype(np.float32))
    rl = 0.2 * n ** (-1 / 3)
    _, lab = jz.fof.fof_labels(t, rl, boxsize=1.0, cfg=fcfg, output_order="input", dist_dtype=torch.float32)
    ref = oracle.fof(x32, rl, boxsize=1.0, key=np.arange(n), method="cells")
    assert np.array_equal(_canon(lab.cpu().numpy()), ref)


def test_mixed_precision_2e28(dev, oracle):
    import jztree_amd as jz
    n, k = 1 << 28, 32
    nh = int(0.6 * n / 421)
    t0 = time.time()
    x = clustered_torch(n, nhalos=nh, halo_frac=0.6, mmax=2 ** 18, seed=5, device=dev)
    torch.cuda.synchronize()
    print(f"[c5] generated {n} points ({nh} halos) in {time.time() - t0:.1f}s", flush=True)
    kcfg, fcfg = _cfgs(leaf=128)  # 4.2M leaf slots x 256 list entries: int32 lists hold
    xs = clustered_torch(1 << 26, nhalos=nh // 4, halo_frac=0.6, mmax=2 ** 18, seed=6, device=dev, dtype=torch.float32)
    f, tt = jz.proposer.training_data([xs], k, kcfg, boxsize=1.0)
    del xs
    prop = jz.proposer.RadiusProposer(k, device=dev, seed=0)
    prop.train(f, tt, steps=500, batch=8192, lr=3e-3, lr_final=3e-4)
    prop.calibrate(f, tt, quantile=0.999)
    del f, tt
    q = torch.from_numpy(np.sort(np.random.default_rng(4).choice(n, 256, replace=False))).to(dev)

    def digest(r, i):
        """sampled rows and exact integer checksums of the full outputs (two k=32 outputs at 2^28
        would hold 137 GB at once)"""
        ci = cr = 0
        for s0 in range(0, n, 1 << 22):  # row slices: the int64 products of a full output are 68 GB
            w = (torch.arange(s0, min(n, s0 + (1 << 22)), device=dev, dtype=torch.int64) % 1009 + 1)[:, None]
            ci += int((i[s0:s0 + (1 << 22)].to(torch.int64) * w).sum())
            cr += int((r[s0:s0 + (1 << 22)].view(torch.int32).to(torch.int64) * w).sum())
        return i[q].cpu().numpy().astype(np.int64), r[q].cpu().numpy(), ci, cr

    t0 = time.time()
    r1, i1, retry = jz.proposer.knn_learned(x, k, prop, boxsize=1.0, cfg=kcfg, return_retry=True,
                                           dist_dtype=torch.float32)
    torch.cuda.synchronize()
    print(f"[c5] learned kNN {time.time() - t0:.1f}s, fallback {float(retry.float().mean()):.2e}", flush=True)
    got_i, got_r, ci1, cr1 = digest(r1, i1)
    del r1, i1, retry
    torch.cuda.empty_cache()
    r0, i0 = jz.knn.knn(x, k, boxsize=1.0, cfg=kcfg, dist_dtype=torch.float32)
    torch.cuda.synchronize()
    pi, pr, ci0, cr0 = digest(r0, i0)
    del r0, i0
    torch.cuda.empty_cache()
    assert np.array_equal(pi, got_i) and np.array_equal(pr, got_r) and (ci0, cr0) == (ci1, cr1), \
        "learned differs from plain"
    x32 = x.to(torch.float32).cpu().numpy()
    print("[c5] oracle grid kNN of the sampled queries", flush=True)
    t0 = time.time()
    r2o, io = oracle.knn(x32, k, xq=x32[q.cpu().numpy()], boxsize=1.0, method="grid")
    print(f"[c5] oracle grid kNN of {len(q)} queries: {time.time() - t0:.1f}s", flush=True)
    assert np.array_equal(got_i, io)
    assert np.array_equal(got_r, np.sqrt(r2o).astype(np.float32))
    # FoF: groups inside a sub-box, against the oracle on the points of the sub-box
    rl = 0.2 * n ** (-1 / 3)
    print("[c5] FoF", flush=True)
    _, lab = jz.fof.fof_labels(x, rl, boxsize=1.0, cfg=fcfg, output_order="input", dist_dtype=torch.float32)
    lab = lab.cpu().numpy().astype(np.int64)
    print("[c5] FoF done", flush=True)
    del x
    torch.cuda.empty_cache()
    lo, hi = 0.30, 0.42
    sel = np.nonzero(np.all((x32 >= lo) & (x32 < hi), axis=1))[0]
    xs = x32[sel]
    t0 = time.time()
    ref = oracle.fof(xs, rl, boxsize=0.0, key=np.arange(len(sel)), method="cells")
    print(f"[c5] oracle FoF of {len(sel)} sub-box points: {time.time() - t0:.1f}s", flush=True)
    # oracle groups all of whose members are farther than 2 rlink from the sub-box faces
    inner = np.all((xs >= lo + 2 * rl) & (xs < hi - 2 * rl), axis=1)
    bad_group = np.zeros(len(sel), bool)
    np.logical_or.at(bad_group, ref, ~inner)
    keep = ~bad_group[ref]
    assert keep.sum() > 0.5 * len(sel)
    g_gpu = lab[sel[keep]]
    g_ref = ref[keep]
    # same partition on the kept points: the maps gpu label <-> oracle label are functions
    order = np.lexsort((g_gpu, g_ref))
    a, b = g_ref[order], g_gpu[order]
    assert np.all((a[1:] != a[:-1]) | (b[1:] == b[:-1])), "an oracle group is split on the GPU"
    order = np.lexsort((g_ref, g_gpu))
    a, b = g_gpu[order], g_ref[order]
    assert np.all((a[1:] != a[:-1]) | (b[1:] == b[:-1])), "GPU group merges oracle groups"
    # and a GPU group of kept points has no member outside the kept set
    sizes_gpu = np.bincount(lab, minlength=n)
    kept_sizes = np.bincount(g_gpu, minlength=n)
    assert np.array_equal(sizes_gpu[np.unique(g_gpu)], kept_sizes[np.unique(g_gpu)])
