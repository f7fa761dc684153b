"""Benchmark of the tree hot path on one MI355X.

Workload: 2^26 clustered float32 points in the periodic unit box (70% in 65,536 Plummer halos with
scale radius 0.002 and power-law member weights dN/dM ~ M^-1.9 on [32, 2^16], 30% uniform; seeded,
synthetic).  One step = exact kNN (k=32, periodic, neighbours in input order) with the learned
radius proposer, followed by FoF labels (linking length 0.2 x mean spacing, periodic), each from
positions resident in HBM: z-order sort, tree build, (features + MLP bound,) dual tree walk, leaf
kernels.  Every timed call is a replay of a HIP graph captured from the library's calls (no host
synchronisation inside; list-overflow checks are evaluated after the replays).  The plain exact kNN
(guaranteed radii, no proposer) is timed the same way on the same points; both return
bit-identical neighbours, which the run checks.

The proposers (k=32 and k=64) are trained first, on the GPU with the library's own kernels, on the
plain kNN radii of 4 other seeds of the same configuration (generated on the host in parallel);
their margins are calibrated on those seeds.

Also reported: the same points with k=64 (learned and plain), 2^24 clustered points with 4096
halos (FoF and plain kNN k=32), 2^28 points with float64 positions searched in float32 (learned and
plain kNN, FoF), and the batched workload (256 uniform sets, kNN k=8 + FoF).

Prints one JSON line.  value = learned kNN queries/s (k=32); roofline = the leaf kernel of the
learned search (the step's dominant kernel) timed with HIP events on the launch stream; fof_roofline
= FoF end to end and its particle-link stage; cpu_baseline = scipy cKDTree (the reference's own CPU
comparator) on a bounded sample of the same points, fof_cpu_baseline = the oracle's cell FoF.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import multiprocessing as mp
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
# HBM traffic of the leaf kernel per launch, from two rocprofv3 --pmc passes over this bench
# (FETCH_SIZE x 2.0 and WRITE_SIZE x 1.0, the factors a plain copy kernel of known size reads);
# keyed by (kernel, n, k, halos, leaf size, XCD order knob); None where not measured.
LEAF_TRAFFIC_BYTES = {("k_knn_leaf", 1 << 26, 32, 65536, 128, 1): 37_532_876_800}
# algorithmic bytes per point of the FoF stages (f32 positions, u32 labels): end to end
# (key generation, 8 radix passes, permutation, leaf build, particle links) and the link stage alone
FOF_BYTES_END_TO_END = 288
FOF_BYTES_LINK = 28


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=1 << 26)
    p.add_argument("--halos", type=int, default=65536)
    p.add_argument("--k", type=int, default=32)
    p.add_argument("--k2", type=int, default=64, help="second k on the same points, learned and plain (0 = skip)")
    p.add_argument("--b", type=float, default=0.2)
    p.add_argument("--train-seeds", type=int, default=4)
    p.add_argument("--train-steps", type=int, default=2000)
    # margin = this quantile of the training residuals; 0.999 measured fastest on the default
    # workload (0.99: 11k short queries, 0.999: 1.1k; the retry pass costs more than the
    # slightly larger leaf lists)
    p.add_argument("--quantile", type=float, default=0.999)
    p.add_argument("--n-mid", type=int, default=1 << 24, help="points of the 4096-halo set (0 = skip)")
    p.add_argument("--halos-mid", type=int, default=4096)
    p.add_argument("--c5-n", type=int, default=1 << 28,
                   help="points of the float64-key / float32-distance leg (0 = skip)")
    p.add_argument("--cpu-side", type=float, default=0.5,
                   help="CPU baselines run on the points in [0, side)^3 of the benchmark set")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--batch-sets", type=int, default=256, help="sets of the batched workload (0 = skip it)")
    p.add_argument("--batch-k", type=int, default=8)
    p.add_argument("--no-tune", action="store_true",
                   help="skip the autotuner (fixed leaf 128 for kNN and FoF, waves 4, tile 16, XCD order on)")
    p.add_argument("--copy-bytes", type=int, default=0,
                   help="only time a plain device copy of this many bytes (calibrates the HBM byte counters)")
    return p.parse_args()


def copy_probe(nbytes, steps):
    """A plain device-to-device copy of nbytes (torch's copy kernel), timed with HIP events: the
    reference point for what rocprofv3's FETCH_SIZE / WRITE_SIZE counters report per byte."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a)
    ms = Timer(torch)(lambda: b.copy_(a), steps, 2)
    print(json.dumps({"copy_bytes": a.numel() * 4, "ms": ms, "GB_per_s": 2 * a.numel() * 4 / (ms * 1e-3) / 1e9}),
          flush=True)


def _batch_workload(nsets, seed=7):
    """256 independent uniform sets in the periodic unit box, sizes floor(2^U(10,20)) (seeded)."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(2.0 ** rng.uniform(10, 20, nsets)).astype(np.int64)
    x = rng.random((int(sizes.sum()), 3), dtype=np.float32)
    return sizes, x


def _gen(args):
    from gen import clustered
    n, halos, seed = args
    return clustered(n, nhalos=halos, seed=seed)


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


class Timer:
    """Mean time of a callable over `steps` calls after `warmup` calls, from HIP events recorded on
    the current stream around each call (a graph replay is enqueued on that stream)."""

    def __init__(self, torch):
        self.torch = torch

    def __call__(self, fn, steps, warmup):
        torch = self.torch
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        evs = []
        for _ in range(steps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            evs.append((a, b))
        torch.cuda.synchronize()
        return float(np.mean([a.elapsed_time(b) for a, b in evs]))


def graph_time(jz, torch, fn, ref, steps, warmup, timer):
    """Capture fn as a HIP graph, check a replay against the eager result `ref`, time the replays.
    Returns (ms, identical, node counts)."""
    g = jz.graph.Captured(fn)
    out = g.run()
    torch.cuda.synchronize()
    g.check()
    flat = lambda o: [t for t in (o if isinstance(o, (tuple, list)) else (o,)) if torch.is_tensor(t)]
    same = all(bool(torch.equal(a, b)) for a, b in zip(flat(out), flat(ref)))
    ms = timer(g.run, steps, warmup)
    g.check()
    nodes = dict(g.node_counts)
    del g, out
    torch.cuda.empty_cache()
    return ms, same, nodes


def train_proposer(jz, torch, dev, sets, k, kcfg, box, steps, quantile):
    t0 = time.time()
    feats, targets = [], []
    for xs in sets:
        f, t = jz.proposer.training_data([torch.from_numpy(xs).to(dev)], k, kcfg, boxsize=box)
        feats.append(f)
        targets.append(t)
    feats, targets = torch.cat(feats), torch.cat(targets)
    prop = jz.proposer.RadiusProposer(k, device=dev, seed=0)
    rep = prop.train(feats, targets, steps=steps, batch=8192, lr=3e-3, lr_final=1e-4)
    prop.calibrate(feats, targets, quantile=quantile)
    torch.cuda.synchronize()
    info = {"k": k, "leaves": int(len(targets)), "steps": steps, "loss_first": rep.loss_first,
            "loss_last": rep.loss_last, "margin_log2": prop.margin, "quantile": quantile,
            "seconds": round(time.time() - t0, 1)}
    del feats, targets
    return prop, info


def learned_vs_plain(jz, torch, dev, x, k, prop, kcfg, box, steps, warmup, timer):
    """Exactness, fallback rate and log2-radius RMSE of the learned search, then both searches timed
    as graph replays."""
    n = len(x)
    plain = lambda: jz.knn.knn(x, k=k, boxsize=box, cfg=kcfg)
    learned = lambda: jz.proposer.knn_learned(x, k, prop, boxsize=box, cfg=kcfg)
    r0, i0 = plain()
    r1, i1, retry = jz.proposer.knn_learned(x, k, prop, boxsize=box, cfg=kcfg, return_retry=True)
    torch.cuda.synchronize()
    identical = bool(torch.equal(i0, i1) and torch.equal(r0, r1))
    fallback = float(retry.float().mean())
    partz, originz, th = jz.tree.zsort_and_tree(x, kcfg.tree, data=jz.data.RankIdx(
        rank=None, idx=torch.arange(n, dtype=torch.int32, device=dev)))
    nl = int(th.num(0))
    target = jz.proposer.leaf_log2_radius(th, r0[originz.idx.long()], k)[:nl]
    pred = prop.predict_log2r(jz.proposer.leaf_features(th)[:nl])
    rmse = float(torch.sqrt(torch.mean((pred.double() - target.double()) ** 2)))
    del partz, originz, th, target, pred, retry
    t_plain, same_p, nodes_p = graph_time(jz, torch, plain, (r0, i0), steps, warmup, timer)
    t_learn, same_l, nodes_l = graph_time(jz, torch, learned, (r1, i1), steps, warmup, timer)
    del r0, i0, r1, i1
    torch.cuda.empty_cache()
    e_plain = timer(plain, max(2, steps // 2), 1)
    e_learn = timer(learned, max(2, steps // 2), 1)
    return {"k": k, "plain_ms": t_plain, "learned_ms": t_learn, "learned_speedup": t_plain / t_learn,
            "plain_queries_per_s": n / (t_plain * 1e-3), "learned_queries_per_s": n / (t_learn * 1e-3),
            "learned_identical": identical and same_l and same_p, "fallback_rate": fallback,
            "log2_radius_rmse": rmse, "eager_plain_ms": e_plain, "eager_learned_ms": e_learn,
            "graph_nodes_plain": nodes_p, "graph_nodes_learned": nodes_l}


def fof_stage_times(jz, torch, x, rlink, fcfg, box, timer, steps):
    """FoF end to end as a graph replay, and its particle-link stage (FofLeaf2Leaf: leaf labels to
    particles, the leaf-pair link kernel, label compression) timed alone on the launch stream."""
    n = len(x)
    fof = lambda: jz.fof.fof_labels(x, rlink, boxsize=box, cfg=fcfg)
    ref = fof()
    t_fof, same, nodes = graph_time(jz, torch, fof, ref, steps, 2, timer)
    del ref
    partz, idz = jz.tree.zsort(x)
    th = jz.tree.build_tree_hierarchy(partz, cfg_tree=fcfg.tree)
    node_data, ilist = jz.fof._fof_dual_walk(th, rlink=rlink, boxsize=box, alloc_fac_ilist=fcfg.alloc_fac_ilist)
    jz.checks.flush()
    posz = jz.data.get_pos(partz)
    link = lambda: jz.fof._fof_leaf2leaf(node_data, ilist, posz, rlink=rlink, boxsize=box)
    t_link = timer(link, steps, 1)
    del partz, posz, idz, th, node_data, ilist
    torch.cuda.empty_cache()
    e2e = n * FOF_BYTES_END_TO_END / (t_fof * 1e-3) / 1e9
    lk = n * FOF_BYTES_LINK / (t_link * 1e-3) / 1e9
    return {"fof_ms": t_fof, "fof_mparticles_per_s": n / (t_fof * 1e-3) / 1e6, "graph_identical": same,
            "graph_nodes": nodes, "link_stage_ms": t_link,
            "roofline_end_to_end": {"bound": "hbm", "bytes_per_particle": FOF_BYTES_END_TO_END, "achieved": e2e,
                                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": e2e / HBM_PEAK_GBS},
            "roofline_link_stage": {"bound": "hbm", "bytes_per_particle": FOF_BYTES_LINK, "achieved": lk,
                                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": lk / HBM_PEAK_GBS}}


def bench_batch(jz, torch, dev, nsets, k, steps, warmup, timer):
    """Batched kNN (k) + FoF (b = 0.2 of each set's own mean spacing) over `nsets` sets: one
    pipeline over all sets (eager and as a replayed HIP graph) against the same pipeline run one
    set at a time.  All three give identical results, which is checked."""
    sizes, xh = _batch_workload(nsets)
    n = int(sizes.sum())
    x = torch.from_numpy(xh).to(dev)
    rl = [0.2 * float(s) ** (-1.0 / 3.0) for s in sizes]
    off = np.concatenate([[0], np.cumsum(sizes)])
    box = 1.0
    ctree, tres = jz.autotune.tune_pipeline_sets(x, sizes, k, rl, boxsize=box, leaf_sizes=(16, 24, 32, 48),
                                                 leaf_waves=(2, 4), sort_kpts=(12, 16), reps=2)
    eager = lambda: jz.batch.pipeline_sets(x, sizes, k, rl, boxsize=box, cfg_tree=ctree)
    ref = [t.clone() for t in eager()]
    g = jz.batch.CapturedSets(x, sizes, k, rl, boxsize=box, cfg_tree=ctree)
    got = g.run()
    torch.cuda.synchronize()
    g.check()
    same_graph = all(bool(torch.equal(a, b)) for a, b in zip(got, ref))

    def sequential():
        outs = []
        for s in range(nsets):
            outs.append(jz.batch.pipeline_sets(x[off[s]:off[s + 1]], sizes[s:s + 1], k, rl[s:s + 1], boxsize=box,
                                               cfg_tree=ctree))
        return outs

    seq = sequential()
    same_seq = all(bool(torch.equal(torch.cat([o[i] for o in seq]), ref[i])) for i in range(3))
    del seq
    t_eager = timer(eager, steps, warmup)
    t_graph = timer(g.run, steps, warmup)
    t_seq = timer(sequential, max(1, steps // 2), 1)
    got = g.run()  # a replay after many others, with eager work in between, must still agree
    torch.cuda.synchronize()
    g.check()
    same_graph = same_graph and all(bool(torch.equal(a, b)) for a, b in zip(got, ref))
    nodes = dict(g.node_counts)
    del g, ref, got
    torch.cuda.empty_cache()
    return {
        "workload": f"{nsets} uniform sets in the periodic unit box, sizes floor(2^U(10,20)) (seed 7), "
                    f"{n} points; kNN k={k} + FoF b=0.2 of each set's mean spacing, on one shared forest",
        "n_points": n, "n_sets": nsets, "k": k,
        "batched_ms": t_eager, "graph_ms": t_graph, "one_set_at_a_time_ms": t_seq,
        "graph_points_per_s": n / (t_graph * 1e-3), "batched_points_per_s": n / (t_eager * 1e-3),
        "one_set_at_a_time_points_per_s": n / (t_seq * 1e-3),
        "graph_identical": same_graph, "one_set_at_a_time_identical": same_seq, "graph_nodes": nodes,
        "autotune": {"leaf_size": tres.leaf_size, "leaf_waves": tres.leaf_waves, "sort_kpt": tres.sort_kpt},
    }


def bench_c5(jz, torch, dev, args, timer, kcfg):
    """2^28 clustered points (60% in Plummer halos, member weights on [32, 2^18]), float64
    positions: z-order keys and tree in float64, kNN (k, learned) and FoF on the float32-rounded
    positions.  Points drawn on the GPU; the proposer is trained on a 2^26 float32 set of the same
    mix.  Learned and plain kNN are compared through exact checksums of their outputs."""
    from gen import clustered_torch
    n, k = args.c5_n, args.k
    nh = int(0.6 * n / 421)
    x = clustered_torch(n, nhalos=nh, halo_frac=0.6, mmax=2 ** 18, seed=11, device=dev)
    xs = clustered_torch(1 << 26, nhalos=max(1, nh * (1 << 26) // n), halo_frac=0.6, mmax=2 ** 18, seed=12,
                         device=dev, dtype=torch.float32)
    ccfg = dataclasses.replace(kcfg, alloc_fac_ilist=256.0)
    f, t = jz.proposer.training_data([xs], k, ccfg, boxsize=1.0)
    del xs
    prop = jz.proposer.RadiusProposer(k, device=dev, seed=0)
    prop.train(f, t, steps=1000, batch=8192, lr=3e-3, lr_final=1e-4)
    prop.calibrate(f, t, quantile=args.quantile)
    del f, t
    torch.cuda.empty_cache()
    f32 = torch.float32
    box = 1.0

    def digest(out):
        r, i = out[0], out[1]
        ci = cr = 0
        for s0 in range(0, n, 1 << 22):
            w = (torch.arange(s0, min(n, s0 + (1 << 22)), device=dev, dtype=torch.int64) % 1009 + 1)[:, None]
            ci += int((i[s0:s0 + (1 << 22)].to(torch.int64) * w).sum())
            cr += int((r[s0:s0 + (1 << 22)].view(torch.int32).to(torch.int64) * w).sum())
        return ci, cr

    res = {}
    for name, fn in (("learned", lambda: jz.proposer.knn_learned(x, k, prop, boxsize=box, cfg=ccfg, dist_dtype=f32)),
                     ("plain", lambda: jz.knn.knn(x, k, boxsize=box, cfg=ccfg, dist_dtype=f32))):
        g = jz.graph.Captured(fn)
        ms = timer(g.run, max(2, args.steps // 2), 1)
        g.check()
        res[name] = (ms, digest(g.out), dict(g.node_counts))
        del g
        torch.cuda.empty_cache()
        log(f"c5 {name} kNN: {ms:.1f} ms")
    rlink = args.b * n ** (-1.0 / 3.0)
    fcfg = jz.config.FofConfig()
    fcfg = dataclasses.replace(fcfg, tree=dataclasses.replace(fcfg.tree, alloc_fac_nodes=2.0))
    g = jz.graph.Captured(lambda: jz.fof.fof_labels(x, rlink, boxsize=box, cfg=fcfg, dist_dtype=f32))
    t_fof = timer(g.run, max(2, args.steps // 2), 1)
    g.check()
    del g
    torch.cuda.empty_cache()
    t_l, t_p = res["learned"][0], res["plain"][0]
    return {"workload": f"clustered {n} points ({nh} halos, 60% in halos, member weights on [32, 2^18]), periodic "
                        f"unit box, float64 positions; float64 keys and tree, float32 distances; kNN k={k} learned "
                        f"and plain, FoF b={args.b}",
            "n_points": n, "k": k, "knn_learned_ms": t_l, "knn_plain_ms": t_p, "learned_speedup": t_p / t_l,
            "learned_queries_per_s": n / (t_l * 1e-3), "learned_identical": res["learned"][1] == res["plain"][1],
            "fof_ms": t_fof, "fof_mparticles_per_s": n / (t_fof * 1e-3) / 1e6,
            "fof_roofline_frac": n * (FOF_BYTES_END_TO_END + 12) / (t_fof * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "graph_nodes_learned": res["learned"][2]}


def cpu_baselines(xh, k, b, side):
    """The points of the benchmark set inside [0, side)^3 (same density and clustering, non-periodic),
    timed on this host's cores: scipy cKDTree build + k-nearest query of every one of them
    (workers=-1; the reference's own CPU comparator), and the oracle's cell FoF with the same
    linking length."""
    from scipy.spatial import cKDTree
    from oracle import oracle as O
    sub = xh[np.all(xh < np.float32(side), axis=1)]
    ns = len(sub)
    cores = len(os.sched_getaffinity(0))
    c0 = time.perf_counter()
    tree = cKDTree(sub)
    tree.query(sub, k, workers=-1)
    ct = time.perf_counter() - c0
    del tree
    knn_cpu = {"value": ns / ct, "unit": "queries/s", "cores": cores, "kind": "reference",
               "sample": f"scipy cKDTree build + query(k={k}, workers=-1) of the {ns} benchmark points in "
                         f"[0,{side})^3 (non-periodic), {ct:.1f} s"}
    rl = b * len(xh) ** (-1.0 / 3.0)
    c0 = time.perf_counter()
    O.fof(sub, rl, boxsize=0.0, key=np.arange(ns), method="cells")
    ft = time.perf_counter() - c0
    fof_cpu = {"value": ns / ft / 1e6, "unit": "Mparticles/s", "cores": O.num_threads(), "kind": "port",
               "sample": f"oracle cell FoF (rlink {rl:.3g}) of the {ns} benchmark points in [0,{side})^3 "
                         f"(non-periodic), {ft:.1f} s"}
    return knn_cpu, fof_cpu


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this build has no "
              "multi-GPU path.", file=sys.stderr)
        sys.exit(2)

    if args.copy_bytes:
        copy_probe(args.copy_bytes, args.steps)
        return
    n, k = args.n, args.k
    # point sets first, in worker processes started before anything touches the GPU
    t0 = time.time()
    seeds = [0] + [1000 + s for s in range(args.train_seeds)]
    jobs = [(n, args.halos, s) for s in seeds]
    if args.n_mid:
        jobs.append((args.n_mid, args.halos_mid, 5))
    with ProcessPoolExecutor(max_workers=min(len(jobs), 6), mp_context=mp.get_context("fork")) as ex:
        sets = list(ex.map(_gen, jobs))
    xh = sets[0]
    x_mid = sets[len(seeds)] if args.n_mid else None
    train_sets = sets[1:len(seeds)]
    del sets
    log(f"generated {len(jobs)} point sets in {time.time() - t0:.1f}s")

    import torch
    import jztree_amd as jz

    dev = torch.device("cuda:0")
    timer = Timer(torch)
    x = torch.from_numpy(xh).to(dev)

    # tree settings: a leaf allocation large enough for strongly clustered sets; the leaf size, the
    # leaf kernel's waves per EU, the sort tile and the XCD order of the leaf kernels are picked by
    # the autotuner on these points (results do not depend on any of them, which the tuner checks)
    kcfg = jz.config.KNNConfig()
    kcfg = dataclasses.replace(kcfg, tree=dataclasses.replace(kcfg.tree, alloc_fac_nodes=2.0))
    fcfg = jz.config.FofConfig()
    fcfg = dataclasses.replace(fcfg, tree=dataclasses.replace(fcfg.tree, alloc_fac_nodes=2.0))
    rlink = args.b * n ** (-1.0 / 3.0)
    box = 1.0
    t0 = time.time()
    if args.no_tune:
        kcfg = dataclasses.replace(kcfg, tree=dataclasses.replace(kcfg.tree, max_leaf_size=128))
        fcfg = dataclasses.replace(fcfg, tree=dataclasses.replace(fcfg.tree, max_leaf_size=128))
        jz.autotune.set_knobs(leaf_waves=4, sort_kpt=16, xcd=1)
        tune_info = {"skipped": True, "leaf_size_knn": 128, "leaf_size_fof": 128, "leaf_waves": 4, "sort_kpt": 16,
                     "xcd": 1}
    else:
        xcd_ms = {}
        for xcd in (0, 1):  # XCD order of the leaf kernels, timed on the default kNN
            jz.autotune.set_knobs(xcd=xcd)
            xcd_ms[xcd] = timer(lambda: jz.knn.knn(x, k=k, boxsize=box, cfg=dataclasses.replace(
                kcfg, tree=dataclasses.replace(kcfg.tree, max_leaf_size=128))), 2, 1)
        best_xcd = min(xcd_ms, key=xcd_ms.get)
        jz.autotune.set_knobs(xcd=best_xcd)
        fcfg, ftune = jz.autotune.tune_fof(x, rlink, boxsize=box, cfg=fcfg, leaf_sizes=(32, 64, 96, 128),
                                           sort_kpts=(12, 16), reps=2)
        kcfg, ktune = jz.autotune.tune_knn(x, k, boxsize=box, cfg=kcfg, leaf_sizes=(64, 96, 128, 160),
                                           leaf_waves=(2, 4), sort_kpts=(12, 16), reps=2)
        tune_info = {"leaf_size_knn": ktune.leaf_size, "leaf_waves": ktune.leaf_waves, "sort_kpt": ktune.sort_kpt,
                     "leaf_size_fof": ftune.leaf_size, "fof_sort_kpt": ftune.sort_kpt, "xcd": best_xcd,
                     "knn_ms_by_xcd_order": xcd_ms, "knn_candidates": len(ktune.table),
                     "fof_candidates": len(ftune.table), "seconds": round(time.time() - t0, 1)}
    knobs = jz.autotune.get_knobs()
    log(f"autotune: {tune_info}")

    # ---- proposers: trained on other seeds ----
    prop, train_info = train_proposer(jz, torch, dev, train_sets, k, kcfg, box, args.train_steps, args.quantile)
    log(f"proposer k={k}: {train_info}")
    prop2 = train2 = None
    if args.k2:
        prop2, train2 = train_proposer(jz, torch, dev, train_sets, args.k2, kcfg, box, args.train_steps, args.quantile)
        log(f"proposer k={args.k2}: {train2}")
    del train_sets

    # ---- headline: learned vs plain kNN (k), graphs ----
    wall0 = time.perf_counter()
    head = learned_vs_plain(jz, torch, dev, x, k, prop, kcfg, box, args.steps, args.warmup, timer)
    log(f"k={k}: {head}")
    # ---- FoF on the same points ----
    fof = fof_stage_times(jz, torch, x, rlink, fcfg, box, timer, args.steps)
    log(f"fof: {fof}")
    # ---- the whole step (learned kNN, then FoF) as one graph ----
    step = lambda: (*jz.proposer.knn_learned(x, k, prop, boxsize=box, cfg=kcfg),
                    jz.fof.fof_labels(x, rlink, boxsize=box, cfg=fcfg)[1])
    sref = step()
    t_step, same_step, nodes_step = graph_time(jz, torch, step, sref, args.steps, args.warmup, timer)
    del sref
    torch.cuda.empty_cache()
    wall = time.perf_counter() - wall0
    log(f"step graph {t_step:.2f} ms (identical {same_step}; timed legs wall {wall:.1f} s)")

    # ---- second k on the same points ----
    second = None
    if prop2 is not None:
        second = learned_vs_plain(jz, torch, dev, x, args.k2, prop2, kcfg, box, max(2, args.steps // 2), 1, timer)
        second["proposer"] = train2
        log(f"k={args.k2}: {second}")

    # ---- dominant kernel: the leaf kernel of the learned search, timed alone on the launch stream ----
    partz, originz, th = jz.tree.zsort_and_tree(x, kcfg.tree, data=jz.data.RankIdx(
        rank=None, idx=torch.arange(n, dtype=torch.int32, device=dev)))
    bound = prop.bound_r2(th)
    ilb = jz.knn._knn_radius_walk(th, k, jz.knn._plane_radii(th, bound), boxsize=box,
                                  alloc_fac_ilist=kcfg.alloc_fac_ilist)
    spl = th.splits_leaf_to_part()
    jz.checks.flush()
    leaf = lambda: jz.knn._knn_leaf2leaf(ilb, spl, partz, k=k, boxsize=box, src_ids=originz.idx,
                                        dst_rows=originz.idx, init_r2=bound)
    t_leaf = timer(leaf, 5, 1)
    alg_bytes = n * (16 + 8 * k)  # read query (12) + origin index (4), write k (radius, id) pairs
    achieved = alg_bytes / (t_leaf * 1e-3) / 1e9
    traffic = LEAF_TRAFFIC_BYTES.get(("k_knn_leaf", n, k, args.halos, kcfg.tree.max_leaf_size, knobs["xcd"]))
    del partz, originz, th, bound, ilb, spl
    torch.cuda.empty_cache()

    # ---- 2^24 points, 4096 halos: FoF and plain kNN ----
    mid = None
    if x_mid is not None:
        xm = torch.from_numpy(x_mid).to(dev)
        nm = len(x_mid)
        rl_m = args.b * nm ** (-1.0 / 3.0)
        mid = fof_stage_times(jz, torch, xm, rl_m, fcfg, box, timer, args.steps)
        knn_m = lambda: jz.knn.knn(xm, k=k, boxsize=box, cfg=kcfg)
        r = knn_m()
        mid["knn_plain_ms"], same_m, _ = graph_time(jz, torch, knn_m, r, args.steps, args.warmup, timer)
        mid["knn_graph_identical"] = same_m
        mid["workload"] = (f"clustered {nm} points ({args.halos_mid} halos), periodic unit box; FoF b={args.b} and "
                           f"plain kNN k={k}")
        del xm, r
        torch.cuda.empty_cache()
        log(f"mid: {mid}")

    c5 = None
    if args.c5_n:
        del x
        torch.cuda.empty_cache()
        c5 = bench_c5(jz, torch, dev, args, timer, kcfg)
        log(f"c5: {c5}")

    batched = None
    if args.batch_sets > 0:
        batched = bench_batch(jz, torch, dev, args.batch_sets, args.batch_k, args.steps, args.warmup, timer)
        log(f"batched: {batched}")

    cpu = fof_cpu = None
    if not args.no_cpu:
        cpu, fof_cpu = cpu_baselines(xh, k, args.b, args.cpu_side)
        log(f"cpu: {cpu} {fof_cpu}")

    t_learn = head["learned_ms"]
    out = {
        "metric": "kNN queries/s (k=32, 2^26 pts) learned vs plain; FoF Mparticles/s; %HBM peak",
        "value": n / (t_learn * 1e-3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded clustered point sets, generated in-process; proposers trained in-process)",
        "config": {"workload": f"clustered 2^{int(np.log2(n))} points ({args.halos} Plummer halos, 70% in halos), "
                               f"periodic unit box; exact kNN k={k} with the learned radius proposer, "
                               f"then FoF b={args.b}; every call a HIP graph replay",
                   "n_points": n, "k": k, "fof_b": args.b, "leaf_size_knn": kcfg.tree.max_leaf_size,
                   "leaf_size_fof": fcfg.tree.max_leaf_size},
        "autotune": tune_info,
        "knobs": knobs,
        "plain_queries_per_s": head["plain_queries_per_s"],
        "learned_speedup": head["learned_speedup"],
        "learned_identical": head["learned_identical"],
        "fallback_rate": head["fallback_rate"],
        "log2_radius_rmse": head["log2_radius_rmse"],
        "proposer": train_info,
        "knn_plain_ms": head["plain_ms"],
        "knn_learned_ms": t_learn,
        "knn_eager_plain_ms": head["eager_plain_ms"],
        "knn_eager_learned_ms": head["eager_learned_ms"],
        "step_graph_ms": t_step,
        "step_graph_identical": same_step,
        "step_graph_nodes": nodes_step,
        "fof_ms": fof["fof_ms"],
        "fof_mparticles_per_s": fof["fof_mparticles_per_s"],
        "fof": fof,
        "second_k": second,
        "roofline": {"bound": "hbm", "kernel": "k_knn_leaf (KnnLeaf2Leaf, bounded)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel_ms": t_leaf, "alg_bytes_per_launch": alg_bytes},
        "fof_roofline": fof["roofline_end_to_end"],
        "cpu_baseline": cpu,
        "fof_cpu_baseline": fof_cpu,
        "mid": mid,
        "c5": c5,
        "batched": batched,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
