"""Benchmark of the tree hot path on one MI355X.

Workload: 2^26 clustered float32 points in the periodic unit box (70% in 65,536 Plummer halos with
scale radius 0.002 and power-law member weights dN/dM ~ M^-1.9 on [32, 2^16], 30% uniform; seeded,
synthetic).  One step = exact kNN (k=32, periodic, neighbours in input order) followed by FoF
labels (linking length 0.2 x mean spacing, periodic), each from positions resident in HBM: z-order
sort, tree build, dual tree walk, leaf kernels.

Prints one JSON line.  value = kNN queries/s of the step's kNN call; FoF throughput is reported
next to it.  roofline = the kNN leaf kernel (the dominant kernel of the step), timed with HIP
events on the launch stream; cpu_baseline = the oracle's grid kNN on the host cores for a bounded
sample of queries of the same point set.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=1 << 26)
    p.add_argument("--halos", type=int, default=65536)
    p.add_argument("--k", type=int, default=32)
    p.add_argument("--b", type=float, default=0.2)
    p.add_argument("--cpu-queries", type=int, default=1 << 20)
    p.add_argument("--no-cpu", action="store_true")
    return p.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this build has no "
              "multi-GPU path.", file=sys.stderr)
        sys.exit(2)

    import torch
    import jztree_amd as jz
    from gen import clustered

    dev = torch.device("cuda:0")
    n, k = args.n, args.k
    t0 = time.time()
    xh = clustered(n, nhalos=args.halos, seed=0)
    print(f"# generated {n} points in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    x = torch.from_numpy(xh).to(dev)

    # tree settings for this workload: leaves of <= 64 points (one per lane of a wave64) and a
    # leaf allocation large enough for strongly clustered sets; results do not depend on either
    kcfg = jz.config.KNNConfig()
    kcfg = dataclasses.replace(kcfg, tree=dataclasses.replace(kcfg.tree, alloc_fac_nodes=2.0, max_leaf_size=64))
    fcfg = jz.config.FofConfig()
    fcfg = dataclasses.replace(fcfg, tree=dataclasses.replace(fcfg.tree, alloc_fac_nodes=2.0, max_leaf_size=64))
    rlink = args.b * n ** (-1.0 / 3.0)

    def ev():
        return torch.cuda.Event(enable_timing=True)

    def step():
        e0, e1, e2 = ev(), ev(), ev()
        e0.record()
        rnn, inn = jz.knn.knn(x, k=k, boxsize=1.0, cfg=kcfg)
        e1.record()
        _, ig = jz.fof.fof_labels(x, rlink, boxsize=1.0, cfg=fcfg)
        e2.record()
        return e0, e1, e2

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        print(f"# warmup {i} done", file=sys.stderr, flush=True)

    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    evs = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    t_knn = float(np.mean([a.elapsed_time(b) for a, b, _ in evs]))
    t_fof = float(np.mean([b.elapsed_time(c) for _, b, c in evs]))
    print(f"# timed {args.steps} steps: knn {t_knn:.2f} ms, fof {t_fof:.2f} ms, wall {wall:.2f} s",
          file=sys.stderr, flush=True)

    # dominant kernel: the kNN leaf kernel, timed alone on the same stream (one launch per call)
    partz, originz, th = jz.tree.zsort_and_tree(x, kcfg.tree, data=jz.data.RankIdx(
        rank=None, idx=torch.arange(n, dtype=torch.int32, device=dev)))
    ilist = jz.knn._knn_dual_walk(th, k, boxsize=1.0, alloc_fac_ilist=kcfg.alloc_fac_ilist)
    spl = th.splits_leaf_to_part()
    jz.checks.flush()
    jz.knn._knn_leaf2leaf(ilist, spl, partz, k=k, boxsize=1.0, src_ids=originz.idx)
    torch.cuda.synchronize()
    reps = 5
    ka, kb = ev(), ev()
    ka.record()
    for _ in range(reps):
        jz.knn._knn_leaf2leaf(ilist, spl, partz, k=k, boxsize=1.0, src_ids=originz.idx)
    kb.record()
    torch.cuda.synchronize()
    t_leaf = ka.elapsed_time(kb) / reps
    alg_bytes = n * (16 + 8 * k)  # read query (12) + origin index (4), write k (radius, id) pairs
    achieved = alg_bytes / (t_leaf * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        nq = min(args.cpu_queries, n)
        qsel = np.random.default_rng(1).choice(n, nq, replace=False)
        c0 = time.perf_counter()
        O.knn(xh, k, xq=xh[qsel], boxsize=1.0)
        ct = time.perf_counter() - c0
        cpu = {"value": nq / ct, "unit": "queries/s", "cores": O.num_threads(), "kind": "port",
               "sample": f"oracle grid kNN (k={k}, periodic) of {nq} random queries against all {n} points, "
                         f"incl. grid build ({ct:.1f} s)"}

    out = {
        "metric": "kNN queries/s (k=32, 2^26 pts) learned vs plain; FoF Mparticles/s; %HBM peak",
        "value": n / (t_knn * 1e-3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_knn + t_fof,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded clustered point set, generated in-process)",
        "config": {"workload": f"clustered 2^{int(np.log2(n))} points ({args.halos} Plummer halos, 70% in halos), "
                               f"periodic unit box; kNN k={k} (plain) + FoF b={args.b}",
                   "n_points": n, "k": k, "fof_b": args.b, "knn_method": "plain"},
        "knn_ms": t_knn,
        "fof_ms": t_fof,
        "fof_mparticles_per_s": n / (t_fof * 1e-3) / 1e6,
        "roofline": {"bound": "hbm", "kernel": "k_knn_leaf (KnnLeaf2Leaf)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel_ms": t_leaf, "alg_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
