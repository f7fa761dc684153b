"""Autotuner: search logic and knob plumbing on the CPU (the knobs are host-side library state),
tuning runs on the GPU with every candidate's output checked equal."""
import json

import numpy as np
import pytest
import torch

from gen import uniform


def test_knobs_roundtrip_and_validation():
    import jztree_amd as jz
    before = jz.autotune.get_knobs()
    assert before["leaf_waves"] in (1, 2, 4) and before["sort_kpt"] in (8, 12, 16) and before["xcd"] in (0, 1)
    try:
        jz.autotune.set_knobs(leaf_waves=2, sort_kpt=16, xcd=0)
        assert jz.autotune.get_knobs() == {"leaf_waves": 2, "sort_kpt": 16, "xcd": 0}
        with pytest.raises(RuntimeError):
            jz.autotune.set_knobs(leaf_waves=3)
        with pytest.raises(RuntimeError):
            jz.autotune.set_knobs(sort_kpt=10)
        with pytest.raises(RuntimeError):
            jz.autotune.set_knobs(xcd=2)
        with pytest.raises(ValueError):
            jz.autotune.set_knobs(tile=4)
        assert jz.autotune.get_knobs() == {"leaf_waves": 2, "sort_kpt": 16, "xcd": 0}
    finally:
        jz.autotune.set_knobs(**before)


def test_search_picks_fastest_and_checks_results():
    import jztree_amd as jz
    cost = {}
    for kpt in (8, 12, 16):
        for ls in (16, 32, 64):
            for w in (1, 2, 4):
                cost[(ls, w, kpt)] = abs(kpt - 12) * 3 + abs(ls - 32) / 8 + abs(w - 2) + 10
    state = {}
    out = torch.arange(5)

    def run(s):
        state["s"] = s
        return out

    def time(fn):
        fn()
        s = state["s"]
        return cost[(s["leaf_size"], s["leaf_waves"], s["sort_kpt"])]

    best, ms, table = jz.autotune.search(run, time, (8, 12, 16), (16, 32, 64), (1, 2, 4))
    assert best == {"leaf_size": 32, "leaf_waves": 2, "sort_kpt": 12} and ms == 10
    assert len(table) == 3 + 9 - 1  # the stage-1 candidate with the best tile is not timed twice

    calls = {"n": 0}

    def run_bad(s):
        calls["n"] += 1
        return torch.arange(5) + (1 if s["leaf_size"] == 64 else 0)

    with pytest.raises(RuntimeError, match="changed the result"):
        jz.autotune.search(run_bad, time, (12,), (16, 64), (2,))


def test_signature_and_cache(tmp_path):
    import jztree_amd as jz
    sig = jz.autotune.signature("knn", (1 << 26) + 5, 32, torch.float32, 3, 1.0)
    assert sig == "knn:n2^26:k32:float32:d3:periodic:sets1"
    path = str(tmp_path / "tune.json")
    res = jz.autotune.TuneResult(sig, 48, 2, 8, 12.5)
    jz.autotune._store_cache(path, sig, res)
    assert json.load(open(path))[sig] == {"leaf_size": 48, "leaf_waves": 2, "sort_kpt": 8, "ms": 12.5}
    before = jz.autotune.get_knobs()
    try:
        got = jz.autotune._tune("knn", (1 << 26) + 5, 32, torch.float32, 3, 1.0, 1, None, path, 1, (12,), (32,), (4,))
        assert got.cached and (got.leaf_size, got.leaf_waves, got.sort_kpt) == (48, 2, 8)
        got_knobs = jz.autotune.get_knobs()
        assert (got_knobs["leaf_waves"], got_knobs["sort_kpt"]) == (2, 8)
    finally:
        jz.autotune.set_knobs(**before)


@pytest.mark.gpu
def test_tune_knn_and_batch_on_gpu(dev, tmp_path):
    import jztree_amd as jz
    before = jz.autotune.get_knobs()
    try:
        x = torch.from_numpy(uniform(1 << 16, seed=3)).to(dev)
        print("[autotune] knn 2^16 k=16", flush=True)
        cfg, res = jz.autotune.tune_knn(x, 16, boxsize=1.0, leaf_sizes=(32, 64), leaf_waves=(2, 4),
                                        sort_kpts=(8, 16), reps=1, cache=str(tmp_path / "t.json"))
        print(f"[autotune] chose {res}", flush=True)
        assert cfg.tree.max_leaf_size == res.leaf_size and len(res.table) >= 5
        kn = jz.autotune.get_knobs()
        assert (kn["leaf_waves"], kn["sort_kpt"]) == (res.leaf_waves, res.sort_kpt)
        r1, i1 = jz.knn.knn(x, 16, boxsize=1.0, cfg=cfg)
        jz.autotune.set_knobs(**before)
        r0, i0 = jz.knn.knn(x, 16, boxsize=1.0)
        assert torch.equal(i0, i1) and torch.equal(r0, r1)
        sizes = np.array([3000, 1200, 5000, 2048])
        xb = torch.from_numpy(np.concatenate([uniform(int(s), seed=40 + i) for i, s in enumerate(sizes)])).to(dev)
        rl = [0.2 * float(s) ** (-1 / 3) for s in sizes]
        print("[autotune] pipeline_sets", flush=True)
        ct, rb = jz.autotune.tune_pipeline_sets(xb, sizes, 8, rl, boxsize=1.0, leaf_sizes=(16, 32), leaf_waves=(4,),
                                                sort_kpts=(12,), reps=1)
        assert ct.max_leaf_size == rb.leaf_size
    finally:
        jz.autotune.set_knobs(**before)
