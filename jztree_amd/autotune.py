"""Autotuner for the tree hot path (extension: the reference has no autotuner).

Picks, by timing candidates on the caller's own points:

  leaf_size   TreeConfig.max_leaf_size: the query and source granularity of the leaf kernels
  leaf_waves  waves per EU the leaf kNN kernel is compiled for (k <= 32): its register budget
              against occupancy (library knob JZT_TUNE_LEAF_WAVES)
  sort_kpt    keys per thread of the onesweep radix passes, i.e. the sort's tile of 256 * kpt
              keys (library knob JZT_TUNE_SORT_KPT)

None of them changes a result: the tuner checks that every candidate's output equals the first
candidate's, bit for bit, and raises otherwise.  The chosen knobs are applied to the library, the
chosen leaf size is returned in a config, and the choice can be cached in a JSON file keyed by a
signature of the workload (kind, log2 of the point count, k, dtype, dimension, periodicity, set
count), so that later runs skip the timing.

    cfg, res = jz.autotune.tune_knn(pos, k=32, boxsize=1.0, cache="tune.json")
    rnn, inn = jz.knn.knn(pos, k=32, boxsize=1.0, cfg=cfg)
"""
from __future__ import annotations

import dataclasses
import json
import math
import os
from dataclasses import dataclass, field
from typing import Callable, Sequence

import torch

from . import _lib as L
from .config import KNNConfig

KNOBS = {"leaf_waves": 0, "sort_kpt": 1, "xcd": 2}  # JZT_TUNE_LEAF_WAVES, JZT_TUNE_SORT_KPT, JZT_TUNE_XCD
ALLOWED = {"leaf_waves": (1, 2, 4), "sort_kpt": (8, 12, 16), "xcd": (0, 1)}


def set_knobs(**values):
    """Set library tuning knobs by name (leaf_waves, sort_kpt, xcd)."""
    for name, v in values.items():
        if name not in KNOBS:
            raise ValueError(f"unknown knob {name!r}; known: {sorted(KNOBS)}")
        L.call("jzt_set_tuning", KNOBS[name], int(v))


def get_knobs() -> dict:
    return {name: int(L.LIB.jzt_get_tuning(i)) for name, i in KNOBS.items()}


@dataclass
class TuneResult:
    signature: str
    leaf_size: int
    leaf_waves: int
    sort_kpt: int
    ms: float                       # time of the chosen setting
    table: list = field(default_factory=list)  # (setting dict, ms) of every candidate timed
    cached: bool = False


def signature(kind: str, n: int, k: int, dtype, dim: int, boxsize, nsets: int = 1) -> str:
    lg = int(round(math.log2(max(n, 1))))
    return f"{kind}:n2^{lg}:k{k}:{str(dtype).replace('torch.', '')}:d{dim}:{'periodic' if boxsize else 'open'}:sets{nsets}"


def _timer(fn: Callable, reps: int) -> float:
    """Median of `reps` timed calls (HIP events on the current stream) after one untimed call."""
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b))
    ts.sort()
    return ts[len(ts) // 2]


def _same(a, b) -> bool:
    if isinstance(a, (tuple, list)):
        return len(a) == len(b) and all(_same(x, y) for x, y in zip(a, b))
    return bool(torch.equal(a, b))


def search(run: Callable[[dict], object], time: Callable[[Callable], float], sort_kpts: Sequence[int],
           leaf_sizes: Sequence[int], leaf_waves: Sequence[int]):
    """Two-stage search.  Stage 1 times the sort tiles with the first leaf size and leaf-wave count;
    stage 2 times every (leaf size, leaf waves) pair with the best tile.  `run(setting)` computes
    the workload for a setting and returns its output; `time(fn)` times a callable in ms.  Every
    output must equal the first.  Returns (best setting, best ms, table)."""
    table = []
    ref = []

    def measure(setting):
        out = run(setting)
        if not ref:
            ref.append(out)
        elif not _same(out, ref[0]):
            raise RuntimeError(f"autotune: setting {setting} changed the result")
        del out
        ms = time(lambda: run(setting))
        table.append((dict(setting), ms))
        return ms

    base = {"leaf_size": leaf_sizes[0], "leaf_waves": leaf_waves[0]}
    kpt = min(sort_kpts, key=lambda v: measure({**base, "sort_kpt": v}))
    best, best_ms = None, math.inf
    for ls in leaf_sizes:
        for w in leaf_waves:
            s = {"leaf_size": ls, "leaf_waves": w, "sort_kpt": kpt}
            hit = [ms for st, ms in table if st == s]
            ms = hit[0] if hit else measure(s)
            if ms < best_ms:
                best, best_ms = s, ms
    return best, best_ms, table


def _load_cache(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def _store_cache(path, sig, res: TuneResult):
    if not path:
        return
    data = _load_cache(path)
    data[sig] = {"leaf_size": res.leaf_size, "leaf_waves": res.leaf_waves, "sort_kpt": res.sort_kpt, "ms": res.ms}
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _with_leaf(cfg_tree, leaf_size):
    return dataclasses.replace(cfg_tree, max_leaf_size=int(leaf_size))


def _tune(kind, n, k, dtype, dim, boxsize, nsets, run, cache, reps, sort_kpts, leaf_sizes, leaf_waves):
    sig = signature(kind, n, k, dtype, dim, boxsize, nsets)
    hit = _load_cache(cache).get(sig)
    if hit is not None:
        set_knobs(leaf_waves=hit["leaf_waves"], sort_kpt=hit["sort_kpt"])
        return TuneResult(sig, hit["leaf_size"], hit["leaf_waves"], hit["sort_kpt"], hit.get("ms", float("nan")),
                          cached=True)
    before = get_knobs()
    try:
        best, ms, table = search(run, lambda fn: _timer(fn, reps), sort_kpts, leaf_sizes, leaf_waves)
    except BaseException:
        set_knobs(**before)
        raise
    set_knobs(leaf_waves=best["leaf_waves"], sort_kpt=best["sort_kpt"])
    res = TuneResult(sig, best["leaf_size"], best["leaf_waves"], best["sort_kpt"], ms, table)
    _store_cache(cache, sig, res)
    return res


def tune_knn(pos: torch.Tensor, k: int, boxsize: float | None = None, cfg: KNNConfig = KNNConfig(),
             leaf_sizes: Sequence[int] = (32, 48, 64), leaf_waves: Sequence[int] = (1, 2, 4),
             sort_kpts: Sequence[int] = (8, 12, 16), reps: int = 3, cache: str | None = None):
    """Tune knn() on these points.  Returns (cfg with the chosen leaf size, TuneResult); the chosen
    knobs are applied to the library."""
    from .knn import knn

    def run(s):
        set_knobs(leaf_waves=s["leaf_waves"], sort_kpt=s["sort_kpt"])
        c = dataclasses.replace(cfg, tree=_with_leaf(cfg.tree, s["leaf_size"]))
        return knn(pos, k=k, boxsize=boxsize, cfg=c)

    res = _tune("knn", len(pos), k, pos.dtype, pos.shape[-1], boxsize, 1, run, cache, reps, sort_kpts,
                leaf_sizes, leaf_waves)
    return dataclasses.replace(cfg, tree=_with_leaf(cfg.tree, res.leaf_size)), res


def tune_pipeline_sets(pos: torch.Tensor, sizes, k: int, rlink, boxsize: float = 0., cfg_tree=None,
                       leaf_sizes: Sequence[int] = (16, 24, 32, 48), leaf_waves: Sequence[int] = (1, 2, 4),
                       sort_kpts: Sequence[int] = (8, 12, 16), reps: int = 3, cache: str | None = None):
    """Tune batch.pipeline_sets (sort, forest, kNN, FoF of a batch of sets).  Returns (tree config
    with the chosen leaf size, TuneResult); the chosen knobs are applied to the library."""
    from .batch import pipeline_sets
    cfg_tree = cfg_tree if cfg_tree is not None else KNNConfig().tree
    nsets = len(sizes) if not torch.is_tensor(sizes) or not sizes.is_cuda else len(sizes) - 1

    def run(s):
        set_knobs(leaf_waves=s["leaf_waves"], sort_kpt=s["sort_kpt"])
        return pipeline_sets(pos, sizes, k, rlink, boxsize=boxsize, cfg_tree=_with_leaf(cfg_tree, s["leaf_size"]))

    res = _tune("pipeline_sets", len(pos), k, pos.dtype, pos.shape[-1], boxsize, nsets, run, cache, reps,
                sort_kpts, leaf_sizes, leaf_waves)
    return _with_leaf(cfg_tree, res.leaf_size), res


def tune_fof(pos: torch.Tensor, rlink: float, boxsize: float = 0., cfg=None,
             leaf_sizes: Sequence[int] = (32, 64, 96, 128), sort_kpts: Sequence[int] = (8, 12, 16), reps: int = 3,
             cache: str | None = None):
    """Tune fof_labels on these points over leaf size and sort tile (the leaf kNN waves knob is
    left as it is).  Returns (FofConfig with the chosen leaf size, TuneResult)."""
    from .config import FofConfig
    from .fof import fof_labels
    cfg = cfg if cfg is not None else FofConfig()
    waves = get_knobs()["leaf_waves"]

    def run(s):
        set_knobs(sort_kpt=s["sort_kpt"])
        return fof_labels(pos, rlink, boxsize=boxsize,
                          cfg=dataclasses.replace(cfg, tree=_with_leaf(cfg.tree, s["leaf_size"])))[1]

    res = _tune("fof", len(pos), 0, pos.dtype, pos.shape[-1], boxsize, 1, run, cache, reps, sort_kpts,
                leaf_sizes, (waves,))
    return dataclasses.replace(cfg, tree=_with_leaf(cfg.tree, res.leaf_size)), res
