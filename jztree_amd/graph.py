"""Whole calls captured as HIP graphs and replayed (extension).

The reference's FFI handlers are command-buffer compatible (src/jztree_cuda/generated/ffi_knn.cu:186,
`kCmdBufferCompatible`): XLA may record a jitted sort -> tree -> kNN -> FoF chain once and replay
it.  Here the same holds for the library's calls: inside `checks.deferred()` no call reads device
data on the host (list overflows and other device-side conditions are recorded, passes that redo
short queries are launched unconditionally), every constant is written by a fill kernel, and the
library clears memory with fill kernels, so a capture holds no memset node.

    g = Captured(lambda: (knn(x, 32, boxsize=1.0), fof_labels(x, rl, boxsize=1.0)))
    (rnn, inn), (xz, lab) = g.run()   # replays; g.check() raises if a list overflowed

The outputs are the graph's own buffers, overwritten by the next replay.  Inputs are read from the
tensors the callable closes over: copy new data into them (same shapes) between replays.
"""
from __future__ import annotations

from typing import Any, Callable

import torch

from . import _lib as L
from . import checks


class Captured:
    def __init__(self, fn: Callable[[], Any], warmup: bool = True):
        self._fn = fn
        if warmup:
            # eager run on a side stream: one-time set-up (library loads, allocator growth) stays
            # out of the capture, and its result is checked
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                try:
                    fn()
                    checks.flush()
                except BaseException:
                    checks.clear()
                    raise
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph(keep_graph=True)
        checks.clear()
        try:
            with checks.deferred(), torch.cuda.graph(self.graph):
                self.out = fn()
        except BaseException:
            checks.clear()
            raise
        self._conds = checks.take()
        counts = torch.zeros(4, dtype=torch.int64)
        L.call("jzt_graph_count_nodes", self.graph.raw_cuda_graph(), counts.data_ptr())
        self.node_counts = dict(zip(("kernel", "memcpy", "memset", "other"), counts.tolist()))
        self.graph.instantiate()

    def memset_nodes(self) -> int:
        return self.node_counts["memset"]

    def run(self):
        """Replay the captured work; returns the outputs of the captured call."""
        self.graph.replay()
        return self.out

    def check(self):
        """Raise if the last replay recorded a device-side error (synchronises)."""
        checks.evaluate(self._conds)
