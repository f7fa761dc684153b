"""Deferred device-side error checks (the role of jax_ext.raise_if, src/jztree/jax_ext.py:174-204).

The reference raises from an io_callback when a device-side condition fires (allocation
overflow).  Here a condition is a device tensor; it is recorded without synchronising and all
recorded conditions are read back with one copy when the public entry point returns
(``flush``), so the hot path has at most one host synchronisation per call and none inside.
"""
from __future__ import annotations

import contextlib
import threading

import torch

_tls = threading.local()


def _pending():
    if not hasattr(_tls, "pending"):
        _tls.pending = []
    return _tls.pending


def raise_if(cond: torch.Tensor, msg: str, **values):
    """Record `cond` (bool device tensor); `msg` is formatted with `values` (device scalars)."""
    _pending().append((cond.reshape(()), msg, values))


def clear():
    _pending().clear()


def take() -> list:
    """Hand over the recorded conditions (and forget them here), e.g. to evaluate them after each
    replay of a captured graph whose buffers they live in."""
    items = list(_pending())
    _pending().clear()
    return items


@contextlib.contextmanager
def deferred():
    """Within this block nothing reads device data on the host, so the calls can be captured into
    a HIP graph: flush() keeps the recorded conditions for take() (to be evaluated after each
    replay), and passes whose launch depends on device data (retries of short queries) are
    launched unconditionally; they do nothing when there is nothing to redo."""
    prev = getattr(_tls, "deferred", False)
    _tls.deferred = True
    try:
        yield
    finally:
        _tls.deferred = prev


def is_deferred() -> bool:
    return getattr(_tls, "deferred", False)


def flush():
    """Raise RuntimeError for the first recorded condition that holds; then forget all.  Inside
    deferred() the conditions are kept instead."""
    if is_deferred():
        return
    evaluate(take())


def evaluate(items: list):
    """Raise RuntimeError for the first of the given recorded conditions that holds."""
    if not items:
        return
    conds = torch.stack([c.to(torch.bool) for c, _, _ in items]).cpu()
    bad = [i for i in range(len(items)) if bool(conds[i])]
    if bad:
        _, msg, values = items[bad[0]]
        vals = {k: (v.item() if torch.is_tensor(v) else v) for k, v in values.items()}
        raise RuntimeError(msg.format(**vals))
