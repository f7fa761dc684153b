"""CPU restatement of the learned radius proposer (test infrastructure only).

The proposer has no counterpart in the reference; this file states in float64 numpy what the HIP
kernels of jztree_amd/csrc/proposer.hip compute, so that tests can check them:

  leaf_features   the 12 per-leaf features
  bf16_round      float32 -> bf16 (round to nearest, ties to even) -> float32
  forward         12-64-64-64-1 MLP, GELU (erf form), weights as given (callers pass the bf16 values)
  loss_and_grad   mean squared error over a batch and its gradient with respect to the weights
  adam_step       one bias-corrected Adam step

Parity pins (tests/test_proposer.py): loss_and_grad against central finite differences of
forward; GELU's derivative against finite differences; bf16_round against hand-worked halfway
cases; adam_step against its closed form for the first step (w - lr * sign(g)); leaf_features
against a hand-computed tree.  Only tests/, smoke() and bench.py's cpu_baseline leg use it.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

NF, NH = 12, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + NF * NH
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH * NH
OFF_W4 = OFF_B3 + NH
OFF_B4 = OFF_W4 + NH
NPARAM = OFF_B4 + 1


def unpack(p: np.ndarray):
    """Flat parameter vector -> (W1 (12,64), b1, W2 (64,64), b2, W3, b3, W4 (64,), b4), input-major."""
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(NF, NH), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(NH, NH), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3].reshape(NH, NH), p[OFF_B3:OFF_W4], p[OFF_W4:OFF_B4], p[OFF_B4])


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 values rounded to bf16 (nearest, ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan_inf = (u & 0x7F800000) == 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    quiet = np.where((u & 0xFFFF) != 0, np.uint64(0x40), np.uint64(0))
    r = np.where(nan_inf, (u >> 16) | quiet, r)
    return (r.astype(np.uint32) << 16).view(np.float32)


def forward(p: np.ndarray, x: np.ndarray, return_cache: bool = False):
    """y (batch,) for standardised features x (batch, 12)."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ W1 + b1
    a1 = gelu(z1)
    z2 = a1 @ W2 + b2
    a2 = gelu(z2)
    z3 = a2 @ W3 + b3
    a3 = gelu(z3)
    y = a3 @ W4 + b4
    if return_cache:
        return y, (x, z1, a1, z2, a2, z3, a3)
    return y


def loss_and_grad(p: np.ndarray, x: np.ndarray, t: np.ndarray):
    """loss = mean (y - t)^2 and its gradient (NPARAM,) with respect to p."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    y, (x, z1, a1, z2, a2, z3, a3) = forward(p, x, return_cache=True)
    b = len(y)
    e = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(e * e))
    dy = 2.0 * e / b
    g = np.zeros(NPARAM)
    g[OFF_W4:OFF_B4] = a3.T @ dy
    g[OFF_B4] = dy.sum()
    d3 = np.outer(dy, W4) * gelu_grad(z3)
    g[OFF_W3:OFF_B3] = (a2.T @ d3).ravel()
    g[OFF_B3:OFF_W4] = d3.sum(0)
    d2 = (d3 @ W3.T) * gelu_grad(z2)
    g[OFF_W2:OFF_B2] = (a1.T @ d2).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(0)
    d1 = (d2 @ W2.T) * gelu_grad(z1)
    g[OFF_W1:OFF_B1] = (x.T @ d1).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(0)
    return loss, g


def adam_step(w, m, v, g, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One bias-corrected Adam step; returns (w, m, v) as float64."""
    w, m, v, g = (np.asarray(a, dtype=np.float64) for a in (w, m, v, g))
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def _lg(n):
    return np.log2(np.maximum(np.asarray(n, dtype=np.float64), 1.0))


def leaf_features(npart0, lvl0, leaf_parent, npart1, lvl1, n2c1, parent_parent, npart2, lvl2, nleaves, size):
    """(size, 12) features as k_leaf_features defines them (slots >= nleaves are zero)."""
    f = np.zeros((size, NF))
    for i in range(min(nleaves, size)):
        p = min(max(int(leaf_parent[i]), 0), size - 1)
        g = min(max(int(parent_parent[p]), 0), size - 1) if parent_parent is not None else -1
        n = npart0[i]
        lv = lvl0[i] / 3.0
        im = i - 1 if i > 0 else i
        ip = i + 1 if i + 1 < nleaves else i
        nsum = sum(int(npart0[d]) for d in range(max(i - 2, 0), min(i + 2, nleaves - 1) + 1))
        f[i] = [_lg(n), lv, _lg(npart1[p]), lvl1[p] / 3.0,
                _lg(npart2[g]) if g >= 0 else _lg(npart1[p]), lvl2[g] / 3.0 if g >= 0 else lvl1[p] / 3.0,
                _lg(npart0[im]), _lg(npart0[ip]), lvl0[im] / 3.0 - lv, lvl0[ip] / 3.0 - lv,
                _lg(n2c1[p + 1] - n2c1[p]), _lg(nsum)]
    return f
