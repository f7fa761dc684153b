"""Seeded synthetic point sets (numpy default_rng; float32 in [0, 1) unless noted).

clustered(): a fraction of the points in Plummer-profile halos (scale radius a, radius
a/sqrt(u^(-2/3)-1) truncated at u <= 0.999, isotropic), the rest uniform, in a periodic unit box.
Halo member counts follow dN/dM ~ M^-1.9 on [mmin, mmax]; since sum(M) of the drawn halos is far
below the requested halo fraction, the drawn counts are used as weights and the halo particles are
assigned to halos by multinomial sampling proportional to them (shape kept, totals exact).
"""
from __future__ import annotations

import numpy as np


def uniform(n: int, dim: int = 3, seed: int = 0, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x = rng.random((n, dim), dtype=np.float32)
    if lo != 0.0 or hi != 1.0:
        x = (np.float32(lo) + x * np.float32(hi - lo)).astype(np.float32)
    return x


def _powerlaw(rng, n, alpha, mmin, mmax):
    # inverse CDF of dN/dM ~ M^-alpha on [mmin, mmax]
    u = rng.random(n)
    a1 = 1.0 - alpha
    return (mmin ** a1 + u * (mmax ** a1 - mmin ** a1)) ** (1.0 / a1)


def clustered(n: int, nhalos: int = 4096, halo_frac: float = 0.7, a: float = 0.002, mmin: float = 32,
              mmax: float = 2 ** 16, alpha: float = 1.9, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    nh = int(round(halo_frac * n))
    w = _powerlaw(rng, nhalos, alpha, mmin, mmax)
    counts = rng.multinomial(nh, w / w.sum())
    cent = rng.random((nhalos, 3))
    hid = np.repeat(np.arange(nhalos), counts)
    u = rng.uniform(1e-12, 0.999, nh)
    r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    v = rng.normal(size=(nh, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    xh = cent[hid] + r[:, None] * v
    xb = rng.random((n - nh, 3))
    x = np.concatenate([xh, xb]) % 1.0
    x = x.astype(np.float32)
    x[x >= np.float32(1.0)] = np.nextafter(np.float32(1.0), np.float32(0.0))
    return x[rng.permutation(n)]


def clustered_torch(n: int, nhalos: int, halo_frac: float = 0.6, a: float = 0.002, mmin: float = 32,
                    mmax: float = 2 ** 18, alpha: float = 1.9, seed: int = 0, device="cuda", dtype=None):
    """clustered() drawn on the GPU with a torch generator (for sizes where the numpy generator
    takes minutes): the same construction -- power-law halo weights, halo of each halo particle
    by inverse CDF of the weights, Plummer radii truncated at u <= 0.999, isotropic directions,
    uniform background, periodic unit box, a random permutation -- in float64, returned in
    `dtype` (default float64) on `device`.  Reproducible for a seed on the same GPU and torch."""
    import torch
    dtype = torch.float64 if dtype is None else dtype
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    f64 = dict(dtype=torch.float64, device=device)
    nh = int(round(halo_frac * n))
    u = torch.rand(nhalos, generator=g, **f64)
    a1 = 1.0 - alpha
    w = (mmin ** a1 + u * (mmax ** a1 - mmin ** a1)) ** (1.0 / a1)
    cdf = torch.cumsum(w, 0)
    cdf = cdf / cdf[-1]
    cent = torch.rand((nhalos, 3), generator=g, **f64)
    x = torch.empty((n, 3), **f64)
    step = 1 << 24  # halo particles drawn in slices to bound the temporaries
    for s in range(0, nh, step):
        m = min(step, nh - s)
        hid = torch.searchsorted(cdf, torch.rand(m, generator=g, **f64)).clamp_(max=nhalos - 1)
        uu = torch.rand(m, generator=g, **f64) * (0.999 - 1e-12) + 1e-12
        r = a / torch.sqrt(uu ** (-2.0 / 3.0) - 1.0)
        v = torch.randn((m, 3), generator=g, **f64)
        v /= torch.linalg.norm(v, dim=1, keepdim=True)
        x[s:s + m] = cent[hid] + r[:, None] * v
    x[nh:] = torch.rand((n - nh, 3), generator=g, **f64)
    x = torch.remainder(x, 1.0).to(dtype)
    one = torch.tensor(1.0, dtype=dtype, device=device)
    x = torch.where(x >= one, torch.nextafter(one, torch.zeros_like(one)), x)
    return x[torch.randperm(n, generator=g, device=device)]
