"""CenterOfMass and mass-centred trees (tree.cuh:378-429, tree.py:479-499, :736-754).

- jzt_center_of_mass bit-exact against the oracle's C restatement on random ranges (empty ranges,
  zero masses, long ranges), float32/float64, 2-D/3-D, with and without Kahan compensation;
- build_tree_hierarchy(mass_centered=True): every plane's masses and mass centres bit-exact against
  the oracle's plane recursion over the GPU's own node splits, every mass centre inside its cell;
- the reference's NaN-padding test (checks/tests/test_tree.py:135-154): padding with NaN rows and
  num_total leaves the tree, masses and mass centres unchanged.
"""
import dataclasses

import numpy as np
import pytest
import torch

from gen import clustered

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kahan", [True, False])
def test_center_of_mass_bit_exact(dev, oracle, dtype, dim, kahan):
    import jztree_amd as jz
    rng = np.random.default_rng(11 + dim)
    lens = rng.integers(0, 300, 5000)
    lens[::97] = 0
    lens[5] = 20000
    spl = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    n = int(spl[-1])
    x = rng.random((n, dim)).astype(dtype)
    w = (rng.random(n) * 2).astype(dtype)
    w[rng.random(n) < 0.01] = 0
    w[spl[7]:spl[8]] = 0  # a range of zero total mass
    print(f"[com] n={n} nodes={len(lens)} dtype={np.dtype(dtype).name} dim={dim} kahan={kahan}", flush=True)
    cg, mg = jz.tree.center_of_mass(torch.from_numpy(spl).to(dev), torch.from_numpy(x).to(dev),
                                    torch.from_numpy(w).to(dev), kahan_summation=kahan)
    co, mo = oracle.center_of_mass(spl, x, w, kahan=kahan)
    assert np.array_equal(cg.cpu().numpy(), co, equal_nan=True)
    assert np.array_equal(mg.cpu().numpy(), mo, equal_nan=True)


def _tree(dev, n, seed, pad=0):
    import jztree_amd as jz
    x = clustered(n, nhalos=64, seed=seed)
    w = (np.random.default_rng(seed).random(n) + 0.5).astype(np.float32)
    xz, idz = jz.tree.zsort(torch.from_numpy(x).to(dev))
    wz = torch.from_numpy(w).to(dev)[idz.long()]
    if pad:
        xz = torch.cat([xz, torch.full((pad, 3), float("nan"), device=dev)])
        wz = torch.cat([wz, torch.full((pad,), float("nan"), device=dev)])
    part = jz.data.PosMass(pos=xz.contiguous(), mass=wz.contiguous(), num_total=n if pad else None)
    cfg = dataclasses.replace(jz.config.TreeConfig(), mass_centered=True, alloc_fac_nodes=2.0)
    return part, jz.tree.build_tree_hierarchy(part, cfg)


def test_mass_centred_tree(dev, oracle):
    part, th = _tree(dev, 1 << 18, 3)
    isp = th.ispl_n2n
    c_prev, m_prev = oracle.center_of_mass(isp.get(0).cpu().numpy(), part.pos.cpu().numpy(), part.mass.cpu().numpy())
    for level in range(th.num_planes()):
        if level > 0:
            c_prev, m_prev = oracle.center_of_mass(isp.get(level, size=len(m_prev) + 1).cpu().numpy(), c_prev, m_prev)
        num = int(isp.num(level)) - 1
        got_c = th.mass_cent.get(level, num).cpu().numpy()
        got_m = th.mass.get(level, num).cpu().numpy()
        assert np.array_equal(got_c, c_prev[:num], equal_nan=True), f"plane {level}"
        assert np.array_equal(got_m, m_prev[:num], equal_nan=True), f"plane {level}"
        # every mass centre lies inside its node's cell, up to the float32 rounding of the centre
        # (cells of the tightest clumps are a few float32 units in the last place wide)
        cent = th.geom_cent.get(level, num).double().cpu().numpy()
        lvl = th.lvl.get(level, num).cpu().numpy()
        o = (lvl + 2000 * 3) // 3 - 2000
        r = lvl - o * 3
        half = np.stack([2.0 ** (o + (r >= 3 - i)) for i in range(3)], axis=1) * 0.5
        slack = 4 * np.spacing(np.abs(cent).astype(np.float32)).astype(np.float64)
        assert np.all(np.abs(got_c.astype(np.float64) - cent) <= half * (1 + 1e-6) + slack), f"plane {level}"
    top = th.num_planes() - 1
    assert abs(float(th.mass.get(top, int(isp.num(top)) - 1).double().sum()) -
               float(part.mass.double().sum())) <= 1e-4 * float(part.mass.double().sum())


def test_mass_centred_tree_nan_padding(dev):
    n = 50000
    _, th1 = _tree(dev, n, 4)
    _, th2 = _tree(dev, n, 4, pad=1356)
    for name in ("ispl_n2l", "ispl_n2n", "geom_cent", "lvl", "mass", "mass_cent"):
        a, b = getattr(th1, name), getattr(th2, name)
        assert torch.equal(a.ispl, b.ispl), name
        for level in range(len(a.ispl) - 1):
            num = int(a.num(level))
            va, vb = a.get(level, num).cpu(), b.get(level, num).cpu()
            if va.is_floating_point():
                va, vb = va.nan_to_num(7.0), vb.nan_to_num(7.0)
            assert torch.equal(va, vb), name
