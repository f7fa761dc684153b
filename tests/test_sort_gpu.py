"""Z-order sort on the GPU vs the oracle (bit-exact permutation)."""
import numpy as np
import pytest
import torch

from conftest import ZSORT_GOLDEN_IN, ZSORT_GOLDEN_OUT
from gen import clustered, uniform

pytestmark = pytest.mark.gpu


def _gpu_zsort(x, dev):
    import jztree_amd as jz
    print(f"[zsort] n={len(x)} dim={x.shape[1]} dtype={x.dtype}", flush=True)
    t = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    pz, idz = jz.tree.zsort(t)
    torch.cuda.synchronize()
    return pz.cpu().numpy(), idz.cpu().numpy()


def test_zsort_golden(dev):
    """checks/tests/test_tree.py:85-108"""
    xin = np.array(ZSORT_GOLDEN_IN, np.float32)
    pz, idz = _gpu_zsort(xin, dev)
    assert np.array_equal(pz[:-1], np.array(ZSORT_GOLDEN_OUT, np.float32)[:-1])
    assert np.all(np.isnan(pz[-1]))


@pytest.mark.parametrize("case", ["uniform16", "clustered", "dups", "negative", "specials", "2d", "wide",
                                  "f64", "tiny", "one"])
def test_zsort_matches_oracle(dev, oracle, case):
    rng = np.random.default_rng(3)
    if case == "uniform16":
        x = uniform(1 << 16, seed=0)
    elif case == "clustered":
        x = clustered(1 << 17, nhalos=256, seed=1)
    elif case == "dups":
        x = uniform(20000, seed=2)
        x = np.concatenate([x, x[:5000], x[:3000]])
        x[::7, 1] = x[0, 1]
        x[::11, 0] = 0.0
        x[::13, 2] = np.float32(1e-30)
    elif case == "negative":
        x = (rng.random((50000, 3)) * 4 - 2).astype(np.float32)
    elif case == "specials":
        x = (rng.random((5000, 3)) * 2 - 1).astype(np.float32)
        x[::9, 0] = np.inf
        x[::10, 1] = -np.inf
        x[::15, 2] = np.nan
        x[::23, 2] = np.float32(1e-42)
        x[::29, 1] = -0.0
    elif case == "2d":
        x = uniform(40000, dim=2, seed=4)
    elif case == "wide":
        x = uniform(40000, seed=5, lo=0.0, hi=1e6)
    elif case == "f64":
        x = rng.random((30000, 3))
    elif case == "tiny":
        x = uniform(5, seed=6)
    else:
        x = uniform(1, seed=7)
    pz, idz = _gpu_zsort(x, dev)
    opz, oidz = oracle.zsort(x)
    assert np.array_equal(idz.astype(np.int64), oidz)
    assert np.array_equal(pz, x[oidz], equal_nan=True)


def test_zsort_large_is_permutation_and_sorted(dev, oracle):
    """Full-size property checks: a bijection, and adjacent pairs never inverted."""
    x = clustered(1 << 22, nhalos=4096, seed=9)
    pz, idz = _gpu_zsort(x, dev)
    assert np.array_equal(np.sort(idz), np.arange(len(x)))
    assert np.array_equal(pz, x[idz])
    sample = np.random.default_rng(0).integers(0, len(x) - 1, 20000)
    assert all(oracle.z_exact_cmp(pz[i], pz[i + 1]) <= 0 for i in sample)


def test_zsort_gradient(dev):
    """custom VJP (tree.py:111-138): the gradient scatters back through the permutation."""
    import jztree_amd as jz
    x = torch.from_numpy(uniform(1000, seed=8)).to(dev).requires_grad_(True)
    pz, idz = jz.tree.zsort(x)
    w = torch.arange(pz.numel(), dtype=torch.float32, device=dev).reshape(pz.shape)
    (pz * w).sum().backward()
    g = torch.empty_like(w)
    g[idz.long()] = w
    assert torch.equal(x.grad, g)


@pytest.mark.gpu
@pytest.mark.parametrize("kpt", [8, 16])
def test_zsort_tile_sizes(dev, oracle, kpt):
    """Every onesweep tile (keys per thread) sorts exactly as the default one and the oracle."""
    import jztree_amd as jz
    from gen import clustered
    x = clustered(200_003, nhalos=50, seed=9)
    x[:1000] = x[1000:2000]  # equal keys
    t = torch.from_numpy(x).to(dev)
    before = jz.autotune.get_knobs()
    try:
        jz.autotune.set_knobs(sort_kpt=kpt)
        print(f"[zsort tile] kpt={kpt}", flush=True)
        pz, iz = jz.tree.zsort(t)
        torch.cuda.synchronize()
    finally:
        jz.autotune.set_knobs(**before)
    p0, i0 = jz.tree.zsort(t)
    assert torch.equal(iz, i0) and torch.equal(pz, p0)
    assert np.array_equal(iz.cpu().numpy().astype(np.int64), oracle.zsort(x)[1])
