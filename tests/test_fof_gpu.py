"""FoF on the GPU vs the oracle: labels bit-exact in z-order (min z-index per group) and after
canonicalisation (min input index per group)."""
import numpy as np
import pytest
import torch

from conftest import COSMO2D_LABELS40, COSMO2D_NGROUPS_GE10, cosmo2d
from gen import clustered, uniform

pytestmark = pytest.mark.gpu


def _cfg():
    """Strongly clustered sets need a larger leaf allocation than the FofConfig default (1.1)."""
    import dataclasses
    from jztree_amd.config import FofConfig
    c = FofConfig()
    return dataclasses.replace(c, tree=dataclasses.replace(c.tree, alloc_fac_nodes=2.0))


def _fof(x, rlink, dev, boxsize=0.0, output_order="z"):
    import jztree_amd as jz
    print(f"[fof] n={len(x)} dim={x.shape[1]} rlink={rlink} box={boxsize} {output_order}", flush=True)
    t = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    pz, ig = jz.fof.fof_labels(t, rlink, boxsize=boxsize, output_order=output_order, cfg=_cfg())
    torch.cuda.synchronize()
    return pz.cpu().numpy(), ig.cpu().numpy()


def test_fof_cosmo2d_known_answer(dev):
    """docs/quickstart.md:115-131 and :219-251 of the reference."""
    pos = cosmo2d()
    pz, ig = _fof(pos, 0.2 * 100. / 128, dev, boxsize=100.0)
    assert list(ig[:40]) == COSMO2D_LABELS40
    assert np.sum(np.bincount(ig, minlength=len(ig)) >= 10) == COSMO2D_NGROUPS_GE10


@pytest.mark.parametrize("case", ["uniform", "clustered", "periodic", "dense"])
def test_fof_matches_oracle(dev, oracle, case):
    if case == "uniform":
        x, b, box = uniform(1 << 16, seed=1), 0.7, 0.0
    elif case == "clustered":
        x, b, box = clustered(1 << 17, nhalos=256, seed=2), 0.2, 0.0
    elif case == "periodic":
        x, b, box = clustered(1 << 17, nhalos=256, seed=3), 0.2, 1.0
    else:
        x, b, box = uniform(1 << 15, seed=4), 1.2, 1.0
    rl = b * len(x) ** (-1.0 / 3.0)
    pz, ig = _fof(x, rl, dev, boxsize=box)
    opz, oidz = oracle.zsort(x)
    assert np.array_equal(pz, opz)
    ref_z = oracle.fof(opz, rl, boxsize=box, key=np.arange(len(x)))
    assert np.array_equal(ig.astype(np.int64), ref_z), f"{np.sum(ig != ref_z)} labels differ"


def test_fof_input_order_and_canonical(dev, oracle):
    """output_order='input' gives the input index of each group's first z-order member
    (fof.py:810-813); canonicalised it equals the oracle's minimum input index per group."""
    x = clustered(1 << 16, nhalos=128, seed=5)
    rl = 0.2 * len(x) ** (-1.0 / 3.0)
    _, ig = _fof(x, rl, dev, boxsize=1.0, output_order="input")
    canon = oracle.canonical_labels(ig)
    ref = oracle.fof(x, rl, boxsize=1.0)
    assert np.array_equal(canon, ref)
    _, idz = oracle.zsort(x)
    zpos = np.empty(len(x), np.int64)
    zpos[idz] = np.arange(len(x))
    assert np.all(zpos[ig] <= zpos)  # label points at the group member that comes first in z-order


@pytest.mark.gpu
@pytest.mark.parametrize("boxsize", [0.0, 1.0])
def test_fof_float64(dev, boxsize):
    """FoF on float64 positions (sort, tree and FoF kernels in double): the partition equals float64
    brute-force connected components of r2 <= float32(rlink^2) (random points: no pair lies within
    an ulp of the threshold)."""
    import numpy as np
    import torch
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    import jztree_amd as jz
    rng = np.random.default_rng(31)
    x = rng.random((6000, 3))
    rl = 0.7 * len(x) ** (-1 / 3)
    print(f"[fof f64] box={boxsize}", flush=True)
    _, lab = jz.fof.fof_labels(torch.from_numpy(x).to(dev), rl, boxsize=boxsize, output_order="input")
    lab = lab.cpu().numpy()
    d = x[:, None, :] - x[None, :, :]
    if boxsize:
        d = d - boxsize * np.round(d / boxsize)
    d2 = (d * d).sum(-1)
    i, j = np.nonzero(d2 <= float(np.float32(rl * rl)))
    ncomp, comp = connected_components(coo_matrix((np.ones(len(i)), (i, j)), shape=(len(x), len(x))), directed=False)
    # same partition: canonical label = smallest member index
    def canon(l):
        first = {}
        return np.array([first.setdefault(v, k) for k, v in enumerate(l)])
    assert np.array_equal(canon(lab), canon(comp))
    assert ncomp < len(x)
