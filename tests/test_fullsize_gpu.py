"""Full configuration sizes (2^24 clustered points, periodic, kNN k=32 and FoF b=0.2), checked
through properties that do not depend on the size: a float64 brute force over all points for a
sample of queries (kNN) and of particles (FoF), on the GPU."""
import dataclasses

import numpy as np
import pytest
import torch

from gen import clustered

pytestmark = pytest.mark.gpu

N = 1 << 24
BOX = 1.0


@pytest.fixture(scope="module")
def points(dev):
    x = clustered(N, nhalos=4096, seed=5)
    return x, torch.from_numpy(x).to(dev)


def _wrap(d):
    return d - BOX * torch.round(d / BOX)


def _d2_all(xall64, q64):
    """(nq, N) float64 squared periodic distances, in query chunks."""
    return torch.cat([(_wrap(xall64[None, :, :] - q64[i:i + 8, None, :]) ** 2).sum(-1)
                      for i in range(0, len(q64), 8)])


def test_knn_2e24_against_brute_force(dev, points):
    import jztree_amd as jz
    x, t = points
    k = 32
    cfg = jz.config.KNNConfig()
    cfg = dataclasses.replace(cfg, tree=dataclasses.replace(cfg.tree, alloc_fac_nodes=2.0, max_leaf_size=128))
    print(f"[full] knn n={N} k={k}", flush=True)
    rnn, inn = jz.knn.knn(t, k=k, boxsize=BOX, cfg=cfg)
    torch.cuda.synchronize()
    rng = np.random.default_rng(0)
    qi = torch.from_numpy(rng.choice(N, 128, replace=False)).to(dev)
    x64 = t.double()
    d2 = _d2_all(x64, x64[qi])                                   # (128, N)
    r = rnn[qi].double()                                         # (128, k)
    ids = inn[qi].long()
    # radii: within 2 ulp of the float64 distance to the reported neighbour
    dn = torch.sqrt(d2.gather(1, ids))
    ulp = torch.from_numpy(np.spacing(rnn[qi].cpu().numpy())).to(dev).double()
    assert torch.all((dn - r).abs() <= 2 * ulp), "radius differs from the float64 distance"
    # completeness: every point clearly inside the k-th radius is reported, and no more than k are
    rk = r[:, -1:]
    inside = d2 < (rk * (1 - 1e-6)) ** 2
    assert torch.all(inside.sum(1) <= k)
    rep = torch.zeros_like(inside)
    rep.scatter_(1, ids, True)
    assert torch.all(~inside | rep), "a point inside the k-th radius is missing"
    # ascending radii, distinct ids per row
    assert torch.all(r[:, 1:] >= r[:, :-1])
    assert torch.all(ids.sort(1).values[:, 1:] != ids.sort(1).values[:, :-1])


def test_fof_2e24_properties(dev, points):
    import jztree_amd as jz
    x, t = points
    rl = 0.2 * N ** (-1.0 / 3.0)
    cfg = jz.config.FofConfig()
    cfg = dataclasses.replace(cfg, tree=dataclasses.replace(cfg.tree, alloc_fac_nodes=2.0))
    print(f"[full] fof n={N} rlink={rl}", flush=True)
    _, lab = jz.fof.fof_labels(t, rl, boxsize=BOX, cfg=cfg, output_order="input")
    torch.cuda.synchronize()
    lab = lab.long()
    # labels name a member of the group, which is its own label
    assert torch.all(lab[lab] == lab)
    # sampled particles: everything within the linking length shares the label (float32 canonical
    # r2 vs float64 here: pairs within 1e-6 of the threshold are left out of the check)
    rng = np.random.default_rng(1)
    pi = torch.from_numpy(rng.choice(N, 128, replace=False)).to(dev)
    d2 = _d2_all(t.double(), t.double()[pi])
    r2link = float(np.float32(rl * rl))
    near = d2 <= r2link * (1 - 1e-6)
    same = lab[None, :] == lab[pi][:, None]
    assert torch.all(~near | same), "a particle within the linking length is in another group"
    # a group is connected: a sampled particle with a groupmate has another particle within the
    # linking length (near counts the particle itself)
    alone = (lab == lab[pi][:, None]).sum(1) == 1
    has_friend = (d2 <= r2link * (1 + 1e-6)).sum(1) > 1
    assert torch.all(alone | has_friend), "a particle in a group has no friend within the linking length"
