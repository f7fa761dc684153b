"""Learned radius proposer for exact kNN (extension; no reference counterpart).

A 12-64-64-64-1 GELU MLP (bf16 weights, fp32 arithmetic; csrc/proposer.hip) predicts, per leaf,
log2 of the largest k-th neighbour distance among the leaf's particles from 12 cheap features of
the tree (leaf_features).  The prediction plus a calibrated margin becomes the starting bound of
the exact search (knn._knn_bounded): leaf lists are built for that radius instead of the
guaranteed one, the leaf kernel starts its top-k at it, and every query that finds fewer than k
neighbours inside it is recomputed without a bound.  knn_learned therefore returns exactly what
knn() returns, whatever the network predicts; a good prediction only makes it faster.

Training runs on the GPU through the library's own kernels (jzt_mlp_grad: forward + backward of
the mean squared error in log2 radius; jzt_adam_step), on radii that the plain kNN produced for
other point sets.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L
from . import checks
from .config import KNNConfig
from .data import RankIdx, TreeHierarchy, get_pos
from .tools import inverse_of_splits

NFEAT = 12
NPARAM = 9217
_OFF_B1 = NFEAT * 64
_OFF_W2 = _OFF_B1 + 64
_OFF_B2 = _OFF_W2 + 64 * 64
_OFF_W3 = _OFF_B2 + 64
_OFF_B3 = _OFF_W3 + 64 * 64
_OFF_W4 = _OFF_B3 + 64
_OFF_B4 = _OFF_W4 + 64


def _plane_parent(n2c: torch.Tensor, size: int) -> torch.Tensor:
    return torch.clamp(inverse_of_splits(n2c, size), 0, size - 1).to(torch.int32).contiguous()


def leaf_features(th: TreeHierarchy) -> torch.Tensor:
    """(th.size(), 12) float32 features of every leaf slot (zeros past the leaf count); see
    k_leaf_features for the list."""
    size = th.size()
    dev = th.lvl.data.device
    npart0 = th.npart(0, size=size).to(torch.int32).contiguous()
    lvl0 = th.lvl.get(0, size).to(torch.int32).contiguous()
    nleaves = th.num(0).to(torch.int32).reshape(1).contiguous()
    if th.num_planes() >= 2:
        n2c1 = th.ispl_n2n.get(1, size + 1).to(torch.int32).contiguous()
        par1 = _plane_parent(n2c1, size)
        npart1 = th.npart(1, size=size).to(torch.int32).contiguous()
        lvl1 = th.lvl.get(1, size).to(torch.int32).contiguous()
    else:  # a single plane: each leaf is its own parent
        n2c1 = torch.arange(size + 1, dtype=torch.int32, device=dev)
        par1 = torch.arange(size, dtype=torch.int32, device=dev)
        npart1, lvl1 = npart0, lvl0
    par2 = npart2 = lvl2 = None
    if th.num_planes() >= 3:
        par2 = _plane_parent(th.ispl_n2n.get(2, size + 1).to(torch.int32).contiguous(), size)
        npart2 = th.npart(2, size=size).to(torch.int32).contiguous()
        lvl2 = th.lvl.get(2, size).to(torch.int32).contiguous()
    feat = torch.empty((size, NFEAT), dtype=torch.float32, device=dev)
    L.call("jzt_leaf_features", L.stream(), L.ptr(npart0), L.ptr(lvl0), L.ptr(par1), L.ptr(npart1), L.ptr(lvl1),
           L.ptr(n2c1), L.ptr(par2), L.ptr(npart2), L.ptr(lvl2), L.ptr(nleaves), size, L.ptr(feat))
    return feat


def leaf_log2_radius(th: TreeHierarchy, rz: torch.Tensor, k: int) -> torch.Tensor:
    """Training target per leaf slot: log2 of the largest k-th neighbour distance among the leaf's
    particles (rz: z-ordered radii rows of the plain kNN).  Slots past the leaf count: -inf."""
    n = rz.shape[0]
    nl = int(th.num(0))
    spl = th.splits_leaf_to_part()[: nl + 1].contiguous()
    leafof = inverse_of_splits(spl, n).long()
    rk = rz[:, k - 1].double().clamp(min=2.0 ** -60)
    out = torch.full((th.size(),), -math.inf, dtype=torch.float64, device=rz.device)
    out[:nl] = 0.0
    out = out.scatter_reduce(0, leafof, rk, "amax", include_self=False)
    return torch.log2(out).float()


@dataclass
class TrainReport:
    steps: int
    loss_first: float
    loss_last: float
    losses: list


class RadiusProposer:
    """The network, its normalisation and margin, and Adam state.  Weights live on `device`:
    fp32 master copy `w` and the bf16 weights `wq` that every forward pass uses."""

    def __init__(self, k: int, device="cuda", seed: int = 0):
        self.k = int(k)
        self.device = torch.device(device)
        g = torch.Generator().manual_seed(seed)
        w = torch.zeros(NPARAM, dtype=torch.float32)
        w[:_OFF_B1] = torch.randn(NFEAT * 64, generator=g) * math.sqrt(2.0 / NFEAT)
        w[_OFF_W2:_OFF_B2] = torch.randn(64 * 64, generator=g) * math.sqrt(2.0 / 64)
        w[_OFF_W3:_OFF_B3] = torch.randn(64 * 64, generator=g) * math.sqrt(2.0 / 64)
        w[_OFF_W4:_OFF_B4] = torch.randn(64, generator=g) * math.sqrt(1.0 / 64)
        self.w = w.to(self.device)
        self.wq = torch.empty(NPARAM, dtype=torch.bfloat16, device=self.device)
        L.call("jzt_to_bf16", L.stream(), L.ptr(self.w), L.ptr(self.wq), NPARAM)
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)
        self.step = 0
        self.norm = torch.cat([torch.zeros(NFEAT), torch.ones(NFEAT), torch.tensor([0.0, 1.0])]).to(self.device)
        self.norm_fitted = False
        self.margin = 0.0

    # -- normalisation ---------------------------------------------------------------------------
    def fit_normalisation(self, feats: torch.Tensor, targets: torch.Tensor):
        f = feats.double()
        fm, fs = f.mean(0), f.std(0).clamp(min=1e-3)
        t = targets.double()
        self.norm = torch.cat([fm, fs, torch.stack([t.mean(), t.std().clamp(min=1e-3)])]).float().to(self.device)
        self.norm_fitted = True

    def _standardise(self, feats, targets=None):
        x = ((feats.to(self.device).float() - self.norm[:NFEAT]) / self.norm[NFEAT:2 * NFEAT]).contiguous()
        if targets is None:
            return x
        t = ((targets.to(self.device).float() - self.norm[2 * NFEAT]) / self.norm[2 * NFEAT + 1]).contiguous()
        return x, t

    # -- training --------------------------------------------------------------------------------
    def grad(self, x: torch.Tensor, t: torch.Tensor):
        """(loss, gradient) of the mean squared error on standardised (x, t), from jzt_mlp_grad."""
        grad = torch.empty(NPARAM, dtype=torch.float32, device=self.device)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        L.call("jzt_mlp_grad", L.stream(), L.ptr(self.wq), L.ptr(x.contiguous()), L.ptr(t.contiguous()), len(t),
               L.ptr(grad), L.ptr(loss))
        return loss, grad

    def adam(self, grad: torch.Tensor, lr: float, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        L.call("jzt_adam_step", L.stream(), L.ptr(self.w), L.ptr(self.wq), L.ptr(self.m), L.ptr(self.v),
               L.ptr(grad), NPARAM, float(lr), float(beta1), float(beta2), float(eps), self.step)

    def train(self, feats: torch.Tensor, targets: torch.Tensor, steps: int = 3000, batch: int = 8192,
              lr: float = 2e-3, lr_final: float = 1e-4, seed: int = 0, log_every: int = 100) -> TrainReport:
        """Adam on random mini-batches of (feats, targets) (rows = leaves), learning rate decayed
        geometrically from lr to lr_final.  Fits the normalisation first if it is still neutral."""
        if not self.norm_fitted:
            self.fit_normalisation(feats, targets)
        x, t = self._standardise(feats, targets)
        n = len(t)
        gen = torch.Generator(device=self.device).manual_seed(seed)
        losses = []
        for i in range(steps):
            idx = torch.randint(0, n, (min(batch, n),), generator=gen, device=self.device)
            loss, grad = self.grad(x[idx], t[idx])
            self.adam(grad, lr * (lr_final / lr) ** (i / max(steps - 1, 1)))
            if i % log_every == 0 or i == steps - 1:
                losses.append(loss)
        losses = [float(v) for v in torch.cat(losses).cpu()]
        return TrainReport(steps, losses[0], losses[-1], losses)

    # -- inference -------------------------------------------------------------------------------
    def predict_log2r(self, feats: torch.Tensor) -> torch.Tensor:
        feats = feats.to(self.device).float().contiguous()
        out = torch.empty(len(feats), dtype=torch.float32, device=self.device)
        L.call("jzt_mlp_infer", L.stream(), L.ptr(self.wq), L.ptr(self.norm), L.ptr(feats), None, len(feats), 0.0,
               None, L.ptr(out))
        return out

    def calibrate(self, feats: torch.Tensor, targets: torch.Tensor, quantile: float = 0.98) -> float:
        """margin = the `quantile` of (target - prediction) over the given leaves, so that the bound
        covers about that fraction of them."""
        res = (targets.to(self.device).double() - self.predict_log2r(feats).double()).cpu().numpy()
        self.margin = float(np.quantile(res, quantile))
        return self.margin

    def bound_r2(self, th: TreeHierarchy, feats: torch.Tensor | None = None) -> torch.Tensor:
        """Starting bound r2_0 per leaf slot: (2^(prediction + margin))^2, 0 for empty slots."""
        if feats is None:
            feats = leaf_features(th)
        size = th.size()
        npart0 = th.npart(0, size=size).to(torch.int32).contiguous()
        out = torch.empty(size, dtype=torch.float32, device=self.device)
        L.call("jzt_mlp_infer", L.stream(), L.ptr(self.wq), L.ptr(self.norm), L.ptr(feats.contiguous()),
               L.ptr(npart0), size, float(self.margin), L.ptr(out), None)
        return out

    # -- persistence -----------------------------------------------------------------------------
    def state(self) -> dict:
        return {"k": np.array(self.k), "w": self.w.cpu().numpy(), "norm": self.norm.cpu().numpy(),
                "margin": np.array(self.margin)}

    def save(self, path: str):
        np.savez(path, **self.state())

    @classmethod
    def load(cls, path: str, device="cuda") -> "RadiusProposer":
        d = np.load(path, allow_pickle=False)
        p = cls(int(d["k"]), device=device)
        p.w = torch.from_numpy(d["w"].astype(np.float32)).to(p.device)
        L.call("jzt_to_bf16", L.stream(), L.ptr(p.w), L.ptr(p.wq), NPARAM)
        p.norm = torch.from_numpy(d["norm"].astype(np.float32)).to(p.device)
        p.norm_fitted = True
        p.margin = float(d["margin"])
        return p


def training_data(points: list, k: int, cfg: KNNConfig = KNNConfig(), boxsize: float | None = None):
    """(features, log2 radius targets) of the leaves of each point set, from the plain kNN."""
    from .knn import _knn_dual_walk, _knn_leaf2leaf
    from .tree import zsort_and_tree_origin
    fs, ts = [], []
    for x in points:
        dev = x.device
        partz, originz, th = zsort_and_tree_origin(x, cfg.tree)
        il = _knn_dual_walk(th, k, boxsize=boxsize, alloc_fac_ilist=cfg.alloc_fac_ilist)
        rz, _ = _knn_leaf2leaf(il, th.splits_leaf_to_part(), get_pos(partz), k=k, boxsize=boxsize or 0.,
                               src_ids=originz.idx)
        checks.flush()
        nl = int(th.num(0))
        fs.append(leaf_features(th)[:nl])
        ts.append(leaf_log2_radius(th, rz, k)[:nl])
    return torch.cat(fs), torch.cat(ts)


def knn_learned(part, k: int, proposer: RadiusProposer, boxsize: float | None = None,
                cfg: KNNConfig = KNNConfig(), output_order: str = "input", return_retry: bool = False,
                dist_dtype: torch.dtype | None = None):
    """Exact kNN (radii, global indices) with the proposer's starting bound; identical to
    knn(part, k, boxsize, output_order=...).  With return_retry, also the per-query flags of the
    queries that the first (proposed-radius) round could not finish (z order).  dist_dtype as in
    knn(): float32 distances on float64 positions."""
    from .knn import _distance_view, _knn_bounded
    from .tree import zsort_and_tree_origin
    assert output_order in ("z", "input")
    assert k == proposer.k, "the proposer was trained for another k"
    pos = get_pos(part)
    dev = pos.device
    size = len(pos)
    try:
        partz, originz, th = zsort_and_tree_origin(part, cfg.tree)
        th, posz = _distance_view(th, get_pos(partz), dist_dtype)
        bound = proposer.bound_r2(th)
        if output_order == "input":  # rows and reported ids in input order, as knn()
            sid, dst = originz.idx, originz.idx
        else:  # z order: rows and ids are z-indices
            sid, dst = None, None
        rnn, inn, retry = _knn_bounded(th, posz, k, bound, boxsize=boxsize,
                                       alloc_fac_ilist=cfg.alloc_fac_ilist, src_ids=sid, dst_rows=dst)
        checks.flush()
    except BaseException:
        checks.clear()
        raise
    if return_retry:
        return rnn, inn, retry
    return rnn, inn
