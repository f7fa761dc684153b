"""Learned radius proposer: oracle pins (CPU), kernels against the oracle and end-to-end exactness
(GPU)."""
import numpy as np
import pytest
import torch

from gen import clustered
from oracle import proposer_oracle as P


# ------------------------------------------------------------------------------------------------
# oracle pins (CPU)
# ------------------------------------------------------------------------------------------------

def test_oracle_grad_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(size=P.NPARAM) * 0.3
    x = rng.normal(size=(37, P.NF))
    t = rng.normal(size=37)
    loss, g = P.loss_and_grad(p, x, t)
    assert np.isclose(loss, np.mean((P.forward(p, x) - t) ** 2))
    idx = np.concatenate([rng.choice(P.NPARAM, 300, replace=False),
                          [P.OFF_B1, P.OFF_W2, P.OFF_B2, P.OFF_W3, P.OFF_B3, P.OFF_W4, P.OFF_B4, P.NPARAM - 2]])
    h = 1e-6
    for i in idx:
        e = np.zeros(P.NPARAM)
        e[i] = h
        fd = (P.loss_and_grad(p + e, x, t)[0] - P.loss_and_grad(p - e, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_gelu_derivative_and_values():
    z = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (P.gelu(z + h) - P.gelu(z - h)) / (2 * h)
    assert np.allclose(P.gelu_grad(z), fd, atol=1e-8)
    assert P.gelu(0.0) == 0.0 and P.gelu_grad(0.0) == 0.5
    assert np.isclose(P.gelu(10.0), 10.0) and abs(P.gelu(-10.0)) < 1e-20


def test_bf16_round_halfway_cases():
    f = np.float32
    cases = {
        1.0: 1.0,
        1.0 + 2 ** -8: 1.0,                   # halfway between 1 and 1+2^-7: ties to even (1)
        1.0 + 3 * 2 ** -8: 1.0 + 2 ** -6,     # halfway between 1+2^-7 (odd) and 1+2^-6 (even)
        1.0 + 2 ** -8 + 2 ** -20: 1.0 + 2 ** -7,
        -1.5: -1.5,
    }
    bits = lambda b: np.array([b], np.uint32).view(np.float32)[0]
    cases[bits(0x7F61B1E6)] = bits(0x7F620000)  # low half 0xB1E6 > 0x8000: rounds up
    cases[bits(0x3F818000)] = bits(0x3F820000)  # exact half, odd (0x3F81) -> even
    cases[bits(0x3F828000)] = bits(0x3F820000)  # exact half, even stays
    for a, b in cases.items():
        assert P.bf16_round(np.array([a], f))[0] == f(b), (a, P.bf16_round(np.array([a], f))[0], b)
    r = P.bf16_round(np.array([np.inf, -np.inf, np.nan], f))
    assert r[0] == np.inf and r[1] == -np.inf and np.isnan(r[2])
    # every bf16 value is a fixed point, and rounding is monotone
    v = (np.arange(0, 0x7F80, 7, dtype=np.uint32) << 16).view(np.float32)
    assert np.array_equal(P.bf16_round(v), v)
    w = np.sort(np.random.default_rng(1).normal(size=5000).astype(f))
    assert np.all(np.diff(P.bf16_round(w)) >= 0)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(2)
    w0 = rng.normal(size=100)
    g = rng.normal(size=100)
    w1, m, v = P.adam_step(w0, np.zeros(100), np.zeros(100), g, step=1, lr=0.01)
    assert np.allclose(w1, w0 - 0.01 * np.sign(g), atol=1e-9)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_leaf_features_known_answer():
    # 4 leaves in 2 plane-1 parents (no plane 2), 2 empty slots
    npart0 = np.array([1, 2, 4, 8, 0, 0])
    lvl0 = np.array([-30, -33, -36, -36, 0, 0])
    par1 = np.array([0, 0, 1, 1, 5, 5])
    npart1 = np.array([3, 12, 0, 0, 0, 0])
    lvl1 = np.array([-27, -30, 0, 0, 0, 0])
    n2c1 = np.array([0, 2, 4, 4, 4, 4, 4])
    f = P.leaf_features(npart0, lvl0, par1, npart1, lvl1, n2c1, None, None, None, nleaves=4, size=6)
    l3 = np.log2(3)
    assert np.allclose(f[0], [0, -10, l3, -9, l3, -9, 0, 1, 0, -1, 1, np.log2(7)])
    assert np.allclose(f[2], [2, -12, np.log2(12), -10, np.log2(12), -10, 1, 3, 1, 0, 1, np.log2(15)])
    assert np.allclose(f[3], [3, -12, np.log2(12), -10, np.log2(12), -10, 2, 3, 0, 0, 1, np.log2(14)])
    assert np.all(f[4:] == 0)


# ------------------------------------------------------------------------------------------------
# kernels against the oracle (GPU)
# ------------------------------------------------------------------------------------------------

def _tree(dev, n=1 << 16, seed=3, leaf=32):
    import dataclasses
    import jztree_amd as jz
    x = torch.from_numpy(clustered(n, nhalos=64, seed=seed)).to(dev)
    cfg = jz.config.KNNConfig()
    cfg = dataclasses.replace(cfg, tree=dataclasses.replace(cfg.tree, alloc_fac_nodes=2.0, max_leaf_size=leaf))
    partz, originz, th = jz.tree.zsort_and_tree(x, cfg.tree, data=jz.data.RankIdx(
        rank=None, idx=torch.arange(n, dtype=torch.int32, device=dev)))
    return x, cfg, partz, originz, th


@pytest.mark.gpu
@pytest.mark.parametrize("n,leaf", [(1 << 19, 16), (1 << 17, 32), (3000, 64)])
def test_leaf_features_vs_oracle(dev, n, leaf):
    import jztree_amd as jz
    from jztree_amd.proposer import _plane_parent
    _, _, _, _, th = _tree(dev, n=n, leaf=leaf)
    nplanes = th.num_planes()
    print(f"[features] n={n} leaf={leaf} planes={nplanes}", flush=True)
    f = jz.proposer.leaf_features(th).cpu().numpy()
    size = th.size()
    g = lambda t: t.cpu().numpy().astype(np.int64)
    npart0, lvl0 = g(th.npart(0, size=size)), g(th.lvl.get(0, size))
    if nplanes >= 2:
        n2c1 = th.ispl_n2n.get(1, size + 1).to(torch.int32)
        par1, npart1, lvl1, n2c1 = g(_plane_parent(n2c1, size)), g(th.npart(1, size=size)), g(th.lvl.get(1, size)), g(n2c1)
    else:
        par1, npart1, lvl1, n2c1 = np.arange(size), npart0, lvl0, np.arange(size + 1)
    par2 = npart2 = lvl2 = None
    if nplanes >= 3:
        par2 = g(_plane_parent(th.ispl_n2n.get(2, size + 1).to(torch.int32), size))
        npart2, lvl2 = g(th.npart(2, size=size)), g(th.lvl.get(2, size))
    ref = P.leaf_features(npart0, lvl0, par1, npart1, lvl1, n2c1, par2, npart2, lvl2, nleaves=int(th.num(0)),
                          size=size)
    assert np.allclose(f, ref, rtol=1e-6, atol=1e-6)


@pytest.mark.gpu
def test_mlp_infer_vs_oracle(dev):
    import jztree_amd as jz
    prop = jz.proposer.RadiusProposer(k=16, device=dev, seed=4)
    rng = np.random.default_rng(4)
    feats = torch.from_numpy(rng.normal(size=(1000, 12)).astype(np.float32) * 3 + 1).to(dev)
    prop.norm = torch.from_numpy(np.concatenate([np.ones(12), 3 * np.ones(12), [-7.0, 2.0]]).astype(np.float32)).to(dev)
    y = prop.predict_log2r(feats).cpu().numpy()
    wq = prop.wq.float().cpu().numpy()
    assert np.array_equal(wq, P.bf16_round(prop.w.cpu().numpy()))
    x = (feats.cpu().numpy().astype(np.float64) - 1) / 3
    ref = P.forward(wq, x) * 2.0 - 7.0
    assert np.allclose(y, ref, rtol=1e-4, atol=1e-4), np.max(np.abs(y - ref))


@pytest.mark.gpu
@pytest.mark.parametrize("batch", [1, 64, 200, 4096])
def test_mlp_grad_vs_oracle(dev, batch):
    import jztree_amd as jz
    prop = jz.proposer.RadiusProposer(k=16, device=dev, seed=batch)
    rng = np.random.default_rng(batch)
    x = rng.normal(size=(batch, 12)).astype(np.float32)
    t = rng.normal(size=batch).astype(np.float32)
    loss, grad = prop.grad(torch.from_numpy(x).to(dev), torch.from_numpy(t).to(dev))
    wq = prop.wq.float().cpu().numpy()
    lref, gref = P.loss_and_grad(wq, x, t)
    assert np.isclose(float(loss), lref, rtol=1e-4)
    g = grad.cpu().numpy()
    scale = np.abs(gref).max()
    assert np.allclose(g, gref, rtol=1e-3, atol=1e-4 * scale), np.max(np.abs(g - gref)) / scale


@pytest.mark.gpu
def test_adam_vs_oracle(dev):
    import jztree_amd as jz
    prop = jz.proposer.RadiusProposer(k=16, device=dev, seed=5)
    rng = np.random.default_rng(5)
    w = prop.w.cpu().numpy().astype(np.float64)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    for step in range(1, 4):
        g = rng.normal(size=w.shape).astype(np.float32)
        prop.adam(torch.from_numpy(g).to(dev), lr=1e-2)
        w, m, v = P.adam_step(w, m, v, g, step=step, lr=1e-2)
    assert np.allclose(prop.w.cpu().numpy(), w, rtol=1e-5, atol=1e-6)
    assert np.array_equal(prop.wq.float().cpu().numpy(), P.bf16_round(prop.w.cpu().numpy()))


# ------------------------------------------------------------------------------------------------
# training and the learned search (GPU)
# ------------------------------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("k,boxsize", [(32, 1.0), (16, 0.0)])
def test_trained_proposer_exact_and_saves_work(dev, k, boxsize):
    """Train on 4 seeds, then on a fifth: identical results to the plain search, few fallbacks,
    and much shorter leaf lists than the guaranteed-radius lists of the plain search."""
    import dataclasses
    import jztree_amd as jz
    n = 1 << 17
    cfg = jz.config.KNNConfig()
    cfg = dataclasses.replace(cfg, tree=dataclasses.replace(cfg.tree, alloc_fac_nodes=2.0, max_leaf_size=64))
    train = [torch.from_numpy(clustered(n, nhalos=256, seed=s)).to(dev) for s in (101, 102, 103, 104)]
    feats, targets = jz.proposer.training_data(train, k, cfg, boxsize=boxsize or None)
    prop = jz.proposer.RadiusProposer(k, device=dev, seed=0)
    rep = prop.train(feats, targets, steps=400, batch=4096, lr=3e-3, lr_final=3e-4)
    print(f"[proposer] loss {rep.loss_first:.4f} -> {rep.loss_last:.4f}", flush=True)
    assert rep.loss_last < 0.5 * rep.loss_first
    prop.calibrate(feats, targets, quantile=0.98)

    x = torch.from_numpy(clustered(n, nhalos=256, seed=7)).to(dev)
    r0, i0 = jz.knn.knn(x, k, boxsize=boxsize or None, cfg=cfg)
    r1, i1, retry = jz.proposer.knn_learned(x, k, prop, boxsize=boxsize or None, cfg=cfg, return_retry=True)
    torch.cuda.synchronize()
    assert torch.equal(i0, i1) and torch.equal(r0, r1)
    rate = float(retry.float().mean())
    print(f"[proposer] fallback rate {rate:.4f}", flush=True)
    assert rate < 0.1

    # work: leaf-list entries with the proposed radius vs the guaranteed radius
    partz, originz, th = jz.tree.zsort_and_tree(x, cfg.tree, data=jz.data.RankIdx(
        rank=None, idx=torch.arange(n, dtype=torch.int32, device=dev)))
    plain = jz.knn._knn_dual_walk(th, k, boxsize=boxsize or None, alloc_fac_ilist=cfg.alloc_fac_ilist)
    learned = jz.knn._knn_radius_walk(th, k, jz.knn._plane_radii(th, prop.bound_r2(th)), boxsize=boxsize,
                                      alloc_fac_ilist=cfg.alloc_fac_ilist)
    ne_plain, ne_learned = int(plain.ispl[-1]), int(learned.ispl[-1])
    print(f"[proposer] leaf-list entries plain {ne_plain} learned {ne_learned}", flush=True)
    assert ne_learned < 0.6 * ne_plain
