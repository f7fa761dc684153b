"""Single point set, whole calls captured as one HIP graph (jztree_amd.graph.Captured): plain kNN,
kNN with the learned radius proposer and FoF labels, replayed on new positions copied into the
captured input, each replay equal to the eager calls on the same positions; the capture holds no
memset node."""
import dataclasses

import numpy as np
import pytest
import torch

from gen import clustered

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("boxsize", [1.0, 0.0])
def test_captured_knn_learned_fof_replays(dev, boxsize):
    import jztree_amd as jz
    n, k = 1 << 17, 32
    cfg = jz.config.KNNConfig()
    cfg = dataclasses.replace(cfg, tree=dataclasses.replace(cfg.tree, alloc_fac_nodes=2.0, max_leaf_size=64))
    fcfg = jz.config.FofConfig()
    fcfg = dataclasses.replace(fcfg, tree=dataclasses.replace(fcfg.tree, alloc_fac_nodes=2.0))
    box = boxsize or None
    train = [torch.from_numpy(clustered(n, nhalos=256, seed=s)).to(dev) for s in (301, 302)]
    feats, targets = jz.proposer.training_data(train, k, cfg, boxsize=box)
    prop = jz.proposer.RadiusProposer(k, device=dev, seed=0)
    prop.train(feats, targets, steps=200, batch=4096, lr=3e-3, lr_final=3e-4)
    # a loose quantile: some queries fall back, so the replayed retry pass has work to do
    prop.calibrate(feats, targets, quantile=0.9)
    rl = 0.2 * n ** (-1.0 / 3.0)
    inputs = [torch.from_numpy(clustered(n, nhalos=256, seed=s)).to(dev) for s in (7, 8)]

    def calls(x):
        r0, i0 = jz.knn.knn(x, k, boxsize=box, cfg=cfg)
        r1, i1, retry = jz.proposer.knn_learned(x, k, prop, boxsize=box, cfg=cfg, return_retry=True)
        _, lab = jz.fof.fof_labels(x, rl, boxsize=boxsize, cfg=fcfg, output_order="input")
        return r0, i0, r1, i1, retry, lab

    eager = []
    for x in inputs:
        out = [t.clone() for t in calls(x)]
        assert int(out[4].sum()) > 0, "no query falls back: the retry pass is not exercised"
        assert torch.equal(out[0], out[2]) and torch.equal(out[1], out[3])
        eager.append(out)
    xbuf = inputs[0].clone()
    print(f"[graph] capturing n={n} k={k} box={boxsize}", flush=True)
    g = jz.graph.Captured(lambda: calls(xbuf))
    print(f"[graph] nodes {g.node_counts}", flush=True)
    assert g.memset_nodes() == 0
    for rep, which in enumerate((0, 1, 0, 1)):
        xbuf.copy_(inputs[which])
        got = g.run()
        torch.cuda.synchronize()
        g.check()
        names = ("plain radii", "plain ids", "learned radii", "learned ids", "retry flags", "FoF labels")
        for a, b, name in zip(got, eager[which], names):
            assert torch.equal(a, b), f"replay {rep} (input {which}): {name} differ from the eager call"
        _ = torch.cumsum(inputs[1][:, 0], 0)  # eager work between replays
        torch.cuda.synchronize()
