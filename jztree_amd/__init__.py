"""jztree_amd: the tree hot path of jztree (z-order sort, tree build, exact kNN, friends-of-friends)
as hand-written HIP kernels for AMD Instinct MI355X (gfx950), behind jztree's Python API.

    import jztree_amd as jz
    posz, idz = jz.tree.zsort(pos)                 # pos: (N, 3) float32 CUDA/HIP tensor
    rnn, inn = jz.knn.knn(pos, k=16)               # exact kNN, input order
    posz, igroup = jz.fof.fof_labels(pos, rlink, boxsize=1.0)

Importing requires the built library jztree_amd/libjztree_amd.so (python jztree_amd/build.py);
there is no CPU fallback.
"""
from . import config, data, tools  # noqa: F401
from . import _lib  # noqa: F401  (loads libjztree_amd.so, raises if missing)
from . import tree, knn, fof, proposer, batch, autotune, graph  # noqa: F401

__all__ = ["config", "data", "tools", "tree", "knn", "fof", "proposer", "batch", "autotune", "graph"]
