"""Static settings of the tree, kNN and FoF calls.

The class names, field names, defaults and nesting are the public settings API of the reference
(src/jztree/config.py:1-64), kept identical so that existing callers construct them unchanged.
"""
from dataclasses import dataclass, field


@dataclass(unsafe_hash=True)
class LoggingConfig():
    """Verbosity of the library's diagnostic messages and whether they carry the call site."""
    level: int = 1
    show_loc: bool = True


@dataclass(unsafe_hash=True)
class RegularizationConfig():
    """Extra leaf splits for elongated leaves: leaves whose cell volume exceeds max_volume_fac
    times the volume at the given percentile of leaf levels are split further."""
    regularize_percentile: float = 90.
    max_volume_fac: float = 20.


@dataclass(unsafe_hash=True)
class TreeConfig():
    """Shape of the leaf/node hierarchy and the size of its preallocated arrays."""
    # leaf size, node-plane coarsening factor, and the node count at which coarsening stops
    max_leaf_size: int = 32
    coarse_fac: float = 6.0
    stop_coarsen: int = 1024
    regularization: RegularizationConfig | None = None

    # node arrays are sized for alloc_fac_nodes times the expected count
    alloc_fac_nodes: float = 1.0

    # key samples per rank when splitting a multi-GPU domain (unused on one GPU)
    nsamp: int = 1024

    # node centres at the centre of mass instead of the cell centre
    mass_centered: bool = False


@dataclass(unsafe_hash=True)
class FofCatalogueConfig():
    """Groups with fewer than npart_min members are left out of a FoF catalogue."""
    npart_min: int = 20


@dataclass(unsafe_hash=True)
class FofConfig:
    """Settings of fof_labels: interaction-list allocation and the tree it builds."""
    alloc_fac_ilist: float = 32.
    alloc_fac_distr_links: float = 0.01

    tree: TreeConfig = field(default_factory=lambda: TreeConfig(
        max_leaf_size=48,
        coarse_fac=8.,
        alloc_fac_nodes=1.1,
        stop_coarsen=2048,
        mass_centered=False
    ))
    catalogue: FofCatalogueConfig = field(default_factory=FofCatalogueConfig)


@dataclass(unsafe_hash=True)
class KNNConfig:
    """Settings of knn: interaction-list allocation and the tree it builds."""
    alloc_fac_ilist: float = 256.

    tree: TreeConfig = field(default_factory=lambda: TreeConfig(
        max_leaf_size=48,
        coarse_fac=8.,
        alloc_fac_nodes=1.,
        stop_coarsen=2048,
        mass_centered=False,
        regularization=RegularizationConfig()
    ))
