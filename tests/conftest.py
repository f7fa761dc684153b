import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")
    config.addinivalue_line("markers", "slow: larger sizes")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


def cosmo2d():
    return np.load(os.path.join(GOLDEN, "pos2d_cosmo_128_100.npy"), allow_pickle=False)


# labels[0:40] of fof_labels(cosmo_2d_sample(), rlink=0.2*100/128, boxsize=100) in z-order, and the
# number of groups with >= 10 particles: docs/quickstart.md:115-131 and :219-251 of the reference
COSMO2D_LABELS40 = [0, 1, 2, 3, 4, 5, 6, 7, 6, 9, 10, 11, 12, 13, 14, 15, 16, 16, 16, 16, 20, 21, 22, 23, 24, 16,
                    16, 27, 27, 23, 30, 27, 27, 33, 34, 35, 35, 35, 38, 39]
COSMO2D_NGROUPS_GE10 = 164

# checks/tests/test_tree.py:85-108 of the reference (z-order with infinities and NaN)
ZSORT_GOLDEN_IN = [[0.1, 0.1, 0.5], [np.nan, np.nan, np.nan], [0.5, 0.1, 0.1], [np.inf, 0.2, np.inf],
                   [-0.7, np.inf, 0.], [-np.inf, -np.inf, -np.inf], [0.3, 0.3, 0.3]]
ZSORT_GOLDEN_OUT = [[-np.inf, -np.inf, -np.inf], [-0.7, np.inf, 0.], [0.3, 0.3, 0.3], [0.1, 0.1, 0.5],
                    [0.5, 0.1, 0.1], [np.inf, 0.2, np.inf], [np.nan, np.nan, np.nan]]
