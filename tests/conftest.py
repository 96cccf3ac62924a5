import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def csr(d, prefix):
    return sp.csr_matrix((d[prefix + "_val"], d[prefix + "_col"], d[prefix + "_rowptr"]),
                         shape=tuple(int(t) for t in d[prefix + "_shape"]))


def gmg_levels(d):
    """[(A_l, P_l)] of an exported geometric multigrid hierarchy."""
    out = []
    for l in range(int(d["nlevels"])):
        A = csr(d, "A") if l == 0 else csr(d, "level%d_A" % l)
        P = csr(d, "level%d_P" % l) if ("level%d_P_shape" % l) in d else None
        out.append((A, P))
    return out


@pytest.fixture(scope="session")
def golden():
    return load
