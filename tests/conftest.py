import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def qvec():
    """AMDAT's 3D q-vector shells (fixture packed from src/qvectors/qvectors.zip).

    Returns (off, n): shell of bin b holds n[off[b-1]:off[b]]."""
    d = np.load(os.path.join(GOLDEN, "qvectors3d.npz"))
    return d["off"].astype(np.int64), d["n"].astype(np.int32)


def ref_shells(qvec, first_bin, last_bin):
    """Shell table (offsets, vectors) for bins first_bin..last_bin as the
    reference's Wave_Vectors holds them (wn = bin - 2)."""
    off, n = qvec
    rows = n[off[first_bin - 1]:off[last_bin]]
    so = (off[first_bin - 1:last_bin + 1] - off[first_bin - 1]).astype(np.int32)
    return so, np.ascontiguousarray(rows)


@pytest.fixture(scope="session")
def ctx():
    from amdat_amd import _abi
    c = _abi.Context(0)
    yield c
    c.close()
