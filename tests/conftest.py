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


@pytest.fixture(scope="session")
def qvec2d():
    """AMDAT's 2D q-vector shells: shell b holds n[off[b]:off[b+1]] (files .000-.284)."""
    d = np.load(os.path.join(GOLDEN, "qvectors2d.npz"))
    return d["off"].astype(np.int64), d["n"].astype(np.int32)


@pytest.fixture(scope="session")
def qvec1d():
    """AMDAT's 1D q-vector shells: twice the triple of files .001-.300 (row b-1)."""
    return np.load(os.path.join(GOLDEN, "qvectors1d.npz"))["n2"].astype(np.int32)


def write_plane_files(dirpath, qvec2d=None, qvec1d=None, bins=range(2, 101)):
    """The reference's 2D (tab-separated integers) and 1D (%e, CRLF) files
    for the given bins; returns the stems."""
    stems = {}
    if qvec2d is not None:
        off, n = qvec2d
        d = os.path.join(dirpath, "qvectors2d")
        os.makedirs(d, exist_ok=True)
        for b in bins:
            with open(os.path.join(d, "qvector.%03d" % b), "w") as fh:
                for v in n[off[b]:off[b + 1]]:
                    fh.write("\t%d\t%d\t%d\n" % tuple(v))
        stems["2d"] = os.path.join(d, "qvector")
    if qvec1d is not None:
        d = os.path.join(dirpath, "qvectors1d")
        os.makedirs(d, exist_ok=True)
        for b in bins:
            with open(os.path.join(d, "qvector.%03d" % b), "w", newline="") as fh:
                fh.write("   %.7e   %.7e   %.7e\r\n" % tuple(qvec1d[b - 1] / 2.0))
        stems["1d"] = os.path.join(d, "qvector")
    return stems


def atoms_at(lst, t):
    """Atoms of a time-varying list (conv, off, idx) at frame t."""
    conv, off, idx = lst
    c = conv[t]
    if c < 0 or c >= len(off) - 1:
        return idx[:0]
    return idx[off[c]:off[c + 1]]


def compose_pairs(lst, pairs, fn):
    """For a time-varying list: call fn(sub_pairs, atoms) once per list time
    of the pairs' t0 (skipping empty lists) and yield the results."""
    tg, t0, t1 = (np.asarray(v) for v in pairs)
    keys = np.array([lst[0][t] if 0 <= lst[0][t] < len(lst[1]) - 1 else -1 for t in t0])
    for c in np.unique(keys):
        sel = keys == c
        at = atoms_at(lst, int(t0[sel][0]))
        if len(at) == 0:
            continue
        yield fn((tg[sel], t0[sel], t1[sel]), at)


def rho_rel(got, ref, n_atoms):
    """Per-vector error of rho: |d rho| / max(|rho|, sqrt(N)), the relative
    error where the sum is coherent (|rho| above sqrt(N)) and the error on
    the sqrt(N) scale of a sum of N random phases otherwise."""
    return float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), np.sqrt(n_atoms))))


def random_vectors(M, n_shells=3, per_shell=60, seed=0):
    """Integer vectors with components in [-M, M] (at least one at +-M per
    shell), including nz < 0, in shells of per_shell vectors."""
    rng = np.random.default_rng(seed + M)
    n = rng.integers(-M, M + 1, size=(n_shells * per_shell, 3)).astype(np.int32)
    n[::per_shell, 0] = M
    n[1::per_shell, 2] = -M
    n[np.all(n == 0, axis=1)] = (1, 0, 0)
    off = np.arange(0, n_shells * per_shell + 1, per_shell).astype(np.int32)
    return off, n
