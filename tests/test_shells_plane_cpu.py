"""Reader of the reference's 2D and 1D q-vector files (amdat_shells_read_plane):
columns placed on the axes as src/wave_vectors.cpp:266-424 places them,
1D half-integer shells as 2n over a denominator of 2, and the oracle's
k = 2 pi n / l with the doubled length scale equal to the 1D k of the
reference (k = (bin/2) * 2 pi / L)."""
import numpy as np
import pytest

from amdat_amd import _abi
from oracle import oracle as orc
from tests.conftest import write_plane_files


@pytest.fixture(scope="module")
def stems(tmp_path_factory, qvec2d, qvec1d):
    d = tmp_path_factory.mktemp("planes")
    return write_plane_files(str(d), qvec2d, qvec1d)


@pytest.mark.parametrize("plane,perm", [("xy", (0, 1, 2)), ("xz", (0, 2, 1)), ("yz", (1, 2, 0))])
def test_2d_placement(stems, qvec2d, plane, perm):
    off2, n2 = qvec2d
    off, n, den = _abi.shells_read_plane(stems["2d"], plane, 2, 99)
    assert den == 1
    raw = n2[off2[2]:off2[101]]
    assert np.array_equal(off, off2[2:102] - off2[2])
    want = np.zeros_like(raw)
    for col, axis in enumerate(perm):   # file column col lands on axis
        want[:, axis] = raw[:, col]
    assert np.array_equal(n, want)


@pytest.mark.parametrize("plane,axis", [("x", 0), ("y", 1), ("z", 2)])
def test_1d_half_integers(stems, qvec1d, plane, axis):
    off, n, den = _abi.shells_read_plane(stems["1d"], plane, 2, 99)
    assert den == 2
    assert np.array_equal(off, np.arange(100))
    assert np.array_equal(n[:, axis], np.arange(2, 101))          # 2 * (bin / 2)
    assert np.all(np.delete(n, axis, axis=1) == 0)
    assert np.array_equal(qvec1d[1:100, 0], np.arange(2, 101))    # the files hold bin / 2


def test_xyz_plane_equals_3d_reader(tmp_path, qvec):
    off3, n3 = qvec
    d = tmp_path / "q3"
    d.mkdir()
    for b in range(2, 12):
        with open(d / ("qvector.%03d" % b), "w", newline="") as fh:
            for v in n3[off3[b - 1]:off3[b]]:
                fh.write("  %.7e   %.7e   %.7e\r\n" % tuple(v))
    a = _abi.shells_read(str(d / "qvector"), 2, 10)
    b = _abi.shells_read_plane(str(d / "qvector"), "xyz", 2, 10)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and b[2] == 1


def test_errors(tmp_path, stems):
    with pytest.raises(_abi.AmdatError) as e:
        _abi.shells_read_plane(stems["2d"], "xw", 2, 3)
    assert e.value.code == -1
    (tmp_path / "bad.002").write_text("0.25 0 0\n")
    with pytest.raises(_abi.AmdatError) as e:
        _abi.shells_read_plane(str(tmp_path / "bad"), "x", 2, 1)
    assert e.value.code == -5


def test_doubled_length_scale_gives_the_1d_wavenumber():
    """rho with n = 2h at length 2L is exp(i (h 2 pi / L) x) summed: the
    closed form of the reference's 1D vectors."""
    rng = np.random.default_rng(3)
    L = 7.3
    x = rng.uniform(0, L, 50).astype(np.float32)
    y = rng.uniform(0, L, 50).astype(np.float32)
    z = rng.uniform(0, L, 50).astype(np.float32)
    h = np.arange(2, 30) / 2.0
    nv = np.stack([2 * h, 0 * h, 0 * h], 1).astype(np.int32)
    rho = orc.rho_f64(x, y, z, nv, 2 * L)
    want = np.exp(1j * np.outer(h * 2 * np.pi / L, x.astype(np.float64))).sum(1)
    assert np.max(np.abs(rho - want)) < 1e-9
