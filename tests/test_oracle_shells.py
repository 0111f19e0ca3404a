"""Pins for the shell rule (src/qvectors/generate_3d_qvectors.m:6-64) and
the reference's q-vector data (src/wave_vectors.cpp:430-484)."""
import numpy as np
import pytest

from oracle import oracle as orc


def _set(a):
    return set(map(tuple, np.asarray(a).reshape(-1, 3).tolist()))


@pytest.mark.parametrize("b", list(range(2, 20)))
def test_complete_shells_equal_data(qvec, b):
    off, n = qvec
    data = n[off[b - 1]:off[b]]
    enum = orc.shell_enumerate(b)
    assert len(enum) == len(data)
    assert _set(enum) == _set(data)


def test_shell_sizes_match_survey(qvec):
    off, _ = qvec
    sizes = np.diff(off)[1:100]        # bins 2..100 = reference shells wn 0..98
    assert sizes[:17].tolist() == [3, 10, 15, 12, 33, 52, 45, 60, 79, 90, 111, 144, 133, 180, 219, 214, 243]
    assert sizes[17:].tolist() == [300] * 82
    assert sizes.sum() == 26243


@pytest.mark.parametrize("b", [20, 21, 22, 50, 99, 100, 150, 200, 201, 300])
def test_subsampled_shells_obey_rule(qvec, b):
    off, n = qvec
    data = n[off[b - 1]:off[b]]
    assert len(data) == 300
    assert len(_set(data)) == 300                      # no duplicates
    full = _set(orc.shell_enumerate(b))
    assert len(full) > 300
    if b <= 200:                                        # shells wholly inside |n_c| <= 100
        assert _set(data) <= full
    l = np.sqrt((data.astype(np.float64) ** 2).sum(1))
    assert np.all(np.floor((l + 0.25) * 2) == b)
    # half space: z > 0, or z == 0 and x > 0, or x == z == 0 and y > 0
    x, y, z = data.T
    assert np.all((z > 0) | ((z == 0) & (x > 0)) | ((z == 0) & (x == 0) & (y > 0)))


def test_enumeration_is_half_space_and_complete_for_small_bins():
    # brute force over the cube: every half-space vector lands in exactly one bin
    got = {}
    for b in range(1, 12):
        for v in map(tuple, orc.shell_enumerate(b).tolist()):
            assert v not in got
            got[v] = b
    r = range(-7, 8)
    for i in r:
        for j in r:
            for k in r:
                if (i, j, k) == (0, 0, 0):
                    continue
                half = k > 0 or (k == 0 and i > 0) or (k == 0 and i == 0 and j > 0)
                b = int(np.floor((np.sqrt(i * i + j * j + k * k) + 0.25) * 2))
                if b <= 11:
                    assert ((i, j, k) in got) == half
                    if half:
                        assert got[(i, j, k)] == b


def test_nominal_wavenumbers():
    # approx_wavenumber[wn] = spacing + spacing/2 * wn = (pi/L)(wn+2)
    L = np.float32(9.41)
    k = orc.approx_wavenumbers(L, 99)
    sp = np.float32(2.0 * np.pi / float(L))
    assert orc.wave_spacing(L) == sp
    assert np.allclose(k, np.pi / float(L) * (np.arange(99) + 2), rtol=1e-6)
    assert k[20] == np.float32(sp + np.float32(sp / np.float32(2.0)) * np.float32(20))
