"""CPU-side checks of the C ABI: the library loads, exports every symbol
include/amdat_hip.h declares, and its host integer work (time scheme,
pairs, weights, shell files, shell generation) matches the oracle bit for
bit.  No device entry point is called here."""
import os
import re

import numpy as np
import pytest

from amdat_amd import _abi
from oracle import oracle as orc
from tests.conftest import ROOT


def header_symbols():
    text = open(os.path.join(ROOT, "include", "amdat_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(amdat_\w+)\s*\(", text, re.M)))


def test_exports_every_declared_symbol():
    syms = header_symbols()
    assert len(syms) >= 30
    L = _abi.lib()
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_abi.SIGNATURES)


def test_version():
    assert _abi.lib().amdat_version() == 10000


@pytest.mark.parametrize("nexp,nes,frt,fb,limit", [
    (5, 20, 0, 0, 0), (5, 20, 0, 1, 0), (10, 100, 0, 0, 0), (4, 50, 0, 1, 0), (20, 100, 0, 1, 0),
    (3, 2, 1, 1, 0), (3, 2, 0, 1, 1), (6, 5, 0, 1, 3), (1, 1, 0, 0, 0), (7, 1, 1, 0, 2)])
def test_pairs_and_weights_bit_exact(nexp, nes, frt, fb, limit):
    a = _abi.ts_pairs(nexp, nes, frt, fb, limit)
    b = orc.ts_pairs(nexp, nes, frt, fb, limit)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    assert np.array_equal(_abi.ts_weighting(nexp, nes, frt, fb, limit), orc.ts_weighting(nexp, nes, frt, fb, limit))
    assert _abi.ts_n_timegaps(nexp, nes, frt) == orc.ts_n_timegaps(nexp, nes, frt)


@pytest.mark.parametrize("nexp,nes,base,fe,dt", [(5, 20, 1.3, 0, 1.0), (10, 100, 1.2, 0, 1.0),
                                                 (20, 100, 1.15, 0, 0.005), (4, 50, 1.25, 2, 0.1)])
def test_timelist_bit_exact(nexp, nes, base, fe, dt):
    a = _abi.ts_timelist(nexp, nes, base, fe, dt)
    b = orc.ts_timelist(nexp, nes, base, fe, dt)
    assert np.array_equal(a, b)
    for frt in (0, 1):
        assert np.array_equal(_abi.ts_displacement_times(nexp, nes, frt, a), orc.ts_displacement_times(nexp, nes, frt, b))


def test_time_scheme_errors():
    with pytest.raises(_abi.AmdatError) as e:
        _abi.ts_n_timegaps(0, 5, 0)
    assert e.value.code == -1


def _write_qvector_files(d, qvec, bins):
    off, n = qvec
    for b in bins:
        with open(os.path.join(d, "qvector.%03d" % b), "w", newline="") as fh:
            for v in n[off[b - 1]:off[b]]:
                fh.write("  %.7e   %.7e   %.7e\r\n" % tuple(v))
    return os.path.join(d, "qvector")


def test_shells_read_reference_format(tmp_path, qvec):
    bins = list(range(2, 101))
    stem = _write_qvector_files(str(tmp_path), qvec, bins)
    off, nv = _abi.shells_read(stem, 2, 99)
    qoff, qn = qvec
    assert np.array_equal(off, qoff[1:101] - qoff[1])
    assert np.array_equal(nv, qn[qoff[1]:qoff[100]])


def test_shells_read_errors(tmp_path, qvec):
    stem = _write_qvector_files(str(tmp_path), qvec, [2, 3])
    with pytest.raises(_abi.AmdatError) as e:
        _abi.shells_read(stem, 2, 3)       # qvector.004 missing
    assert e.value.code == -5
    with open(str(tmp_path / "bad.002"), "w") as fh:
        fh.write("1.5 0 0\n")
    with pytest.raises(_abi.AmdatError):
        _abi.shells_read(str(tmp_path / "bad"), 2, 1)


def test_shells_generate_rule(qvec):
    off, nv = _abi.shells_generate(2, 120, 300, seed=11)
    qoff, qn = qvec
    sizes = np.diff(off)
    ref_sizes = np.diff(qoff)[1:121]
    assert np.array_equal(sizes, ref_sizes)          # same per-shell sizes as the reference data
    for s, b in enumerate(range(2, 122)):
        v = nv[off[s]:off[s + 1]]
        l = np.sqrt((v.astype(float) ** 2).sum(1))
        assert np.all(np.floor((l + 0.25) * 2) == b)
        if b < 19:
            assert set(map(tuple, v.tolist())) == set(map(tuple, orc.shell_enumerate(b).tolist()))
    off2, nv2 = _abi.shells_generate(2, 120, 300, seed=11)
    assert np.array_equal(nv, nv2)
    off3, nv3 = _abi.shells_generate(2, 120, 300, seed=12)
    assert not np.array_equal(nv, nv3)
