"""Parity of the HIP path (called through the C ABI) with the FP64 oracle.

Gate: S, F, Fs and MSD within 1e-6 relative of FP64, with
    rel = max |y - y64| / max(|y64|, s)
    s = |y| for S and MSD, S(k) of the shell for F, 1 for Fs;
per-vector rho within 1e-6 of max(|rho|, sqrt(N)).  The plain-FP64 device
path is held to 1e-10.  Integer work (pairs, counts) is bit-exact."""
import os

import numpy as np
import pytest

from amdat_amd import _abi, synth
from oracle import oracle as orc
from tests.conftest import ref_shells, rho_rel

pytestmark = pytest.mark.gpu

TOL = 1e-6
TOL64 = 1e-10

orc.set_threads(min(16, os.cpu_count() or 1))


def rel(y, y64, s):
    return float(np.max(np.abs(y - y64) / np.maximum(np.abs(y64), s)))


@pytest.fixture(scope="module")
def plain(ctx):
    print("case: plain N=1000 L=9.41 T=7 (exponential 2 3)", flush=True)
    tr = synth.plain_config(N=1000, L=9.41, nexp=2, nes=3, seed=21)
    d = _abi.Trajectory(ctx, *tr.arrays)
    return tr, d


@pytest.fixture(scope="module")
def shells99(ctx, qvec):
    off, n = ref_shells(qvec, 2, 100)
    return off, n, _abi.Shells(ctx, off, n)


@pytest.fixture(scope="module")
def rho_pair(ctx, plain, shells99):
    tr, d = plain
    off, n, sh = shells99
    L = float(np.float32(9.41))
    r32 = _abi.Rho(ctx, d, None, sh, 0, 98, np.float32(9.41))
    r64 = _abi.Rho(ctx, d, None, sh, 0, 98, np.float32(9.41), precision=_abi.PREC_FP64)
    ref = orc.rho_f64_traj(tr, n, L)
    return r32, r64, ref


def test_rho_per_vector(ctx, rho_pair, plain, shells99):
    r32, r64, ref = rho_pair
    N = plain[0].N
    a, b = r32.download(), r64.download()
    rx = _abi.Rho(ctx, plain[1], None, shells99[2], 0, 98, np.float32(9.41), precision=_abi.PREC_FP32X).download()
    assert a.shape == ref.shape
    e32 = rho_rel(a, ref, N)
    e64 = rho_rel(b, ref, N)
    ex = rho_rel(rx, ref, N)
    print("rho err fp32c %.3e fp32x %.3e fp64 %.3e" % (e32, ex, e64))
    assert e32 < TOL
    assert ex < TOL
    assert e64 < TOL64


def test_sk_symmetric(ctx, rho_pair, plain, shells99):
    r32, r64, ref = rho_pair
    off = shells99[0]
    tr = plain[0]
    S64 = orc.sk_f64(ref, ref, np.full(tr.T, tr.N), off)
    S = _abi.sk(ctx, r32)
    Sd = _abi.sk(ctx, r64)
    print("S rel err fp32c %.3e fp64 %.3e" % (rel(S, S64, np.abs(S64)), rel(Sd, S64, np.abs(S64))))
    assert rel(S, S64, np.abs(S64)) < TOL
    assert rel(Sd, S64, np.abs(S64)) < TOL64


def test_fkt(ctx, rho_pair, plain, shells99):
    r32, r64, ref = rho_pair
    tr = plain[0]
    off = shells99[0]
    for fb in (0, 1):
        tg, t0, t1 = _abi.ts_pairs(2, 3, 0, fb, 0)
        ntg = _abi.ts_n_timegaps(2, 3, 0)
        P = _abi.Pairs(ctx, tg, t0, t1, ntg)
        F64, c64 = orc.fkt_f64(ref, ref, np.full(tr.T, tr.N), (tg, t0, t1), ntg, off)
        F, c = _abi.fkt(ctx, r32, r32, P)
        Fd, _ = _abi.fkt(ctx, r64, r64, P)
        Sk = orc.sk_f64(ref, ref, np.full(tr.T, tr.N), off)
        s = np.broadcast_to(np.abs(Sk), F64.shape)
        assert np.array_equal(c, c64)
        print("F rel err fb=%d fp32c %.3e fp64 %.3e" % (fb, rel(F, F64, s), rel(Fd, F64, s)))
        assert rel(F, F64, s) < TOL
        assert rel(Fd, F64, s) < TOL64
        # wn sub-range selects columns
        Fs_, _ = _abi.fkt(ctx, r32, r32, P, 10, 20)
        assert np.allclose(Fs_, F[:, 10:21], rtol=0, atol=1e-13)


def test_fkt_asymmetric(ctx, plain, shells99):
    """isf ... asymmetric: densities of two lists of unequal size; F(A, B)
    correlates rho_A(t0) with rho_B(t1) and divides by list A's atom count
    (src/intermediate_scattering_function.cpp:551-566)."""
    tr, d = plain
    off, n, sh = shells99
    rng = np.random.default_rng(11)
    A = np.sort(rng.choice(tr.N, 300, replace=False)).astype(np.int32)
    B = np.setdiff1d(np.arange(tr.N), A).astype(np.int32)
    LA, LB = _abi.AtomList(ctx, d, A), _abi.AtomList(ctx, d, B)
    L = float(np.float32(9.41))
    rA = _abi.Rho(ctx, d, LA, sh, 0, 98, np.float32(9.41))
    rB = _abi.Rho(ctx, d, LB, sh, 0, 98, np.float32(9.41))
    refA = orc.rho_f64_traj(tr, n, L, atoms=A)
    refB = orc.rho_f64_traj(tr, n, L, atoms=B)
    for fb in (0, 1):
        tg, t0, t1 = _abi.ts_pairs(2, 3, 0, fb, 0)
        ntg = _abi.ts_n_timegaps(2, 3, 0)
        P = _abi.Pairs(ctx, tg, t0, t1, ntg)
        for r1, r2, ref1, ref2, n1 in ((rA, rB, refA, refB, len(A)), (rB, rA, refB, refA, len(B))):
            F64, c64 = orc.fkt_f64(ref1, ref2, np.full(tr.T, n1), (tg, t0, t1), ntg, off)
            S1 = orc.sk_f64(ref1, ref1, np.full(tr.T, n1), off)
            F, c = _abi.fkt(ctx, r1, r2, P)
            assert np.array_equal(c, c64)
            e = rel(F, F64, np.broadcast_to(np.abs(S1), F64.shape))
            print("asymmetric F rel err fb=%d N1=%d: %.3e" % (fb, n1, e))
            assert e < TOL


@pytest.mark.parametrize("fb", [0, 1])
def test_fs(ctx, plain, shells99, fb):
    tr, d = plain
    off, n, sh = shells99
    ell = np.float32(9.41 / 2)
    tg, t0, t1 = _abi.ts_pairs(2, 3, 0, fb, 0)
    ntg = _abi.ts_n_timegaps(2, 3, 0)
    P = _abi.Pairs(ctx, tg, t0, t1, ntg)
    F64, c64 = orc.fs_f64(tr, (tg, t0, t1), ntg, off, n, float(ell))
    F, c = _abi.fs(ctx, d, None, sh, 0, 98, ell, P)
    Fd, _ = _abi.fs(ctx, d, None, sh, 0, 98, ell, P, precision=_abi.PREC_FP64)
    assert np.array_equal(c, c64)
    print("Fs rel err fp32c %.3e fp64 %.3e" % (rel(F, F64, 1.0), rel(Fd, F64, 1.0)))
    assert rel(F, F64, 1.0) < TOL
    assert rel(Fd, F64, 1.0) < TOL64
    assert np.allclose(F[0], 1.0, atol=1e-7)


def test_msd_ngp(ctx, plain):
    tr, d = plain
    for fb in (0, 1):
        tg, t0, t1 = _abi.ts_pairs(2, 3, 0, fb, 0)
        ntg = _abi.ts_n_timegaps(2, 3, 0)
        P = _abi.Pairs(ctx, tg, t0, t1, ntg)
        m = orc.msd_ngp_f64(tr, (tg, t0, t1), ntg)
        s2, s4, c = _abi.msd_ngp(ctx, d, None, P)
        assert np.array_equal(c, m["count"])
        assert s2[0] == 0.0
        # MSD(0) = 0 exactly; the relative gate applies from the first lag on
        assert rel(s2[1:] / c[1:], m["msd"][1:], np.abs(m["msd"][1:])) < 1e-12
        assert rel(s4[1:], m["s4"][1:], np.abs(m["s4"][1:])) < 1e-12
        assert s4[0] == 0.0


def test_atom_subset_and_ragged(ctx, plain, shells99):
    tr, d = plain
    off, n, sh = shells99
    rng = np.random.default_rng(5)
    idx = np.sort(rng.choice(tr.N, 37, replace=False)).astype(np.int32)
    Lst = _abi.AtomList(ctx, d, idx)
    r = _abi.Rho(ctx, d, Lst, sh, 3, 40, np.float32(9.41), frames=[0, 4, 6])
    ref = np.stack([orc.rho_f64(tr.x[f], tr.y[f], tr.z[f], n[off[3]:off[41]], float(np.float32(9.41)), atoms=idx)
                    for f in (0, 4, 6)])
    got = r.download()
    e = np.abs(got - ref)
    i = np.unravel_index(np.argmax(e), e.shape)
    print("subset rho worst: |err| %.3e |rho| %.3f rel %.3e" % (e[i], abs(ref[i]), e[i] / abs(ref[i])))
    assert rho_rel(got, ref, 37) < TOL
    tg, t0, t1 = _abi.ts_pairs(2, 3)
    P = _abi.Pairs(ctx, tg, t0, t1, _abi.ts_n_timegaps(2, 3, 0))
    F, c = _abi.fs(ctx, d, Lst, sh, 5, 7, np.float32(3.0), P)
    F64, c64 = orc.fs_f64(tr, (tg, t0, t1), P.ntg, off[5:9] - off[5], n[off[5]:off[8]], float(np.float32(3.0)), atoms=idx)
    assert np.array_equal(c, c64) and rel(F, F64, 1.0) < TOL
    s2, s4, c = _abi.msd_ngp(ctx, d, Lst, P)
    m = orc.msd_ngp_f64(tr, (tg, t0, t1), P.ntg, atoms=idx)
    assert s2[0] == 0.0 and rel(s2[1:], m["s2"][1:], np.abs(m["s2"][1:])) < 1e-12


def test_negative_nz_vectors_are_conjugated(ctx, plain):
    tr, d = plain
    rng = np.random.default_rng(6)
    n = rng.integers(-20, 21, size=(700, 3)).astype(np.int32)
    n[np.all(n == 0, axis=1)] = [1, 0, 0]
    off = np.array([0, 300, 700], np.int32)
    sh = _abi.Shells(ctx, off, n)
    r = _abi.Rho(ctx, d, None, sh, 0, 1, np.float32(9.41), frames=[2])
    ref = orc.rho_f64(tr.x[2], tr.y[2], tr.z[2], n, float(np.float32(9.41)))
    assert rho_rel(r.download()[0], ref, tr.N) < TOL


def test_single_atom_and_empty_list(ctx, plain, shells99):
    tr, d = plain
    off, n, sh = shells99
    one = _abi.AtomList(ctx, d, np.array([17], np.int32))
    r = _abi.Rho(ctx, d, one, sh, 0, 98, np.float32(9.41), frames=[1, 3])
    assert np.allclose(np.abs(r.download()), 1.0, atol=1e-6)
    assert np.allclose(_abi.sk(ctx, r), 1.0, atol=1e-6)
    empty = _abi.AtomList(ctx, d, np.zeros(0, np.int32))
    r0 = _abi.Rho(ctx, d, empty, sh, 0, 5, np.float32(9.41), frames=[0])
    assert np.all(r0.download() == 0)
    assert np.all(_abi.sk(ctx, r0) == 0)


def test_error_codes(ctx, plain, shells99):
    tr, d = plain
    off, n, sh = shells99
    with pytest.raises(_abi.AmdatError) as e:
        _abi.Rho(ctx, d, None, sh, 0, 99, np.float32(9.41))
    assert e.value.code == -4
    with pytest.raises(_abi.AmdatError) as e:
        _abi.Rho(ctx, d, None, sh, 0, 3, np.float32(9.41), frames=[tr.T])
    assert e.value.code == -4
    P = _abi.Pairs(ctx, np.array([0]), np.array([0]), np.array([tr.T + 5]), 1)
    with pytest.raises(_abi.AmdatError) as e:
        _abi.msd_ngp(ctx, d, None, P)
    assert e.value.code == -4
    with pytest.raises(_abi.AmdatError) as e:
        _abi.AtomList(ctx, d, np.array([tr.N], np.int32))
    assert e.value.code == -4


def test_list_of_another_trajectory_is_rejected(ctx, plain, shells99):
    """A list built for a larger trajectory holds indices past this one's
    atoms: every entry point that takes a list refuses it before launching."""
    tr, d = plain
    sh = shells99[2]
    big = _abi.Trajectory(ctx, *synth.random_walk(tr.N + 200, 9.41, tr.T, 0.05, seed=4).arrays)
    far = _abi.AtomList(ctx, big, np.arange(tr.N, tr.N + 200, dtype=np.int32))
    tg, t0, t1 = _abi.ts_pairs(2, 3)
    P = _abi.Pairs(ctx, tg, t0, t1, _abi.ts_n_timegaps(2, 3, 0))
    calls = [lambda: _abi.Rho(ctx, d, far, sh, 0, 3, np.float32(9.41)),
             lambda: _abi.fs(ctx, d, far, sh, 0, 3, np.float32(4.0), P),
             lambda: _abi.msd_ngp(ctx, d, far, P),
             lambda: _abi.displacements(ctx, d, far, P, 0, 1.0)]
    for call in calls:
        with pytest.raises(_abi.AmdatError) as e:
            call()
        assert e.value.code == -1 and "another size" in str(e.value)


def test_deterministic(ctx, plain, shells99):
    tr, d = plain
    off, n, sh = shells99
    a = _abi.Rho(ctx, d, None, sh, 0, 98, np.float32(9.41), frames=[0, 5]).download()
    b = _abi.Rho(ctx, d, None, sh, 0, 98, np.float32(9.41), frames=[0, 5]).download()
    assert np.array_equal(a, b)


def test_large_n_against_device_fp64(ctx, qvec):
    """N = 64,000 at the precision configuration's box: S(k) and Fs(k,t)
    over 99 shells from the compensated FP32 path against the plain-FP64
    device path (itself pinned to the oracle above), and Fs(k,0) = 1."""
    print("case: N=64000 L=42.2 T=4", flush=True)
    tr = synth.random_walk(64000, 42.2, 4, 0.02, seed=31)
    d = _abi.Trajectory(ctx, *tr.arrays)
    off, n = ref_shells(qvec, 2, 100)
    sh = _abi.Shells(ctx, off, n)
    L = np.float32(42.2)
    r32 = _abi.Rho(ctx, d, None, sh, 0, 98, L)
    r64 = _abi.Rho(ctx, d, None, sh, 0, 98, L, precision=_abi.PREC_FP64)
    S, Sd = _abi.sk(ctx, r32), _abi.sk(ctx, r64)
    print("N=64000 S rel err %.3e" % rel(S, Sd, np.abs(Sd)))
    assert rel(S, Sd, np.abs(Sd)) < TOL
    tg, t0, t1 = _abi.ts_pairs(1, 3)
    P = _abi.Pairs(ctx, tg, t0, t1, _abi.ts_n_timegaps(1, 3, 0))
    F, _ = _abi.fs(ctx, d, None, sh, 0, 98, np.float32(21.1), P)
    Fd, _ = _abi.fs(ctx, d, None, sh, 0, 98, np.float32(21.1), P, precision=_abi.PREC_FP64)
    print("N=64000 Fs rel err %.3e" % rel(F, Fd, 1.0))
    assert rel(F, Fd, 1.0) < TOL
    assert np.allclose(F[0], 1.0, atol=1e-7)
