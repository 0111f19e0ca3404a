"""The tabled-FP64 form (AMDAT_PREC_FP64T: the compensated-FP32 form's plan
and separable phase tables, with FP64 entries, products and sums) against
the FP64 oracle and the plain-FP64 device path.

Gate: per-vector rho within 1e-12 of max(|rho|, sqrt(N)) of the oracle;
Fs within 1e-12 absolute (|Fs| <= 1); the same plan layouts as the FP32
form (R = 2, 4, 8; unpaired, paired and grouped), table half-widths 24, 50,
100 and 127 (there also the compensated-FP32 form at its 1e-6 gate), an
atom subset and atom splits."""
import os

import numpy as np
import pytest

from amdat_amd import _abi, synth
from oracle import oracle as orc
from tests.conftest import random_vectors, ref_shells, rho_rel

pytestmark = pytest.mark.gpu
orc.set_threads(min(16, os.cpu_count() or 1))

TOL_T64 = 1e-12
PLANS = [(2, 0), (4, 0), (8, 0), (4, 1), (8, 1), (4, 4), (8, 4)]


@pytest.fixture(scope="module")
def data(ctx, qvec):
    print("case: tabled FP64, plain form N=1000 T=101, bins 2..100", flush=True)
    tr = synth.plain_config()
    d = _abi.Trajectory(ctx, *tr.arrays)
    off, n = ref_shells(qvec, 2, 100)
    L = np.float32(9.41)
    frames = np.array([0, 37, 100], np.int32)
    ref = np.stack([orc.rho_f64(tr.x[f], tr.y[f], tr.z[f], n, float(L)) for f in frames])
    tg, t0, t1 = _abi.ts_pairs(5, 20, 0, 0, 0)
    ntg = _abi.ts_n_timegaps(5, 20, 0)
    sel = np.isin(tg, [0, 3, 19, 21, 24])
    pairs = (tg[sel], t0[sel], t1[sel])
    F64, _ = orc.fs_f64(tr, pairs, ntg, off, n, float(np.float32(L / 2)))
    return tr, d, off, n, L, frames, ref, pairs, ntg, F64


@pytest.mark.parametrize("R,pair", PLANS)
def test_tabled64_each_plan(ctx, data, R, pair):
    tr, d, off, n, L, frames, ref, pairs, ntg, F64 = data
    sh = _abi.Shells(ctx, off, n)
    sh.plan_hint(R, pair)
    rho = _abi.Rho(ctx, d, None, sh, 0, 98, L, frames=frames, precision=_abi.PREC_FP64T).download()
    err = rho_rel(rho, ref, tr.N)
    P = _abi.Pairs(ctx, *pairs, ntg)
    F, _ = _abi.fs(ctx, d, None, sh, 0, 98, np.float32(L / 2), P, precision=_abi.PREC_FP64T)
    used = np.unique(pairs[0])
    eF = float(np.max(np.abs(F[used] - F64[used])))
    print("R=%d pair=%d tabled FP64: rho per-vector err %.3e, Fs err %.3e" % (R, pair, err, eF))
    assert err < TOL_T64
    assert eF < TOL_T64


@pytest.mark.parametrize("M", [20, 45, 95, 127])
def test_table_widths(ctx, M):
    """Vectors with components up to M: table half-widths 24, 50, 100, 127
    (each with its own atoms per tile), 300 atoms with a 97-atom subset;
    the tabled-FP64 form at its gate and the compensated-FP32 form at 1e-6,
    both against the oracle."""
    print("case: table width, components up to %d" % M, flush=True)
    tr = synth.random_walk(300, 6.3, 4, 0.05, seed=5 + M)
    d = _abi.Trajectory(ctx, *tr.arrays)
    off, n = random_vectors(M)
    sh = _abi.Shells(ctx, off, n)
    L = float(np.float32(6.3))
    sub = np.sort(np.random.default_rng(M).choice(300, 97, replace=False)).astype(np.int32)
    al = _abi.AtomList(ctx, d, sub)
    last = len(off) - 2
    ref = orc.rho_f64_traj(tr, n, L, atoms=sub)
    tg, t0, t1 = _abi.ts_pairs(1, 3)
    ntg = _abi.ts_n_timegaps(1, 3, 0)
    P = _abi.Pairs(ctx, tg, t0, t1, ntg)
    ell = np.float32(L / 2)
    F64, _ = orc.fs_f64(tr, (tg, t0, t1), ntg, off, n, float(ell), atoms=sub)
    for prec, tol in ((_abi.PREC_FP64T, TOL_T64), (_abi.PREC_FP32C, 1e-6)):
        got = _abi.Rho(ctx, d, al, sh, 0, last, np.float32(L), precision=prec).download()
        err = rho_rel(got, ref, len(sub))
        F, _ = _abi.fs(ctx, d, al, sh, 0, last, ell, P, precision=prec)
        eF = float(np.max(np.abs(F - F64)))
        print("M=%d precision %d: rho err %.3e, Fs err %.3e" % (M, prec, err, eF))
        assert err < tol
        assert eF < tol


def test_tabled64_splits_against_device_fp64(ctx, qvec):
    """N = 20,000 on few frames: the launch splits the atoms (FP64 partial
    rows summed in a second pass); S(k) and Fs against the plain-FP64
    device path."""
    print("case: tabled FP64, N=20000 T=4 with atom splits", flush=True)
    tr = synth.random_walk(20000, 28.0, 4, 0.03, seed=41)
    d = _abi.Trajectory(ctx, *tr.arrays)
    off, n = ref_shells(qvec, 2, 100)
    sh = _abi.Shells(ctx, off, n)
    L = np.float32(28.0)
    rt = _abi.Rho(ctx, d, None, sh, 0, 98, L, precision=_abi.PREC_FP64T)
    rp = _abi.Rho(ctx, d, None, sh, 0, 98, L, precision=_abi.PREC_FP64)
    e = rho_rel(rt.download(), rp.download(), tr.N)
    tg, t0, t1 = _abi.ts_pairs(1, 3)
    P = _abi.Pairs(ctx, tg, t0, t1, _abi.ts_n_timegaps(1, 3, 0))
    Ft, _ = _abi.fs(ctx, d, None, sh, 0, 98, np.float32(14.0), P, precision=_abi.PREC_FP64T)
    Fp, _ = _abi.fs(ctx, d, None, sh, 0, 98, np.float32(14.0), P, precision=_abi.PREC_FP64)
    eF = float(np.max(np.abs(Ft - Fp)))
    print("N=20000 tabled vs plain FP64: rho %.3e, Fs %.3e" % (e, eF))
    assert e < TOL_T64
    assert eF < TOL_T64
    assert np.max(np.abs(Ft[0] - 1.0)) < 1e-13
