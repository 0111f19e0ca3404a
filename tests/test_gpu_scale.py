"""The scale form at its full size (N = 1,048,576, L = 96, T = 2,001, all
21,001 time-origin pairs) on the benchmark's shells (rho, S(k) and F(k,t)
on 120 shells, bins 2..121, 32,543 vectors; Fs on bins 20, 40, 60, 80, 100),
checked by size-independent properties: Fs(k,0) = 1 and F(k,0) = S(k) over
the same origins to FP64 rounding, MSD(0) = 0 and exact atom-pair counts,
MSD(t) = 3 sigma^2 dt to sampling precision, S(k) -> 1 for the uniform
start; and the compensated-FP32 S(k), F(k,t) and Fs against the plain-FP64
device path at full N on frame and pair subsets."""
import numpy as np
import pytest

from amdat_amd import _abi, synth

pytestmark = pytest.mark.gpu

N, L, NEXP, NES, SIGMA = 1048576, 96.0, 20, 100, 0.05


def bench_shells(ctx):
    """bench.py's scale-form shells: F/S bins 2..121, Fs bins 20..100 step 20."""
    off, nv = _abi.shells_generate(2, 120, 300, seed=synth.MASTER_SEED + 14)
    rows = [nv[off[b - 2]:off[b - 1]] for b in (20, 40, 60, 80, 100)]
    offS = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    return _abi.Shells(ctx, off, nv), _abi.Shells(ctx, offS, np.concatenate(rows).astype(np.int32))


@pytest.fixture(scope="module")
def scale(ctx):
    T = NES * NEXP + 1
    print("case: scale form N=%d T=%d (37.8 GB resident)" % (N, T), flush=True)
    cols = synth.device_random_walk(N, L, T, SIGMA, synth.MASTER_SEED + 4)
    d = _abi.Trajectory.from_device(ctx, N, T, [c.data_ptr() for c in cols], np.full((T, 3), L, np.float32))
    tg, t0, t1 = _abi.ts_pairs(NEXP, NES, 0, 1, 0)
    P = _abi.Pairs(ctx, tg, t0, t1, _abi.ts_n_timegaps(NEXP, NES, 0))
    yield d, P, (tg, t0, t1), cols
    del cols


def test_msd_and_counts(ctx, scale):
    d, P, (tg, t0, t1), _ = scale
    s2, s4, cnt = _abi.msd_ngp(ctx, d, None, P)
    assert np.array_equal(cnt, np.bincount(tg, minlength=P.ntg) * float(N))
    assert s2[0] == 0.0 and s4[0] == 0.0
    lag = np.zeros(P.ntg)
    for g, a, b in zip(tg, t0, t1):
        lag[g] = b - a
    msd = s2[1:] / cnt[1:]
    want = 3 * SIGMA ** 2 * lag[1:]
    # sampling error of a mean over >= 1e6 Gaussian walks: well under 1%
    assert np.max(np.abs(msd / want - 1)) < 0.01


def test_fs_and_f_identities(ctx, scale):
    d, P, (tg, t0, t1), _ = scale
    shF, shS = bench_shells(ctx)
    Fs, _ = _abi.fs(ctx, d, None, shS, 0, 4, np.float32(L), P)
    assert np.max(np.abs(Fs[0] - 1.0)) < 1e-12
    assert 1.0 > Fs[1, 0] > Fs[50, 0] > Fs[-1, 0] > 0.0  # decays with the lag
    # F(k,0) over the lag-0 pairs (one per block) and S(k) from the same
    # density (same frames: those pairs' origins) are the same sums
    lag0 = np.flatnonzero(tg == 0)
    origins = np.unique(t0[lag0])
    r0 = _abi.Rho(ctx, d, None, shF, 0, 119, np.float32(L), frames=origins)
    P0 = _abi.Pairs(ctx, tg[lag0], t0[lag0], t1[lag0], P.ntg)
    F, _ = _abi.fkt(ctx, r0, r0, P0)
    S0 = _abi.sk(ctx, r0)
    print("scale F(k,0)/S(k) - 1: %.3e over %d origins, 120 shells" % (np.max(np.abs(F[0] / S0 - 1.0)), len(origins)))
    assert np.max(np.abs(F[0] / S0 - 1.0)) < 1e-12
    assert abs(np.mean(S0[-10:]) - 1.0) < 0.05             # uniform start: S(k) -> 1


def test_sk_and_f_against_fp64_at_full_n(ctx, scale):
    """S(k) and F(k,t) on all 120 shells from the compensated-FP32 density
    against the plain-FP64 density, full N, on 24 pairs spread over the
    lags (their 46 frames)."""
    d, P, (tg, t0, t1), _ = scale
    shF, _ = bench_shells(ctx)
    pick = np.linspace(0, len(tg) - 1, 24).astype(np.int64)
    frames = np.unique(np.concatenate([t0[pick], t1[pick]]))
    r32 = _abi.Rho(ctx, d, None, shF, 0, 119, np.float32(L), frames=frames)
    r64 = _abi.Rho(ctx, d, None, shF, 0, 119, np.float32(L), frames=frames, precision=_abi.PREC_FP64)
    S, S64 = _abi.sk(ctx, r32), _abi.sk(ctx, r64)
    Pp = _abi.Pairs(ctx, tg[pick], t0[pick], t1[pick], P.ntg)
    F, c = _abi.fkt(ctx, r32, r32, Pp)
    F64, _ = _abi.fkt(ctx, r64, r64, Pp)
    rows = np.unique(tg[pick])
    eS = float(np.max(np.abs(S - S64) / np.abs(S64)))
    eF = float(np.max(np.abs(F[rows] - F64[rows]) / np.maximum(np.abs(F64[rows]), np.abs(S64)[None])))
    print("full-N S rel err %.3e, F rel err %.3e (%d frames, %d pairs, 120 shells)" % (eS, eF, len(frames), len(pick)))
    r32.close()
    r64.close()
    assert eS < 1e-6
    assert eF < 1e-6


def test_fs_against_fp64_at_full_n(ctx, scale):
    d, P, (tg, t0, t1), _ = scale
    _, shS = bench_shells(ctx)                             # the five Fs shells, 1,500 vectors
    pick = np.arange(0, len(tg), 97)                       # 217 pairs across all lags
    Pp = _abi.Pairs(ctx, tg[pick], t0[pick], t1[pick], P.ntg)
    F32, _ = _abi.fs(ctx, d, None, shS, 0, 4, np.float32(L), Pp)
    F64, _ = _abi.fs(ctx, d, None, shS, 0, 4, np.float32(L), Pp, precision=_abi.PREC_FP64)
    err = float(np.max(np.abs(F32 - F64)))
    print("full-N Fs error vs plain FP64: %.3e" % err)
    assert err < 1e-6
