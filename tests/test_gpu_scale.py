"""The scale form at its full size (N = 1,048,576, L = 96, T = 2,001, all
21,001 time-origin pairs), checked by size-independent properties:
Fs(k,0) = 1 and F(k,0) = S(k) over the same origins to FP64 rounding,
MSD(0) = 0 and exact atom-pair counts, MSD(t) = 3 sigma^2 dt to sampling
precision, S(k) -> 1 for the uniform start, and the compensated-FP32 Fs
against the plain-FP64 device path at full N on a subset of pairs."""
import numpy as np
import pytest

from amdat_amd import _abi, synth

pytestmark = pytest.mark.gpu

N, L, NEXP, NES, SIGMA = 1048576, 96.0, 20, 100, 0.05


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
    off, nv = _abi.shells_generate(2, 40, 300, seed=synth.MASTER_SEED + 14)
    sh = _abi.Shells(ctx, off, nv)
    Fs, _ = _abi.fs(ctx, d, None, sh, 19, 21, np.float32(L), P)
    assert np.max(np.abs(Fs[0] - 1.0)) < 1e-12
    assert 1.0 > Fs[1, 0] > Fs[50, 0] > Fs[-1, 0] > 0.0  # decays with the lag
    # F(k,0) over the lag-0 pairs and S(k) from the same density (same
    # frames: the lag-0 origins) are the same sums
    origins = np.unique(t0[tg == 0])
    r0 = _abi.Rho(ctx, d, None, sh, 0, 39, np.float32(L), frames=origins)
    sel = tg == 0
    P0 = _abi.Pairs(ctx, tg[sel], t0[sel], t1[sel], P.ntg)
    F, _ = _abi.fkt(ctx, r0, r0, P0)
    S0 = _abi.sk(ctx, r0)
    assert np.max(np.abs(F[0] / S0 - 1.0)) < 1e-12
    assert abs(np.mean(S0[-10:]) - 1.0) < 0.05             # uniform start: S(k) -> 1


def test_fs_against_fp64_at_full_n(ctx, scale):
    d, P, (tg, t0, t1), _ = scale
    off, nv = _abi.shells_generate(20, 1, 300, seed=5)    # one shell, 300 vectors
    sh = _abi.Shells(ctx, off, nv)
    pick = np.arange(0, len(tg), 97)                       # 217 pairs across all lags
    Pp = _abi.Pairs(ctx, tg[pick], t0[pick], t1[pick], P.ntg)
    F32, _ = _abi.fs(ctx, d, None, sh, 0, 0, np.float32(L), Pp)
    F64, _ = _abi.fs(ctx, d, None, sh, 0, 0, np.float32(L), Pp, precision=_abi.PREC_FP64)
    err = float(np.max(np.abs(F32 - F64)))
    print("full-N Fs error vs plain FP64: %.3e" % err)
    assert err < 1e-6
