"""Every plan layout gives the same results: rho per vector and Fs on the
same inputs with R = 2, 4, 8 unpaired, R = 4, 8 in both paired modes
(slots sharing ny or nz; partnerless vectors padded, or in unpaired waves),
in quads, and grouped (groups of 8, 4, 2, 1 slots sharing ny or nz; in
mode 5 with waves of singles and pairs holding fewer slots), each against
the FP64 oracle at the 1e-6 gate, on the plain form
(N = 1,000, L = 9.41, exponential 5 20 1.3) with all 99 reference shells."""
import os

import numpy as np
import pytest

from amdat_amd import _abi, synth
from oracle import oracle as orc
from tests.conftest import ref_shells, rho_rel

pytestmark = pytest.mark.gpu
orc.set_threads(min(16, os.cpu_count() or 1))

TOL = 1e-6
PLANS = [(2, 0), (4, 0), (8, 0), (4, 1), (8, 1), (4, 2), (8, 2), (4, 3), (8, 3), (4, 4), (8, 4), (8, 5)]


@pytest.fixture(scope="module")
def data(ctx, qvec):
    print("case: plan layouts, plain form N=1000 T=101, bins 2..100", flush=True)
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
def test_rho_and_fs_each_plan(ctx, data, R, pair):
    tr, d, off, n, L, frames, ref, pairs, ntg, F64 = data
    sh = _abi.Shells(ctx, off, n)
    sh.plan_hint(R, pair)
    rho = _abi.Rho(ctx, d, None, sh, 0, 98, L, frames=frames).download()
    err = rho_rel(rho, ref, tr.N)
    print("R=%d pair=%d rho per-vector err %.3e" % (R, pair, err))
    assert err < TOL
    P = _abi.Pairs(ctx, *pairs, ntg)
    F, _ = _abi.fs(ctx, d, None, sh, 0, 98, np.float32(L / 2), P)
    used = np.unique(pairs[0])
    assert np.max(np.abs(F[used] - F64[used])) < TOL


def test_hint_rejects_bad_values(ctx, data):
    tr, d, off, n, *_ = data
    sh = _abi.Shells(ctx, off, n)
    for R, pair in [(3, 0), (2, 1), (16, 0), (4, 5), (2, 3), (2, 4), (8, 6)]:
        with pytest.raises(_abi.AmdatError):
            sh.plan_hint(R, pair)


def grouped_vectors(M, seed=3):
    """Vectors with components up to M, most in dense (nx, ny) x nz blocks
    (classes of 11 sharing nx and ny: groups of 8 and a pair) and the rest
    scattered (single slots); at least one component at +-M.  Three
    shells."""
    rng = np.random.default_rng(seed + M)
    rows = []
    for nx in [M] + list(rng.choice(np.arange(-M, M), 4, replace=False)):
        nzs = rng.choice(np.arange(-M, M + 1), 11, replace=False)
        for ny in rng.choice(np.arange(-M, M + 1), 3, replace=False):
            rows += [(nx, ny, nz) for nz in nzs]
    rows += [tuple(v) for v in rng.integers(-M, M + 1, size=(60, 3))]
    n = np.unique(np.array(rows, np.int32), axis=0)
    n = n[np.any(n != 0, axis=1)]
    n = n[rng.permutation(len(n))]
    off = np.linspace(0, len(n), 4).astype(np.int32)
    return off, np.ascontiguousarray(n)


_width_cases = {}


@pytest.mark.parametrize("M", [20, 45, 95, 127])
@pytest.mark.parametrize("R,pair", [(2, 0), (8, 1), (8, 4), (8, 5)])
def test_plans_at_table_widths(ctx, M, R, pair):
    """Each layout at table half-widths 24, 50, 100 and 127 (32, 32, 16 and
    8 atoms per tile; 4, 4, 2 and 1 anchors per lane; 300 atoms, so every
    width ends on a partial tile): rho per vector and Fs against the FP64
    oracle at the 1e-6 gate, 4 frames."""
    print("case: plan R=%d mode %d, components up to %d" % (R, pair, M), flush=True)
    if M not in _width_cases:
        tr = synth.random_walk(300, 6.3, 4, 0.05, seed=7 + M)
        off, n = grouped_vectors(M)
        L = float(np.float32(6.3))
        tg, t0, t1 = _abi.ts_pairs(1, 3)
        ntg = _abi.ts_n_timegaps(1, 3, 0)
        ref = orc.rho_f64_traj(tr, n, L)
        F64, _ = orc.fs_f64(tr, (tg, t0, t1), ntg, off, n, float(np.float32(L / 2)))
        _width_cases[M] = (tr, off, n, L, (tg, t0, t1, ntg), ref, F64)
    tr, off, n, L, (tg, t0, t1, ntg), ref, F64 = _width_cases[M]
    d = _abi.Trajectory(ctx, *tr.arrays)
    sh = _abi.Shells(ctx, off, n)
    sh.plan_hint(R, pair)
    got = _abi.Rho(ctx, d, None, sh, 0, len(off) - 2, np.float32(L)).download()
    err = rho_rel(got, ref, tr.N)
    P = _abi.Pairs(ctx, tg, t0, t1, ntg)
    F, _ = _abi.fs(ctx, d, None, sh, 0, len(off) - 2, np.float32(L / 2), P)
    eF = float(np.max(np.abs(F - F64)))
    print("M=%d R=%d mode %d: rho err %.3e, Fs err %.3e" % (M, R, pair, err, eF))
    assert err < TOL
    assert eF < TOL
