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
