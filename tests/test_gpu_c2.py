"""The precision form at its own size, built as bench.py builds it
(amdat_amd.workloads.precision): the 64,000-atom bead-spring melt, L = 42.2,
`exponential 10 100 1.2` (T = 1,001 frames, 1,046 pairs with fullblock 0);
rho_k(t) on 100 shells (bins 2,4,...,200 of the box-L grid, 28,181
vectors) with S(k) and F(k,t) from it; Fs(k,t) on 99 shells at
max_length_scale 21.05 (26,243 vectors); MSD.

Gate, as for every parity test: S, F, Fs and MSD within 1e-6 of FP64 with
rel = max |y - y64| / max(|y64|, s), s = |y| for S and MSD, S(k) of the
shell for F, 1 for Fs.  The FP64 values are the plain-FP64 device path over
every frame and pair (pinned to the oracle by test_gpu_parity.py), and the
oracle itself on frame and pair subsets at full N."""
import os

import numpy as np
import pytest

from amdat_amd import _abi, workloads
from oracle import oracle as orc
from tests.conftest import rho_rel

pytestmark = pytest.mark.gpu

TOL = 1e-6
orc.set_threads(min(16, os.cpu_count() or 1))


def rel(y, y64, s):
    return float(np.max(np.abs(y - y64) / np.maximum(np.abs(y64), s)))


@pytest.fixture(scope="module")
def c2(ctx):
    print("case: precision form N=64000 T=1001, 28181 F vectors, 26243 Fs vectors", flush=True)
    cfg, tr, (offF, nF), (offS, nS), (tg, t0, t1, ntg) = workloads.precision()
    d = _abi.Trajectory(ctx, *tr.arrays)
    shF = _abi.Shells(ctx, offF, nF)
    shS = _abi.Shells(ctx, offS, nS)
    P = _abi.Pairs(ctx, tg, t0, t1, ntg)
    return dict(cfg=cfg, tr=tr, d=d, offF=offF, nF=nF, offS=offS, nS=nS, shF=shF, shS=shS, P=P,
                pairs=(tg, t0, t1), ntg=ntg, L=np.float32(cfg["L"]), ell=np.float32(cfg["ell_fs"]))


@pytest.fixture(scope="module")
def rho(ctx, c2):
    nsh = len(c2["offF"]) - 1
    r32 = _abi.Rho(ctx, c2["d"], None, c2["shF"], 0, nsh - 1, c2["L"])
    r64 = _abi.Rho(ctx, c2["d"], None, c2["shF"], 0, nsh - 1, c2["L"], precision=_abi.PREC_FP64)
    yield r32, r64
    r32.close()
    r64.close()


def test_c2_sk_and_fkt_against_fp64(ctx, c2, rho):
    r32, r64 = rho
    assert (r32.n_frames, r32.K) == (1001, 28181)
    S, S64 = _abi.sk(ctx, r32), _abi.sk(ctx, r64)
    eS = rel(S, S64, np.abs(S64))
    F, c = _abi.fkt(ctx, r32, r32, c2["P"])
    F64, c64 = _abi.fkt(ctx, r64, r64, c2["P"])
    assert np.array_equal(c, c64)
    assert np.array_equal(c, np.bincount(c2["pairs"][0], minlength=c2["ntg"]) * 64000.0)
    eF = rel(F, F64, np.broadcast_to(np.abs(S64), F64.shape))
    print("C2 S rel err %.3e, F rel err %.3e (%d lags x %d shells)" % (eS, eF, F.shape[0], F.shape[1]))
    assert eS < TOL
    assert eF < TOL


def test_c2_rho_against_oracle_at_full_n(ctx, c2, rho):
    """rho per vector on three frames and every tenth shell against the
    FP64 oracle over all 64,000 atoms."""
    r32, _ = rho
    tr, off, n = c2["tr"], c2["offF"], c2["nF"]
    frames = [0, 500, 1000]
    shells = list(range(0, 100, 10))
    cols = np.concatenate([np.arange(off[s], off[s + 1]) for s in shells])
    got = r32.download()[frames][:, cols]
    L = float(c2["L"])
    ref = np.stack([orc.rho_f64(tr.x[f], tr.y[f], tr.z[f], n[cols], L) for f in frames])
    e = rho_rel(got, ref, tr.N)
    print("C2 rho per-vector err vs oracle %.3e (%d vectors x %d frames)" % (e, len(cols), len(frames)))
    assert e < TOL


def test_c2_fs_against_fp64_and_oracle(ctx, c2):
    nsh = len(c2["offS"]) - 1
    F, c = _abi.fs(ctx, c2["d"], None, c2["shS"], 0, nsh - 1, c2["ell"], c2["P"])
    F64, c64 = _abi.fs(ctx, c2["d"], None, c2["shS"], 0, nsh - 1, c2["ell"], c2["P"], precision=_abi.PREC_FP64)
    assert np.array_equal(c, c64)
    e = rel(F, F64, 1.0)
    print("C2 Fs rel err vs plain FP64 %.3e (%d lags x %d shells)" % (e, F.shape[0], F.shape[1]))
    assert e < TOL
    assert np.max(np.abs(F[0] - 1.0)) < 1e-7
    # the oracle on the pairs of three lags (one in-block, two cross-block)
    # and five shells, all 64,000 atoms
    tg, t0, t1 = c2["pairs"]
    sel = np.isin(tg, [3, 104, 109])
    sh = [0, 30, 60, 90, 98]
    off, n = c2["offS"], c2["nS"]
    o = np.concatenate([[0], np.cumsum([off[s + 1] - off[s] for s in sh])]).astype(np.int32)
    nv = np.concatenate([n[off[s]:off[s + 1]] for s in sh])
    F64o, _ = orc.fs_f64(c2["tr"], (tg[sel], t0[sel], t1[sel]), c2["ntg"], o, nv, float(c2["ell"]))
    Psub = _abi.Pairs(ctx, tg[sel], t0[sel], t1[sel], c2["ntg"])
    Fsub, _ = _abi.fs(ctx, c2["d"], None, c2["shS"], 0, nsh - 1, c2["ell"], Psub)
    rows = [3, 104, 109]
    eo = rel(Fsub[rows][:, sh], F64o[rows], 1.0)
    print("C2 Fs rel err vs oracle %.3e (lags 3, 104, 109; 5 shells)" % eo)
    assert eo < TOL


def test_c2_msd_against_oracle(ctx, c2):
    s2, s4, c = _abi.msd_ngp(ctx, c2["d"], None, c2["P"])
    m = orc.msd_ngp_f64(c2["tr"], c2["pairs"], c2["ntg"])
    assert np.array_equal(c, m["count"])
    assert s2[0] == 0.0
    e = rel(s2[1:] / c[1:], m["msd"][1:], np.abs(m["msd"][1:]))
    print("C2 MSD rel err vs oracle %.3e" % e)
    assert e < 1e-12
