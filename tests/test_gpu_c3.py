"""The batched form at its own size, built as bench.py builds it
(amdat_amd.workloads.batch_members / batch_shells): 256 trajectories with
N uniform in [500, 20,000], density U[1.0, 1.2], `exponential 4 50 1.25`
(T = 201, 207 pairs with fullblock 0); S(k) on 60 shells (bins 2..61,
every frame), Fs on the three shells at each member's first peak, MSD;
all in one persistent launch, autotuned and replayed from a hipGraph.

Gate: S and Fs within 1e-6 of FP64 (s = |S| for S, 1 for Fs), MSD within
1e-12 (both FP64).  FP64 is the plain-FP64 single-trajectory path for every
member, and the oracle for three members."""
import os

import numpy as np
import pytest

from amdat_amd import _abi, synth, workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TOL = 1e-6
orc.set_threads(min(16, os.cpu_count() or 1))
B = workloads.BATCH


def rel(y, y64, s):
    return float(np.max(np.abs(y - y64) / np.maximum(np.abs(y64), s)))


@pytest.fixture(scope="module")
def c3(ctx):
    print("case: batched form, 256 trajectories, N in [500, 20000], exponential 4 50", flush=True)
    members, cols, Ls = workloads.batch_members(ctx, B["n_traj"], B["n_min"], B["n_max"], B["nexp"], B["nes"],
                                                synth.MASTER_SEED + 2)
    off, nv, sk_last, fs_first, fs_last = workloads.batch_shells(Ls)
    sh = _abi.Shells(ctx, off, nv)
    tg, t0, t1 = _abi.ts_pairs(B["nexp"], B["nes"], 0, 0, 0)
    ntg = _abi.ts_n_timegaps(B["nexp"], B["nes"], 0)
    P = _abi.Pairs(ctx, tg, t0, t1, ntg)
    b = _abi.Batch(ctx, members)
    args = dict(sk=(sh, 0, sk_last), fs=(sh, fs_first, fs_last), pairs=P, msd=True)
    print("case: batched job, tuned, graph capture and two replays", flush=True)
    outs = [b.run(Ls, graph=True, tune=True, **args) for _ in range(3)]
    info = b.job_info()
    st = b.job_settings()
    print("job", info, st, flush=True)
    yield dict(members=members, cols=cols, Ls=Ls, off=off, nv=nv, sk_last=sk_last, fs_first=fs_first,
               fs_last=fs_last, sh=sh, P=P, pairs=(tg, t0, t1), ntg=ntg, b=b, outs=outs, info=info, st=st)
    del b
    del cols


def test_c3_job_shape_and_replays(c3):
    outs, info, st = c3["outs"], c3["info"], c3["st"]
    assert len(c3["members"]) == 256
    assert info["graph"]
    # the tuner keeps its settings only when the whole job ran no slower
    assert st["tuned"] and 0.0 < st["tuned_ms"] <= st["model_ms"]
    Ns = np.array([m.N for m in c3["members"]], np.float64)
    assert np.array_equal(outs[0]["S_count"], Ns * 201)
    assert np.array_equal(outs[0]["msd_count"], Ns[:, None] * np.bincount(c3["pairs"][0], minlength=c3["ntg"]))
    # replays of the captured graph give the captured launch's bits
    for o in outs[1:]:
        assert np.array_equal(o["S"], outs[0]["S"])
        assert all(np.array_equal(a, b) for a, b in zip(o["Fs"], outs[0]["Fs"]))
        assert np.array_equal(o["s2"], outs[0]["s2"]) and np.array_equal(o["s4"], outs[0]["s4"])


def test_c3_against_plain_fp64(ctx, c3):
    o = c3["outs"][-1]
    eS = eF = eM = 0.0
    for i, m in enumerate(c3["members"]):
        L = np.float32(c3["Ls"][i])
        r = _abi.Rho(ctx, m, None, c3["sh"], 0, c3["sk_last"], L, precision=_abi.PREC_FP64)
        S64 = _abi.sk(ctx, r)
        r.close()
        eS = max(eS, rel(o["S"][i], S64, np.abs(S64)))
        F64, _ = _abi.fs(ctx, m, None, c3["sh"], int(c3["fs_first"][i]), int(c3["fs_last"][i]), L, c3["P"],
                         precision=_abi.PREC_FP64)
        eF = max(eF, rel(o["Fs"][i], F64, 1.0))
        s2, s4, c = _abi.msd_ngp(ctx, m, None, c3["P"])
        eM = max(eM, rel(o["s2"][i][1:], s2[1:], np.abs(s2[1:])), rel(o["s4"][i][1:], s4[1:], np.abs(s4[1:])))
    print("C3 (256 members) vs plain FP64: S %.3e Fs %.3e MSD/s4 %.3e" % (eS, eF, eM))
    assert eS < TOL
    assert eF < TOL
    assert eM < 1e-12


def test_c3_members_against_oracle(ctx, c3):
    """The three smallest members against the oracle over every frame,
    pair and vector of the job (the oracle's FP64 cost grows with N)."""
    o = c3["outs"][-1]
    Ns = np.array([m.N for m in c3["members"]])
    picks = [int(i) for i in np.argsort(Ns)[:3]]
    tg, t0, t1 = c3["pairs"]
    off, nv = c3["off"], c3["nv"]
    so = off[:c3["sk_last"] + 2]
    nsk = nv[:off[c3["sk_last"] + 1]]
    for i in picks:
        t = workloads.member_traj(c3["cols"][i], c3["Ls"][i])
        ref = orc.rho_f64_traj(t, nsk, t.L)
        S64 = orc.sk_f64(ref, ref, np.full(t.T, t.N), so)
        eS = rel(o["S"][i], S64, np.abs(S64))
        f, l = int(c3["fs_first"][i]), int(c3["fs_last"][i])
        F64, c64 = orc.fs_f64(t, (tg, t0, t1), c3["ntg"], off[f:l + 2] - off[f], nv[off[f]:off[l + 1]], t.L)
        assert np.array_equal(o["Fs_count"][i], c64)
        eF = rel(o["Fs"][i], F64, 1.0)
        m = orc.msd_ngp_f64(t, (tg, t0, t1), c3["ntg"])
        eM = rel(o["s2"][i][1:], m["s2"][1:], np.abs(m["s2"][1:]))
        print("C3 member %d (N=%d) vs oracle: S %.3e Fs %.3e MSD %.3e" % (i, t.N, eS, eF, eM))
        assert eS < TOL and eF < TOL and eM < 1e-12
