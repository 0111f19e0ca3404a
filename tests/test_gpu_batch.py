"""Batched analyses (many trajectories of uneven size in one persistent
launch per analysis) against the FP64 oracle run member by member, and
against the single-trajectory entry points.

Members follow the batched-workload recipe (N_i uniform, box from a
density in [1.0, 1.2], own Gaussian step size); S(k) uses one integer
shell set with each member's own box length, Fs uses three shells
bracketing each member's first peak (bins b*-1..b*+1, b* = round(7.25 L/pi)).
Edge cases: an atom subset, an empty atom list, a frame subset, and a
queue short enough that frame groups hold several frames."""
import os

import numpy as np
import pytest

from amdat_amd import _abi, synth
from oracle import oracle as orc
from tests.conftest import ref_shells

pytestmark = pytest.mark.gpu

TOL = 1e-6
NEXP, NES = 4, 25      # exponential 4 25: T = 101
orc.set_threads(min(16, os.cpu_count() or 1))


def rel(y, y64, s):
    return float(np.max(np.abs(y - y64) / np.maximum(np.abs(y64), s)))


@pytest.fixture(scope="module")
def members(ctx):
    print("case: batch of 24 members, N in [40, 700], exponential 4 25", flush=True)
    trs = synth.batch_configs(24, 40, 700, NEXP, NES, seed=77)
    devs = [_abi.Trajectory(ctx, *t.arrays) for t in trs]
    rng = np.random.default_rng(3)
    sub = np.sort(rng.choice(trs[5].N, trs[5].N // 3, replace=False)).astype(np.int32)
    lists = [None] * len(trs)
    lists[5] = _abi.AtomList(ctx, devs[5], sub)
    lists[9] = _abi.AtomList(ctx, devs[9], np.zeros(0, np.int32))
    atoms = [None] * len(trs)
    atoms[5] = sub
    atoms[9] = np.zeros(0, np.int32)
    b = _abi.Batch(ctx, devs, lists)
    return trs, devs, lists, atoms, b


@pytest.fixture(scope="module")
def sk_shells(ctx, qvec):
    off, n = ref_shells(qvec, 2, 16)     # wn 0..14
    return off, n, _abi.Shells(ctx, off, n)


def _sk_oracle(t, n, off, frames, atoms):
    ref = orc.rho_f64_traj(t, n, t.L, frames=frames, atoms=atoms)
    na = t.N if atoms is None else len(atoms)
    return orc.sk_f64(ref, ref, np.full(len(frames), na), off)


@pytest.mark.parametrize("items", [None, "64"])
def test_batch_sk(ctx, members, sk_shells, items):
    trs, devs, lists, atoms, b = members
    off, n, sh = sk_shells
    frames = np.arange(trs[0].T)
    old = os.environ.pop("AMDAT_BATCH_ITEMS", None)
    if items is not None:
        os.environ["AMDAT_BATCH_ITEMS"] = items   # short queue: several frames per group
    try:
        S, cnt = b.sk(sh, 0, len(off) - 2, [t.L for t in trs])
    finally:
        os.environ.pop("AMDAT_BATCH_ITEMS", None)
        if old is not None:
            os.environ["AMDAT_BATCH_ITEMS"] = old
    worst = 0.0
    for i, t in enumerate(trs):
        na = t.N if atoms[i] is None else len(atoms[i])
        assert cnt[i] == len(frames) * na
        if na == 0:
            assert np.all(S[i] == 0.0)
            continue
        S64 = _sk_oracle(t, n, off, frames, atoms[i])
        worst = max(worst, rel(S[i], S64, np.abs(S64)))
    print("batch S rel err %.3e (items=%s)" % (worst, items))
    assert worst < TOL


def test_batch_sk_frames_and_single_path(ctx, members, sk_shells):
    trs, devs, lists, atoms, b = members
    off, n, sh = sk_shells
    frames = np.array([0, 7, 50, 100], np.int32)
    S, cnt = b.sk(sh, 2, 9, [t.L for t in trs], frames=frames)
    o2, n2 = off[2:11] - off[2], n[off[2]:off[10]]
    for i in (0, 5, 13):
        t = trs[i]
        S64 = _sk_oracle(t, n2, o2, frames, atoms[i])
        assert rel(S[i], S64, np.abs(S64)) < TOL
        r = _abi.Rho(ctx, devs[i], lists[i], sh, 2, 9, np.float32(t.L), frames=frames)
        S1 = _abi.sk(ctx, r)
        # the single-trajectory path may pair slots (another FP32 product
        # order): agreement to FP32 rounding, not bit for bit
        assert rel(S[i], S1, np.abs(S1)) < 1e-7


@pytest.mark.parametrize("fb", [0, 1])
def test_batch_fs(ctx, members, qvec, fb):
    trs, devs, lists, atoms, b = members
    off, n = ref_shells(qvec, 2, 100)
    sh = _abi.Shells(ctx, off, n)
    first, last = [], []
    for t in trs:
        bstar = int(round(7.25 * t.L / np.pi))
        first.append(bstar - 1 - 2)
        last.append(bstar + 1 - 2)
    tg, t0, t1 = _abi.ts_pairs(NEXP, NES, 0, fb, 0)
    ntg = _abi.ts_n_timegaps(NEXP, NES, 0)
    P = _abi.Pairs(ctx, tg, t0, t1, ntg)
    Fs, cnt = b.fs(sh, first, last, [t.L for t in trs], P)
    worst = 0.0
    for i, t in enumerate(trs):
        o = off[first[i]:last[i] + 2] - off[first[i]]
        nv = n[off[first[i]]:off[last[i] + 1]]
        F64, c64 = orc.fs_f64(t, (tg, t0, t1), ntg, o, nv, t.L, atoms=atoms[i])
        assert np.array_equal(cnt[i], c64)
        if atoms[i] is not None and len(atoms[i]) == 0:
            assert np.all(Fs[i] == 0.0)
            continue
        worst = max(worst, rel(Fs[i], F64, 1.0))
        assert np.allclose(Fs[i][0], 1.0, atol=1e-7)
    print("batch Fs rel err %.3e (fullblock=%d)" % (worst, fb))
    assert worst < TOL
    # one member through the single-trajectory entry point
    i = 3
    F1, _ = _abi.fs(ctx, devs[i], lists[i], sh, first[i], last[i], np.float32(trs[i].L), P)
    # the single-trajectory path may pair slots: agreement to FP32 rounding
    assert rel(Fs[i], F1, 1.0) < 1e-7


def test_batch_msd_ngp(ctx, members):
    trs, devs, lists, atoms, b = members
    tg, t0, t1 = _abi.ts_pairs(NEXP, NES, 0, 0, 0)
    ntg = _abi.ts_n_timegaps(NEXP, NES, 0)
    P = _abi.Pairs(ctx, tg, t0, t1, ntg)
    s2, s4, cnt = b.msd_ngp(P)
    for i, t in enumerate(trs):
        m = orc.msd_ngp_f64(t, (tg, t0, t1), ntg, atoms=atoms[i])
        assert np.array_equal(cnt[i], m["count"])
        if cnt[i][1] == 0:
            assert np.all(s2[i] == 0.0) and np.all(s4[i] == 0.0)
            continue
        assert s2[i][0] == 0.0
        assert rel(s2[i][1:], m["s2"][1:], np.abs(m["s2"][1:])) < 1e-12
        assert rel(s4[i][1:], m["s4"][1:], np.abs(m["s4"][1:])) < 1e-12
        s2s, s4s, _ = _abi.msd_ngp(ctx, devs[i], lists[i], P)
        # same FP64 terms; the single path sums a gap's pairs as a tree
        assert rel(s2[i][1:], s2s[1:], np.abs(s2s[1:])) < 1e-13
        assert rel(s4[i][1:], s4s[1:], np.abs(s4s[1:])) < 1e-13


def test_batch_rejects_mixed_frame_counts(ctx, members):
    trs, devs, lists, atoms, b = members
    other = synth.random_walk(50, 5.0, 7, 0.05, seed=1)
    d = _abi.Trajectory(ctx, *other.arrays)
    with pytest.raises(_abi.AmdatError):
        _abi.Batch(ctx, [devs[0], d])


_QVEC = {}


def qvec_cache():
    return _QVEC["q"]


@pytest.fixture(scope="module")
def job_args(ctx, members, qvec, sk_shells):
    _QVEC["q"] = qvec
    trs, devs, lists, atoms, b = members
    off, n = ref_shells(qvec, 2, 100)
    fsh = _abi.Shells(ctx, off, n)
    first = [int(round(7.25 * t.L / np.pi)) - 3 for t in trs]
    last = [f + 2 for f in first]
    tg, t0, t1 = _abi.ts_pairs(NEXP, NES, 0, 0, 0)
    ntg = _abi.ts_n_timegaps(NEXP, NES, 0)
    P = _abi.Pairs(ctx, tg, t0, t1, ntg)
    return dict(fsh=fsh, first=first, last=last, P=P, ell=[t.L for t in trs])


def _msd_close(x, y, count):
    """Same sums to rounding; lag 0 and members without atoms exactly 0."""
    assert np.all(x[:, 0] == 0.0) and np.all(x[count == 0] == 0.0)
    nz = count[:, 1:] > 0
    assert rel(x[:, 1:][nz], y[:, 1:][nz], np.abs(y[:, 1:][nz])) < 1e-12


def _separate(b, sk_shells, a):
    S, Sc = b.sk(sk_shells[2], 0, len(sk_shells[0]) - 2, a["ell"])
    Fs, Fc = b.fs(a["fsh"], a["first"], a["last"], a["ell"], a["P"])
    s2, s4, mc = b.msd_ngp(a["P"])
    return S, Sc, Fs, Fc, s2, s4, mc


@pytest.mark.parametrize("graph", [False, True])
def test_batch_job_matches_separate_calls(ctx, members, sk_shells, job_args, graph):
    """One persistent launch for S(k) + Fs + MSD gives what the three
    separate batched launches give (same items and sum orders)."""
    trs, devs, lists, atoms, b = members
    a = job_args
    S, Sc, Fs, Fc, s2, s4, mc = _separate(b, sk_shells, a)
    for rep in range(3):          # first call plans (and captures); later calls reuse or replay
        o = b.run(a["ell"], sk=(sk_shells[2], 0, len(sk_shells[0]) - 2), fs=(a["fsh"], a["first"], a["last"]),
                  pairs=a["P"], msd=True, graph=graph)
        info = b.job_info()
        assert info["graph"] == graph
        assert np.array_equal(o["S_count"], Sc) and np.array_equal(o["Fs_count"], Fc)
        assert np.array_equal(o["msd_count"], mc)
        full = Sc > 0
        assert rel(o["S"][full], S[full], np.abs(S[full])) < 1e-12
        assert np.all(o["S"][~full] == 0.0)
        for i in range(len(trs)):
            assert rel(o["Fs"][i], Fs[i], 1.0) < 1e-12
        _msd_close(o["s2"], s2, mc)
        _msd_close(o["s4"], s4, mc)


def test_batch_job_tuned_against_oracle(ctx, members, sk_shells, job_args):
    trs, devs, lists, atoms, b = members
    a = job_args
    off, n, sh = sk_shells
    o = b.run(a["ell"], sk=(sh, 0, len(off) - 2), fs=(a["fsh"], a["first"], a["last"]), pairs=a["P"], msd=True,
              graph=True, tune=True)
    info = b.job_info()
    st = b.job_settings()
    print("tuned", info, st)
    assert info["RS"] in (2, 4, 8) and info["RF"] in (2, 4, 8)
    assert st["tuned"] and st["graph"]
    assert 0.0 < st["tuned_ms"] <= st["model_ms"]
    assert st["sk_plan_mode"] in (0, 1, 4) and 1 <= st["sk_frames_per_item"] <= trs[0].T
    assert st["fs_pairs_per_item"] in (0, 1, 2) and st["fs_atom_parts"] in (1, 2, 4)
    assert st["atoms_per_tile"] in (8, 16, 32)
    tg, t0, t1 = _abi.ts_pairs(NEXP, NES, 0, 0, 0)
    ntg = _abi.ts_n_timegaps(NEXP, NES, 0)
    qoff, qn = ref_shells(qvec_cache(), 2, 100)
    for i in (0, 5, 17):
        t = trs[i]
        S64 = _sk_oracle(t, n, off, np.arange(t.T), atoms[i])
        assert rel(o["S"][i], S64, np.abs(S64)) < TOL
        so = qoff[a["first"][i]:a["last"][i] + 2] - qoff[a["first"][i]]
        nv = qn[qoff[a["first"][i]]:qoff[a["last"][i] + 1]]
        F64, _ = orc.fs_f64(t, (tg, t0, t1), ntg, so, nv, t.L, atoms=atoms[i])
        assert rel(o["Fs"][i], F64, 1.0) < TOL
        m = orc.msd_ngp_f64(t, (tg, t0, t1), ntg, atoms=atoms[i])
        assert rel(o["s2"][i][1:], m["s2"][1:], np.abs(m["s2"][1:])) < 1e-12


def test_batch_job_parts_and_errors(ctx, members, sk_shells, job_args):
    trs, devs, lists, atoms, b = members
    a = job_args
    o = b.run(a["ell"], pairs=a["P"], msd=True)
    s2, s4, mc = b.msd_ngp(a["P"])
    _msd_close(o["s2"], s2, mc)
    _msd_close(o["s4"], s4, mc)
    o = b.run(a["ell"], fs=(a["fsh"], a["first"], a["last"]), pairs=a["P"])
    assert set(o) == {"Fs", "Fs_count"}
    with pytest.raises(_abi.AmdatError):
        b.run(a["ell"])                                   # nothing asked
    with pytest.raises(_abi.AmdatError):
        b.run(a["ell"], msd=True)                         # MSD needs pairs
    with pytest.raises(_abi.AmdatError):
        b.run([-1.0] * len(trs), pairs=a["P"], msd=True)  # bad length scale


def test_batch_job_paired_sk_plan(ctx, members, qvec):
    """S(k) on 60 shells (bins 2..61, 14,543 vectors): the cost model pairs
    the S(k) plan (R = 8, slots sharing ny or nz); results against the
    oracle, including the member with an atom subset."""
    trs, devs, lists, atoms, b = members
    off, n = ref_shells(qvec, 2, 61)
    sh = _abi.Shells(ctx, off, n)
    o = b.run([t.L for t in trs], sk=(sh, 0, len(off) - 2), graph=True)
    info = b.job_info()
    print("paired S(k) job", info)
    assert info["RS"] == 8 and info["paired"]
    for i in (0, 5, 17):
        t = trs[i]
        S64 = _sk_oracle(t, n, off, np.arange(t.T), atoms[i])
        assert rel(o["S"][i], S64, np.abs(S64)) < TOL


ITEM_FORMS = [
    dict(sk_R=8, sk_plan_mode=4, sk_frames_per_item=7, fs_R=4, fs_pairs_per_item=1, fs_atom_parts=3),
    dict(sk_R=4, sk_plan_mode=4, sk_frames_per_item=101, fs_R=2, fs_pairs_per_item=2, fs_atom_parts=2),
    dict(sk_R=4, sk_plan_mode=1, fs_R=8, fs_atom_parts=64),
    dict(sk_R=2, fs_pairs_per_item=3),
]


@pytest.mark.parametrize("form", range(len(ITEM_FORMS)))
def test_batch_job_item_forms(ctx, members, sk_shells, job_args, form):
    """Each item form the autotuner can choose (plan mode, frames per S(k)
    item, pairs per Fs item, atom parts per pair chunk; 64 parts leave the
    smallest members' parts empty) against the separate batched calls (to
    FP32 rounding: tile boundaries and product orders differ) and members
    0, 5 (atom subset) and 17 against the oracle."""
    trs, devs, lists, atoms, b = members
    a = job_args
    off, n, sh = sk_shells
    kw = ITEM_FORMS[form]
    S, Sc, Fs, Fc, s2, s4, mc = _separate(b, sk_shells, a)
    b.use_settings(**kw)
    try:
        o = b.run(a["ell"], sk=(sh, 0, len(off) - 2), fs=(a["fsh"], a["first"], a["last"]), pairs=a["P"],
                  msd=True, graph=True, tune=True)
        st = b.job_settings()
    finally:
        b.use_settings()
    print("form", kw, "->", st)
    assert not st["tuned"]
    assert st["sk_R"] == kw["sk_R"] and st["sk_plan_mode"] == kw.get("sk_plan_mode", 0)
    if "sk_frames_per_item" in kw:
        assert st["sk_frames_per_item"] == kw["sk_frames_per_item"]
        assert st["sk_frame_groups"] == -(-trs[0].T // kw["sk_frames_per_item"])
    assert st["fs_R"] == kw.get("fs_R", st["fs_R"]) and st["fs_atom_parts"] == kw.get("fs_atom_parts", 1)
    assert st["fs_pairs_per_item"] == kw.get("fs_pairs_per_item", 0)
    assert np.array_equal(o["S_count"], Sc) and np.array_equal(o["Fs_count"], Fc)
    full = Sc > 0
    assert rel(o["S"][full], S[full], np.abs(S[full])) < 1e-7
    assert np.all(o["S"][~full] == 0.0)
    for i in range(len(trs)):
        assert rel(o["Fs"][i], Fs[i], 1.0) < 1e-7
    assert np.all(o["Fs"][9] == 0.0)
    _msd_close(o["s2"], s2, mc)
    tg, t0, t1 = _abi.ts_pairs(NEXP, NES, 0, 0, 0)
    ntg = _abi.ts_n_timegaps(NEXP, NES, 0)
    qoff, qn = ref_shells(qvec_cache(), 2, 100)
    for i in (0, 5, 17):
        t = trs[i]
        S64 = _sk_oracle(t, n, off, np.arange(t.T), atoms[i])
        assert rel(o["S"][i], S64, np.abs(S64)) < TOL
        so = qoff[a["first"][i]:a["last"][i] + 2] - qoff[a["first"][i]]
        nv = qn[qoff[a["first"][i]]:qoff[a["last"][i] + 1]]
        F64, _ = orc.fs_f64(t, (tg, t0, t1), ntg, so, nv, t.L, atoms=atoms[i])
        assert rel(o["Fs"][i], F64, 1.0) < TOL


def test_batch_use_settings_rejects(ctx, members):
    b = members[4]
    for kw in (dict(sk_R=3), dict(fs_R=16), dict(sk_R=2, sk_plan_mode=4), dict(sk_R=8, sk_plan_mode=2),
               dict(fs_atom_parts=65), dict(fs_pairs_per_item=-1), dict(sk_frames_per_item=-2)):
        with pytest.raises(_abi.AmdatError):
            b.use_settings(**kw)
    b.use_settings()
