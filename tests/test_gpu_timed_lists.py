"""Time-varying atom lists on the device: the Fast/Slow/Average particle
selection against the oracle (bit-exact: offsets, atom order, list-time
conversion), and every entry point that takes a list (rho, S(k), F(k,t),
Fs, MSD/NGP, displacement statistics, van Hove counts) on time-varying
lists against the oracle composed list time by list time.  Workload: the
plain form, N = 1,000, L = 9.41, `exponential 5 20 1.3` (T = 101)."""
import os

import numpy as np
import pytest

from amdat_amd import _abi, synth
from oracle import oracle as orc
from tests.conftest import atoms_at, compose_pairs, ref_shells

pytestmark = pytest.mark.gpu
orc.set_threads(min(16, os.cpu_count() or 1))

TOL = 1e-6
NEXP, NES = 5, 20
L = 9.41


@pytest.fixture(scope="module")
def setup(ctx):
    print("case: timed lists, plain form N=1000 T=101", flush=True)
    tr = synth.plain_config()
    d = _abi.Trajectory(ctx, *tr.arrays)
    tg, t0, t1 = _abi.ts_pairs(NEXP, NES, 0, 0, 0)
    ntg = _abi.ts_n_timegaps(NEXP, NES, 0)
    return tr, d, (tg, t0, t1), ntg, _abi.Pairs(ctx, tg, t0, t1, ntg)


def lag_frames(g):
    """Frames spanned by time gap g (frt = 0)."""
    if g < NES:
        return g
    if g < NES + NEXP - 1:
        return NES * (g - NES + 1)
    return NES * NEXP


def d2_quantile(tr, g, q):
    """A threshold at quantile q of the squared displacements over time gap
    g from frame 0."""
    u = tr.unwrapped64()
    d2 = ((u[lag_frames(g)] - u[0]) ** 2).sum(-1)
    return float(np.quantile(d2, q))


def same(dev, ref):
    conv, off, idx = dev.download()
    rc, ro, ri = ref
    assert np.array_equal(conv, rc)
    assert np.array_equal(off, ro)
    assert np.array_equal(idx, ri)


@pytest.mark.parametrize("g", [0, 3, 19, 20, 22, 24])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_select_matches_oracle(ctx, setup, g, mode):
    tr, d, _, _, _ = setup
    lo, hi = d2_quantile(tr, g, 0.3), d2_quantile(tr, g, 0.8)
    dev = _abi.TimedList.select_displacement(ctx, d, None, NEXP, NES, 0, 0, g, mode, lo, hi)
    ref = orc.select_displacement(tr, NEXP, NES, 0, 0, g, mode, lo, hi)
    same(dev, ref)


def test_select_limit_input_lists(ctx, setup):
    tr, d, _, _, _ = setup
    rng = np.random.default_rng(5)
    sub = rng.permutation(tr.N)[:700].astype(np.int32)
    al = _abi.AtomList(ctx, d, sub)
    lo = d2_quantile(tr, 5, 0.5)
    for limit in (0, 2):
        dev = _abi.TimedList.select_displacement(ctx, d, al, NEXP, NES, 0, limit, 5, 0, lo, 0.0)
        ref = orc.select_displacement(tr, NEXP, NES, 0, limit, 5, 0, lo, 0.0, in_list=sub)
        same(dev, ref)
    # select again from a time-varying list (slow among the fast)
    fast = _abi.TimedList.select_displacement(ctx, d, None, NEXP, NES, 0, 0, 10, 0, d2_quantile(tr, 10, 0.4), 0.0)
    ref_fast = orc.select_displacement(tr, NEXP, NES, 0, 0, 10, 0, d2_quantile(tr, 10, 0.4), 0.0)
    hi = d2_quantile(tr, 3, 0.6)
    dev = _abi.TimedList.select_displacement(ctx, d, fast, NEXP, NES, 0, 0, 3, 1, 0.0, hi)
    same(dev, orc.select_displacement(tr, NEXP, NES, 0, 0, 3, 1, 0.0, hi, in_list=ref_fast))
    # every atom (negative threshold) and no atom (huge threshold)
    dev = _abi.TimedList.select_displacement(ctx, d, None, NEXP, NES, 0, 0, 1, 0, -1.0, 0.0)
    conv, off, idx = dev.download()
    assert list(off) == [tr.N * b for b in range(NEXP + 1)] and np.array_equal(idx, np.tile(np.arange(tr.N), NEXP))
    dev = _abi.TimedList.select_displacement(ctx, d, None, NEXP, NES, 0, 0, 1, 0, 1e30, 0.0)
    assert dev.n == 0
    with pytest.raises(_abi.AmdatError):
        _abi.TimedList.select_displacement(ctx, d, None, NEXP, NES, 0, 0, 25, 0, 1.0, 0.0)


def random_timed(tr, rng, n_times=4, p_out=0.15):
    """A random time-varying list: random subsets per list time, random
    list time per frame, some frames past the list times."""
    conv = rng.integers(0, n_times, tr.T).astype(np.int32)
    conv[rng.random(tr.T) < p_out] = n_times + 1
    sets = [rng.permutation(tr.N)[:rng.integers(0, tr.N)] for _ in range(n_times)]
    sets[1] = sets[1][:0]  # one empty list time
    off = np.concatenate([[0], np.cumsum([len(s) for s in sets])]).astype(np.int32)
    idx = np.concatenate(sets).astype(np.int32)
    return conv, off, idx


@pytest.mark.parametrize("kind", ["fast", "random"])
def test_pair_analyses(ctx, setup, qvec, kind):
    tr, d, pairs, ntg, P = setup
    rng = np.random.default_rng(11)
    if kind == "fast":
        dl = _abi.TimedList.select_displacement(ctx, d, None, NEXP, NES, 0, 0, 5, 0, d2_quantile(tr, 5, 0.7), 0.0)
        lst = dl.download()
    else:
        lst = random_timed(tr, rng)
        dl = _abi.TimedList(ctx, d, *lst)
    # MSD / NGP sums and counts
    s2, s4, cnt = _abi.msd_ngp(ctx, d, dl, P)
    R2, R4, RC = np.zeros(ntg), np.zeros(ntg), np.zeros(ntg)
    for m in compose_pairs(lst, pairs, lambda p, at: orc.msd_ngp_f64(tr, p, ntg, atoms=at)):
        R2 += m["s2"]
        R4 += m["s4"]
        RC += m["count"]
    assert np.array_equal(cnt, RC)
    assert np.allclose(s2, R2, rtol=1e-12, atol=1e-12) and np.allclose(s4, R4, rtol=1e-12, atol=1e-12)
    # Fs at |k| = 7.25 (bin 22)
    off, n = ref_shells(qvec, 22, 22)
    sh = _abi.Shells(ctx, off, n)
    F, fc = _abi.fs(ctx, d, dl, sh, 0, 0, np.float32(L), P)
    SF, SC = np.zeros((ntg, 1)), np.zeros(ntg)
    for f, c in compose_pairs(lst, pairs, lambda p, at: orc.fs_f64(tr, p, ntg, off, n, float(np.float32(L)), at)):
        SF += f * c[:, None]
        SC += c
    assert np.array_equal(fc, SC)
    ok = SC > 0
    assert np.max(np.abs(F[ok] - SF[ok] / SC[ok, None])) < TOL
    assert np.all(F[~ok] == 0.0)
    # displacement sums and van Hove counts (bit-exact)
    sd, sd2, dc, hist = _abi.displacements(ctx, d, dl, P, n_bins=60, max_range=1.5)
    H = np.zeros((ntg, 60), np.uint64)
    D = np.zeros((ntg, 3))
    for (h, w), (a, b, c) in zip(compose_pairs(lst, pairs, lambda p, at: orc.vhs_counts(tr, p, ntg, 60, 1.5, at)),
                                 compose_pairs(lst, pairs, lambda p, at: orc.disp_f64(tr, p, ntg, at))):
        H += h
        D += b
    assert np.array_equal(hist, H)
    assert np.allclose(sd2, D, rtol=1e-12, atol=1e-12)
    assert np.array_equal(dc, RC)


def test_density_sk_fkt(ctx, setup, qvec):
    tr, d, pairs, ntg, P = setup
    rng = np.random.default_rng(3)
    lst = random_timed(tr, rng)
    dl = _abi.TimedList(ctx, d, *lst)
    off, n = ref_shells(qvec, 2, 40)
    sh = _abi.Shells(ctx, off, n)
    ell = float(np.float32(L))
    rho = _abi.Rho(ctx, d, dl, sh, 0, 38, np.float32(L))
    got = rho.download()
    natoms = np.array([len(atoms_at(lst, t)) for t in range(tr.T)])
    ref = np.zeros((tr.T, len(n)), complex)
    for t in range(tr.T):
        at = atoms_at(lst, t)
        if len(at):
            ref[t] = orc.rho_f64(tr.x[t], tr.y[t], tr.z[t], n, ell, at)
    err = np.max(np.abs(got - ref) / np.sqrt(np.maximum(natoms, 1))[:, None])
    assert err < TOL
    S = _abi.sk(ctx, rho)
    S64 = orc.sk_f64(ref, ref, natoms, off)
    assert np.max(np.abs(S - S64) / np.abs(S64)) < TOL
    F, cnt = _abi.fkt(ctx, rho, rho, P)
    F64, c64 = orc.fkt_f64(ref, ref, natoms, pairs, ntg, off)
    assert np.array_equal(cnt, c64)
    ok = c64 > 0
    assert np.max(np.abs(F[ok] - F64[ok]) / np.abs(S64)[None, :]) < TOL
    # no atoms at any frame: S = 0 as the reference sets it
    empty = _abi.TimedList(ctx, d, np.zeros(tr.T, np.int32), np.zeros(2, np.int32), np.zeros(0, np.int32))
    r0 = _abi.Rho(ctx, d, empty, sh, 0, 38, np.float32(L))
    assert np.all(_abi.sk(ctx, r0) == 0.0)


def test_batch_rejects_timed_lists(ctx, setup):
    tr, d, _, _, _ = setup
    dl = _abi.TimedList(ctx, d, *random_timed(tr, np.random.default_rng(1)))
    with pytest.raises(_abi.AmdatError):
        _abi.Batch(ctx, [d], [dl])
