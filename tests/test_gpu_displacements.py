"""amdat_displacements against the oracle: per-gap sums of d and d^2 to
FP64 rounding, and the self van Hove counts bit-exact (the same FP32
reference-form |dr| and bin on both sides); an atom subset, fullblock
pairs and an empty list."""
import numpy as np
import pytest

from amdat_amd import _abi, synth
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def walk(ctx):
    print("case: displacements N=2000 box (6, 7, 8) T=31", flush=True)
    tr = synth.random_walk_box(2000, (6.0, 7.0, 8.0), 31, 0.35, seed=31)
    return tr, _abi.Trajectory(ctx, *tr.arrays)


@pytest.mark.parametrize("fb", [0, 1])
def test_sums_and_counts(ctx, walk, fb):
    tr, d = walk
    tg, t0, t1 = _abi.ts_pairs(5, 6, 0, fb, 0)
    ntg = _abi.ts_n_timegaps(5, 6, 0)
    P = _abi.Pairs(ctx, tg, t0, t1, ntg)
    sd, sd2, cnt, hist = _abi.displacements(ctx, d, None, P, n_bins=300, max_range=3.0)
    osd, osd2, ocnt = orc.disp_f64(tr, (tg, t0, t1), ntg)
    ohist, w = orc.vhs_counts(tr, (tg, t0, t1), ntg, 300, 3.0)
    assert np.array_equal(cnt, ocnt) and np.array_equal(cnt, w)
    scale = np.sqrt(np.maximum(osd2, 1e-300)) * np.sqrt(ocnt)[:, None]
    assert np.max(np.abs(sd - osd) / np.maximum(scale, 1e-12)) < 1e-12
    assert np.max(np.abs(sd2 - osd2) / np.maximum(osd2, 1e-300)) < 1e-12
    assert sd2[0].max() == 0.0
    assert np.array_equal(hist, ohist)
    print("counts in range: %d of %d" % (hist.sum(), cnt.sum()))


def test_subset_and_empty(ctx, walk):
    tr, d = walk
    tg, t0, t1 = _abi.ts_pairs(5, 6, 0, 0, 0)
    ntg = _abi.ts_n_timegaps(5, 6, 0)
    P = _abi.Pairs(ctx, tg, t0, t1, ntg)
    idx = np.arange(3, 2000, 7).astype(np.int32)
    Lst = _abi.AtomList(ctx, d, idx)
    sd, sd2, cnt, hist = _abi.displacements(ctx, d, Lst, P, n_bins=64, max_range=1.5)
    osd, osd2, ocnt = orc.disp_f64(tr, (tg, t0, t1), ntg, atoms=idx)
    ohist, _ = orc.vhs_counts(tr, (tg, t0, t1), ntg, 64, 1.5, atoms=idx)
    assert np.array_equal(cnt, ocnt)
    assert np.max(np.abs(sd2 - osd2) / np.maximum(osd2, 1e-300)) < 1e-12
    assert np.array_equal(hist, ohist)
    E = _abi.AtomList(ctx, d, np.zeros(0, np.int32))
    sd, sd2, cnt, hist = _abi.displacements(ctx, d, E, P, n_bins=8, max_range=1.0)
    assert np.all(sd == 0) and np.all(sd2 == 0) and np.all(cnt == 0) and np.all(hist == 0)
    with pytest.raises(_abi.AmdatError):
        _abi.displacements(ctx, d, None, P, n_bins=20000, max_range=1.0)


@pytest.mark.parametrize("g", [0, 3, 5, 6, 9, 10])
def test_squared_displacement_distribution(ctx, walk, g):
    """amdat_displacement_dist (u2dist, displacement_dist) on the pairs of
    one time gap: counts of the reference-form FP32 squared displacement,
    overflow bin included, bit-exact with the oracle; FP64 sums of |dr|^2
    and |dr|^4 against the FP64 oracle; gap pairs as the reference lists
    them (in-block, cross-block, the last gap)."""
    tr, d = walk
    tg, t0, t1 = _abi.ts_gap_pairs(5, 6, 0, 0, g)
    otg, ot0, ot1 = orc.ts_pairs(5, 6)
    ref = [(a, b) for gg, a, b in zip(otg, ot0, ot1) if gg == g]   # fullblock = 0, same gap
    assert list(zip(t0, t1)) == ref
    ntg = _abi.ts_n_timegaps(5, 6, 0)
    P = _abi.Pairs(ctx, tg, t0, t1, ntg)
    counts, s2, s4, c = _abi.displacement_dist(ctx, d, None, P, 40, 2.0)
    oc, _, _, occ = orc.displacement_dist(tr, (tg, t0, t1), 40, 2.0)
    assert np.array_equal(counts, oc) and c == occ == len(t0) * tr.N
    if g > 0:
        assert counts[-1] > 0 or g < 2                            # long gaps fill the overflow bin
    m = orc.msd_ngp_f64(tr, (tg, t0, t1), ntg)
    assert abs(s2 - m["s2"].sum()) <= 1e-12 * max(m["s2"].sum(), 1e-300)
    assert abs(s4 - m["s4"].sum()) <= 1e-12 * max(m["s4"].sum(), 1e-300)


def test_squared_displacement_distribution_lists(ctx, walk):
    tr, d = walk
    tg, t0, t1 = _abi.ts_gap_pairs(5, 6, 0, 2, 4)               # limit 2
    assert len(t0) == 2
    P = _abi.Pairs(ctx, tg, t0, t1, _abi.ts_n_timegaps(5, 6, 0))
    idx = np.arange(5, 2000, 3)[::-1].astype(np.int32)
    counts, s2, s4, c = _abi.displacement_dist(ctx, d, _abi.AtomList(ctx, d, idx), P, 25, 1.0)
    oc, _, _, occ = orc.displacement_dist(tr, (tg, t0, t1), 25, 1.0, atoms=idx)
    assert np.array_equal(counts, oc) and c == occ
    with pytest.raises(_abi.AmdatError):
        _abi.ts_gap_pairs(5, 6, 0, 0, 11)                         # past the last gap
