"""Pins the oracle's displacement selection (Fast/Slow/Average particle
lists per exponential block) on constructed trajectories whose
displacements are exactly representable, so the membership of every block
follows from the construction: Δ is a multiple of 1/4 along one axis, the
FP32 length is |Δ| exactly and the squared displacement is Δ² exactly.
Covers the squared-threshold comparison, strict and inclusive bounds, image
crossings, list order, limit, cross-block and last time gaps, time-varying
input lists and the list-time conversion."""
import numpy as np
import pytest

from amdat_amd import synth
from oracle import oracle as orc

L = 8.0
NEXP, NES = 3, 4
T = NEXP * NES + 1
# displacement steps used by the construction (multiples of 1/4)
STEPS = np.array([0.0, 0.5, 1.0, 1.25, 1.5, 2.0])


def constructed(frt=0, seed=7):
    """N = 12 atoms; atom j moves along axis j % 3 by s[j][t] (multiples of
    1/4) from a start near the box edge, so image counts change."""
    rng = np.random.default_rng(seed)
    N = 12
    s = np.zeros((N, T))
    for j in range(N):
        # a random walk over the step set; the step taken into frame t
        for t in range(1, T):
            s[j, t] = s[j, t - 1] + rng.choice(STEPS) * rng.choice([-1.0, 1.0])
    start = np.array([7.75 if j % 2 else 0.25 for j in range(N)])
    u = np.zeros((T, N, 3))
    for j in range(N):
        u[:, j, :] = [1.0 + 0.25 * j, 2.0, 3.0]
        u[:, j, j % 3] = start[j] + s[j]
    box = np.full((T, 3), L)
    x = np.empty((T, N), np.float32)
    img = np.empty((T, N), np.int16)
    arrs = []
    for c in range(3):
        fl = np.floor(u[..., c] / L)
        arrs.append(((u[..., c] - L * fl).astype(np.float32), fl.astype(np.int16)))
    tr = synth.Traj(arrs[0][0], arrs[1][0], arrs[2][0], arrs[0][1], arrs[1][1], arrs[2][1], box.astype(np.float32))
    return tr, s


def expected(s, pairs, test, order, nexp=NEXP, nes=NES, frt=0):
    """Per block: atoms of `order` (input order) passing test(Δ²)."""
    blocks = [[] for _ in range(nexp)]
    for t0, t1 in pairs:
        b = int((t0 - frt) / nes)
        for j in order:
            d2 = (s[j, t1] - s[j, t0]) ** 2
            if test(d2):
                blocks[b].append(j)
    off = np.concatenate([[0], np.cumsum([len(b) for b in blocks])])
    return off, np.array([j for b in blocks for j in b], np.int32)


def check(got, want_off, want_idx):
    conv, off, idx = got
    assert np.array_equal(off, want_off)
    assert np.array_equal(idx, want_idx)


def test_images_cross_the_box():
    tr, s = constructed()
    assert len(np.unique(tr.ix)) > 1 and len(np.unique(tr.iy)) > 1 and len(np.unique(tr.iz)) > 1


@pytest.mark.parametrize("mode,lo,hi,test", [
    (0, 1.2, 0.0, lambda d2: d2 > 1.2),
    (0, 0.3, 0.0, lambda d2: d2 > 0.3),          # squared: |Δ| = 0.5 is not fast
    (0, 0.25, 0.0, lambda d2: d2 > 0.25),        # strict bound
    (1, 0.0, 0.3, lambda d2: d2 < 0.3),
    (1, 0.0, 1.0, lambda d2: d2 < 1.0),          # strict bound
    (2, 0.25, 1.5625, lambda d2: 0.25 <= d2 <= 1.5625),  # inclusive bounds
])
def test_in_block_gap(mode, lo, hi, test):
    tr, s = constructed()
    g = 2
    pairs = [(NES * b, NES * b + g) for b in range(NEXP)]
    want = expected(s, pairs, test, range(tr.N))
    check(orc.select_displacement(tr, NEXP, NES, 0, 0, g, mode, lo, hi), *want)


def test_squared_threshold_differs_from_distance_threshold():
    tr, s = constructed()
    g = 3
    pairs = [(NES * b, NES * b + g) for b in range(NEXP)]
    _, idx = expected(s, pairs, lambda d2: d2 > 0.3, range(tr.N))
    _, idx_dist = expected(s, pairs, lambda d2: np.sqrt(d2) > 0.3, range(tr.N))
    assert len(idx) != len(idx_dist)
    assert np.array_equal(orc.select_displacement(tr, NEXP, NES, 0, 0, g, 0, 0.3, 0.0)[2], idx)


def test_cross_block_and_last_gap():
    tr, s = constructed()
    # g = 4: one block apart, blocks 0..nexp-2 only; g = 5: two blocks apart
    for g, btg in ((4, 1), (5, 2)):
        pairs = [(NES * b, NES * b + NES * btg) for b in range(NEXP - btg)]
        want = expected(s, pairs, lambda d2: d2 > 1.2, range(tr.N))
        check(orc.select_displacement(tr, NEXP, NES, 0, 0, g, 0, 1.2, 0.0), *want)
        assert want[0][-1] == want[0][-1 - btg]  # the last blocks stay empty
    want = expected(s, [(0, T - 1)], lambda d2: d2 > 1.2, range(tr.N))
    check(orc.select_displacement(tr, NEXP, NES, 0, 0, 6, 0, 1.2, 0.0), *want)
    with pytest.raises(ValueError):
        orc.select_displacement(tr, NEXP, NES, 0, 0, 7, 0, 1.2, 0.0)


def test_limit():
    tr, s = constructed()
    pairs = [(0, 2)]
    want = expected(s, pairs, lambda d2: d2 > 0.3, range(tr.N))
    check(orc.select_displacement(tr, NEXP, NES, 0, 1, 2, 0, 0.3, 0.0), *want)
    pairs = [(0, 2), (4, 6)]
    want = expected(s, pairs, lambda d2: d2 > 0.3, range(tr.N))
    check(orc.select_displacement(tr, NEXP, NES, 0, 2, 2, 0, 0.3, 0.0), *want)


def test_input_order_and_subset():
    tr, s = constructed()
    order = np.array([11, 3, 7, 0, 5, 9, 1], np.int32)
    pairs = [(NES * b, NES * b + 2) for b in range(NEXP)]
    want = expected(s, pairs, lambda d2: d2 > 0.3, order)
    check(orc.select_displacement(tr, NEXP, NES, 0, 0, 2, 0, 0.3, 0.0, in_list=order), *want)


def test_time_varying_input():
    tr, s = constructed()
    first = orc.select_displacement(tr, NEXP, NES, 0, 0, 2, 0, 0.3, 0.0)
    conv, off, idx = first
    blocks = [[] for _ in range(NEXP)]
    for b in range(NEXP):
        t0, t1 = NES * b, NES * b + 1
        for j in idx[off[b]:off[b + 1]]:
            if (s[j, t1] - s[j, t0]) ** 2 < 1.0:
                blocks[b].append(j)
    want_off = np.concatenate([[0], np.cumsum([len(b) for b in blocks])])
    want_idx = np.array([j for b in blocks for j in b], np.int32)
    check(orc.select_displacement(tr, NEXP, NES, 0, 0, 1, 1, 0.0, 1.0, in_list=first), want_off, want_idx)


def test_time_conversion():
    tr, _ = constructed()
    conv, _, _ = orc.select_displacement(tr, NEXP, NES, 0, 0, 0, 0, -1.0, 0.0)
    assert list(conv) == [0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3]  # last frame: past the list times
    conv, off, idx = orc.select_displacement(tr, NEXP, NES, 1, 0, 0, 0, -1.0, 0.0)
    assert list(conv) == [0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2]  # int() truncates toward zero
    # time gap 0 with a negative threshold: every atom in every block
    assert list(off) == [0, tr.N, 2 * tr.N, 3 * tr.N] and list(idx) == list(range(tr.N)) * 3


def test_frt_pairs():
    tr, s = constructed()
    # frt = 1: blocks start at nes*b + 1; n_timegaps = nes + nexp - 1 = 6;
    # g = 6 is the last gap (0, T-1), g = 7 is out of range
    pairs = [(NES * b + 1, NES * b + 1 + 2) for b in range(NEXP)]
    want = expected(s, pairs, lambda d2: d2 > 0.3, range(tr.N), frt=1)
    check(orc.select_displacement(tr, NEXP, NES, 1, 0, 2, 0, 0.3, 0.0), *want)
    want = expected(s, [(0, T - 1)], lambda d2: d2 > 0.3, range(tr.N), frt=1)
    check(orc.select_displacement(tr, NEXP, NES, 1, 0, 6, 0, 0.3, 0.0), *want)
    with pytest.raises(ValueError):
        orc.select_displacement(tr, NEXP, NES, 1, 0, 7, 0, 0.3, 0.0)


@pytest.mark.parametrize("g", [2, 4, 6])
def test_displacement_dist(g):
    """Squared-displacement histogram (u2dist / displacement_dist): with
    Delta^2 exact multiples of 1/16 and bin size 1/4 (max 4, 16 bins + the
    overflow bin), every count follows from the construction, bin edges
    included (Delta^2 = 0.25 falls in bin 1)."""
    tr, s = constructed()
    if g < NES:
        pairs = [(NES * b, NES * b + g) for b in range(NEXP)]
    elif g < NES + NEXP - 1:
        btg = g - NES + 1
        pairs = [(NES * b, NES * b + NES * btg) for b in range(NEXP - btg)]
    else:
        pairs = [(0, T - 1)]
    t0 = np.array([p[0] for p in pairs], np.int32)
    t1 = np.array([p[1] for p in pairs], np.int32)
    counts, s2, s4, c = orc.displacement_dist(tr, (np.full(len(t0), g, np.int32), t0, t1), 16, 4.0)
    want = np.zeros(17, np.uint64)
    d2 = np.array([(s[j, b] - s[j, a]) ** 2 for a, b in pairs for j in range(tr.N)])
    for v in d2:
        want[min(int(v / 0.25), 16)] += 1
    assert np.array_equal(counts, want)
    assert want[16] > 0                                  # the overflow bin is exercised
    assert c == len(d2) and s2 == d2.sum() and s4 == (d2 ** 2).sum()
    # a subset in a given order counts only its atoms
    sub = np.array([3, 7, 1], np.int32)
    counts, s2, s4, c = orc.displacement_dist(tr, (np.full(len(t0), g, np.int32), t0, t1), 16, 4.0, atoms=sub)
    d2s = np.array([(s[j, b] - s[j, a]) ** 2 for a, b in pairs for j in sub])
    assert c == len(d2s) and s2 == d2s.sum() and counts.sum() == len(d2s)
