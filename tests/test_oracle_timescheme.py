"""Pins for the oracle's time-scheme functions (src/system.cpp:88-141,
218, 3542-3568, 3763-3784, 4090-4179, 4599-4628): closed-form pair counts,
hand-enumerated small schemes, the weighting quirk."""
import numpy as np
import pytest

from oracle import oracle as orc


def test_hand_enumerated_scheme_fullblock0():
    # exponential 3 2 ...: blocks of 2 frames, 3 blocks, T = 7, 5 gaps.
    tg, t0, t1 = orc.ts_pairs(3, 2, 0, 0, 0)
    got = list(zip(tg.tolist(), t0.tolist(), t1.tolist()))
    assert got == [(0, 0, 0), (0, 2, 2), (0, 4, 4),
                   (1, 0, 1), (1, 2, 3), (1, 4, 5),
                   (2, 0, 2), (2, 2, 4),
                   (3, 0, 4),
                   (4, 0, 6)]


def test_hand_enumerated_scheme_fullblock1():
    tg, t0, t1 = orc.ts_pairs(3, 2, 0, 1, 0)
    got = list(zip(tg.tolist(), t0.tolist(), t1.tolist()))
    assert got == [(0, 0, 0), (0, 2, 2), (0, 4, 4),
                   (1, 0, 1), (1, 2, 3), (1, 4, 5),
                   (2, 0, 2), (2, 2, 4), (2, 1, 3), (2, 3, 5),
                   (3, 0, 4), (3, 1, 5),
                   (4, 0, 6)]


def test_limit_caps_pairs_per_gap():
    tg, t0, t1 = orc.ts_pairs(3, 2, 0, 1, 1)
    got = list(zip(tg.tolist(), t0.tolist(), t1.tolist()))
    assert got == [(0, 0, 0), (1, 0, 1), (2, 0, 2), (3, 0, 4), (4, 0, 6)]
    assert orc.ts_weighting(3, 2, 0, 1, 1).tolist() == [1, 1, 1, 1, 1]


@pytest.mark.parametrize("nexp,nes", [(1, 1), (2, 1), (5, 20), (10, 100), (4, 50), (20, 100), (7, 3)])
@pytest.mark.parametrize("fb", [0, 1])
def test_pair_count_closed_form(nexp, nes, fb):
    tg, t0, t1 = orc.ts_pairs(nexp, nes, 0, fb, 0)
    if fb == 0:
        P = nes * nexp + nexp * (nexp - 1) // 2 + 1
    else:
        P = nes * nexp + nes * nexp * (nexp - 1) // 2 + 1
    assert len(tg) == P
    T = nes * nexp + 1
    assert t0.min() >= 0 and t1.max() <= T - 1
    assert orc.ts_n_timegaps(nexp, nes, 0) == nes + nexp
    # every pair of a gap spans the same number of frames
    for g in np.unique(tg):
        assert len(np.unique((t1 - t0)[tg == g])) == 1


def test_config_counts():
    assert len(orc.ts_pairs(5, 20, 0, 0)[0]) == 111
    assert len(orc.ts_pairs(5, 20, 0, 1)[0]) == 301
    assert len(orc.ts_pairs(10, 100, 0, 0)[0]) == 1046
    assert len(orc.ts_pairs(10, 100, 0, 1)[0]) == 5501
    assert len(orc.ts_pairs(20, 100, 0, 1)[0]) == 21001


def test_weighting_quirk_plus_one():
    w = orc.ts_weighting(5, 20, 0, 0)
    assert w[:20].tolist() == [5] * 20
    assert w[20:].tolist() == [5, 4, 3, 2, 1]      # true counts are 4, 3, 2, 1, 1
    tg = orc.ts_pairs(5, 20, 0, 0)[0]
    true = np.bincount(tg, minlength=25)
    assert true[20:24].tolist() == [4, 3, 2, 1]
    w1 = orc.ts_weighting(5, 20, 0, 1)
    assert w1[20:24].tolist() == [4 * 20 + 1, 3 * 20 + 1, 2 * 20 + 1, 1 * 20 + 1]


def test_frt1_drops_final_pair():
    tg, t0, t1 = orc.ts_pairs(3, 2, 1, 0, 0)
    assert orc.ts_n_timegaps(3, 2, 1) == 4
    assert (0, 6) not in list(zip(t0.tolist(), t1.tolist()))
    assert t0.min() == 1


def test_timelist_closed_form():
    tl = orc.ts_timelist(5, 20, 1.3, 0, 1.0)
    assert len(tl) == 101
    base = float(np.float32(1.3))
    block = []
    for e in range(1, 21):
        p = base ** (e - 1)
        block.append(e if p <= e else int(np.floor(p)))
    # block times accumulate from the last time of the previous block
    exp = [0.0]
    start = 0.0
    for b in range(5):
        for v in block:
            exp.append(start + v)
        start = exp[-1]
    assert np.array_equal(tl, np.array(exp, np.float32))
    assert block[:9] == list(range(1, 10)) and block[9:13] == [10, 13, 17, 23]


def test_displacement_times():
    tl = orc.ts_timelist(3, 4, 2.0, 0, 0.5)
    dt = orc.ts_displacement_times(3, 4, 0, tl)
    # in-block gaps: t[g]-t[0]; cross-block: t[4g]-t[0]; last: t[T-1]-t[0]
    assert dt.tolist() == [tl[0], tl[1], tl[2], tl[3], tl[4], tl[8], tl[12]]
