"""amdat_msd_ngp_noalloc, the entry point that allocates nothing (its
partial sums live in the library's per-GPU scratch array), against
amdat_msd_ngp: the same bits on two calls in a row, the second with a
smaller input, and a clear error past the scratch's size or for lists it
does not take."""
import numpy as np
import pytest

from amdat_amd import _abi, synth

pytestmark = pytest.mark.gpu


def _check_same(ctx, d, lst, P):
    a2, a4, ac = _abi.msd_ngp(ctx, d, lst, P)
    b2, b4, bc = _abi.msd_ngp_noalloc(ctx, d, lst, P)
    assert np.array_equal(ac, bc)
    assert np.array_equal(a2, b2) and np.array_equal(a4, b4)
    # lag 0 is exactly 0; every gap with atom pairs has a positive sum
    assert a2[0] == 0.0 and np.all(a2[1:][ac[1:] > 0] > 0.0)


def test_noalloc_matches_msd_ngp_twice(ctx):
    print("case: no-alloc MSD, N=5000 T=101 all pairs, then a 715-atom list and the pairs with t1 <= 50", flush=True)
    tr = synth.random_walk(5000, 16.0, 101, 0.05, seed=61)
    d = _abi.Trajectory(ctx, *tr.arrays)
    tg, t0, t1 = _abi.ts_pairs(4, 25, 0, 1, 0)
    ntg = _abi.ts_n_timegaps(4, 25, 0)
    _check_same(ctx, d, None, _abi.Pairs(ctx, tg, t0, t1, ntg))
    # second call, smaller: an atom subset and the pairs of the first two blocks
    sub = _abi.AtomList(ctx, d, np.arange(0, 5000, 7, dtype=np.int32))
    keep = t1 <= 50
    _check_same(ctx, d, sub, _abi.Pairs(ctx, tg[keep], t0[keep], t1[keep], ntg))


def test_noalloc_limits(ctx):
    tr = synth.random_walk(2048, 12.0, 3, 0.05, seed=62)
    d = _abi.Trajectory(ctx, *tr.arrays)
    # 2,048 atoms: 2 atom blocks, 32 B per pair; 1,100,000 pairs exceed 32 MiB
    n = 1100000
    P = _abi.Pairs(ctx, np.zeros(n, np.int32), np.zeros(n, np.int32), np.ones(n, np.int32), 1)
    with pytest.raises(_abi.AmdatError) as e:
        _abi.msd_ngp_noalloc(ctx, d, None, P)
    assert e.value.code == -4 and "MiB scratch" in str(e.value)
    timed = _abi.TimedList(ctx, d, [0, 0, 0], [0, 3], [0, 1, 2])
    P1 = _abi.Pairs(ctx, np.array([0], np.int32), np.array([0], np.int32), np.array([1], np.int32), 1)
    with pytest.raises(_abi.AmdatError) as e:
        _abi.msd_ngp_noalloc(ctx, d, timed, P1)
    assert e.value.code == -1
    # empty list: zeros and counts of zero, no launch
    empty = _abi.AtomList(ctx, d, np.zeros(0, np.int32))
    s2, s4, c = _abi.msd_ngp_noalloc(ctx, d, empty, P1)
    assert np.all(s2 == 0) and np.all(s4 == 0) and np.all(c == 0)
