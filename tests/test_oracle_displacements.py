"""Pins of the oracle's displacement sums and self van Hove counts:
the squared sums add up to the MSD oracle's second moment; a uniform
translation gives sum d = N v dt and sum d^2 = N (v dt)^2 per component in
closed form, and puts every atom pair of a gap into the one bin
int(|v dt| / bin_size); counts never exceed the weighting and equal it when
max_range exceeds every displacement; counting is blind to the image
representation (the same motion stored with other images)."""
import numpy as np

from amdat_amd import synth
from oracle import oracle as orc


def translated(N, L, T, v, seed=0):
    rng = np.random.default_rng(seed)
    u0 = rng.uniform(0, L, (N, 3))
    u = u0[None] + np.arange(T)[:, None, None] * np.asarray(v)[None, None, :]
    (x, y, z), (ix, iy, iz) = synth._store(u, L)
    return synth.Traj(x, y, z, ix, iy, iz, np.full((T, 3), L, np.float32))


def test_squares_sum_to_msd():
    tr = synth.random_walk(120, 5.0, 21, 0.3, seed=4)
    tg, t0, t1 = orc.ts_pairs(4, 5)
    ntg = orc.ts_n_timegaps(4, 5)
    sd, sd2, cnt = orc.disp_f64(tr, (tg, t0, t1), ntg)
    m = orc.msd_ngp_f64(tr, (tg, t0, t1), ntg)
    assert np.array_equal(cnt, m["count"])
    assert np.allclose(sd2.sum(1), m["s2"], rtol=1e-13, atol=1e-12)


def test_uniform_translation_closed_form():
    v = np.array([0.125, -0.25, 0.0625])       # dyadic: exact in FP32 and FP64
    tr = translated(64, 4.0, 13, v, seed=2)
    tg, t0, t1 = orc.ts_pairs(3, 4)
    ntg = orc.ts_n_timegaps(3, 4)
    sd, sd2, cnt = orc.disp_f64(tr, (tg, t0, t1), ntg)
    dt = np.zeros(ntg)
    for g, a, b in zip(tg, t0, t1):
        dt[g] = b - a                           # every pair of a gap has the same lag here
    npairs = np.bincount(tg, minlength=ntg)
    assert np.allclose(sd, (npairs * 64 * dt)[:, None] * v[None], rtol=1e-6, atol=1e-6)
    assert np.allclose(sd2, (npairs * 64 * dt ** 2)[:, None] * v[None] ** 2, rtol=1e-6, atol=1e-6)
    n_bins, rmax = 40, 4.0
    hist, w = orc.vhs_counts(tr, (tg, t0, t1), ntg, n_bins, rmax)
    assert np.array_equal(w, npairs * 64.0)
    for g in range(ntg):
        r = np.float32(np.linalg.norm(v) * dt[g])
        b = int(r / np.float32(rmax / n_bins))
        want = np.zeros(n_bins, np.uint64)
        if b < n_bins:
            want[b] = npairs[g] * 64
        # a jump of exactly a bin edge may land on either side by FP32 rounding: allow the neighbour
        got = hist[g]
        assert got.sum() == want.sum()
        assert got[b if b < n_bins else 0] == want[b if b < n_bins else 0] or got[max(b - 1, 0)] == want.sum()


def test_counts_bounded_by_weighting():
    tr = synth.random_walk(200, 6.0, 21, 0.2, seed=8)
    tg, t0, t1 = orc.ts_pairs(4, 5, 0, 1)
    ntg = orc.ts_n_timegaps(4, 5)
    hist, w = orc.vhs_counts(tr, (tg, t0, t1), ntg, 50, 0.5)
    assert np.all(hist.sum(1) <= w)
    hist2, w2 = orc.vhs_counts(tr, (tg, t0, t1), ntg, 50, 100.0)
    assert np.array_equal(hist2.sum(1).astype(float), w2)


def test_image_representation_invisible():
    tr = synth.random_walk(80, 5.0, 13, 0.4, seed=3)
    shifted = synth.Traj(tr.x, tr.y, tr.z, tr.ix + 3, tr.iy - 2, tr.iz + 1, tr.box)
    tg, t0, t1 = orc.ts_pairs(3, 4)
    ntg = orc.ts_n_timegaps(3, 4)
    a = orc.disp_f64(tr, (tg, t0, t1), ntg)
    b = orc.disp_f64(shifted, (tg, t0, t1), ntg)
    assert np.allclose(a[0], b[0], rtol=1e-12, atol=1e-9) and np.allclose(a[1], b[1], rtol=1e-12)
