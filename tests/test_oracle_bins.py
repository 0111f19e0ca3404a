"""Pins the oracle's rectangular binning (Trajectory_List_Bins::assign_bins)
by its defining properties: points well inside a bin map to
floor((c - lo) / w) on each axis, a point on an inner bin boundary goes to
the upper bin (the reference adds 10 FLT_EPSILON before dividing), points
on the box faces or outside the region are excluded, and the cell index is
(x ny + y) nz + z.  The distance bins (point, plane, nearest atom of a
second list) against FP64 minimum-image geometry computed here, away from
bin edges: the point and average-plane forms are the minimum-image
distance, above and below the distance upward and downward through the
periodic box (summing to L), the list form the brute-force nearest
neighbour in the box of frame 0; the bin cap and list membership."""
import numpy as np

from amdat_amd import synth
from oracle import oracle as orc


def single_frame(points):
    p = np.asarray(points, np.float32)
    z = np.zeros((1, len(p)), np.int16)
    return synth.Traj(p[None, :, 0].copy(), p[None, :, 1].copy(), p[None, :, 2].copy(), z, z, z,
                      np.full((1, 3), 8.0, np.float32))


def test_interior_points_floor_rule():
    rng = np.random.default_rng(4)
    nb = (4, 2, 5)
    w = np.array([8.0 / 4, 8.0 / 2, 8.0 / 5])
    pts = []
    for _ in range(300):
        cell = [rng.integers(0, n) for n in nb]
        frac = rng.uniform(0.05, 0.95, 3)              # well inside the cell
        pts.append([(c + f) * wi for c, f, wi in zip(cell, frac, w)])
    tr = single_frame(pts)
    lo, hi = np.zeros((1, 3)), np.full((1, 3), 8.0)
    cell = orc.bins_assign(tr, nb, lo, hi)[0]
    p = np.asarray(pts)
    want = (np.floor(p[:, 0] / w[0]) * nb[1] + np.floor(p[:, 1] / w[1])) * nb[2] + np.floor(p[:, 2] / w[2])
    assert np.array_equal(cell, want.astype(np.int32))


def test_boundaries_faces_and_regions():
    # bin width 2 on x (4 bins); y, z one bin
    pts = [[2.0, 4.0, 4.0],      # inner boundary: upper bin
           [4.0, 4.0, 4.0],
           [0.0, 4.0, 4.0],      # on the lower face: outside
           [8.0, 4.0, 4.0],      # on the upper face: outside
           [-0.5, 4.0, 4.0],     # outside
           [1.0, 4.0, 4.0],
           [7.5, 4.0, 4.0]]
    tr = single_frame(pts)
    cell = orc.bins_assign(tr, (4, 1, 1), np.zeros((1, 3)), np.full((1, 3), 8.0))[0]
    assert list(cell) == [1, 2, -1, -1, -1, 0, 3]
    # a region [2, 6] on x: points at 2 and 6 are on its faces
    pts = [[2.0, 4.0, 4.0], [2.5, 4.0, 4.0], [5.9, 4.0, 4.0], [6.0, 4.0, 4.0], [4.0, 4.0, 4.0]]
    tr = single_frame(pts)
    cell = orc.bins_assign(tr, (2, 1, 1), np.array([[2.0, 0.0, 0.0]]), np.array([[6.0, 8.0, 8.0]]))[0]
    assert list(cell) == [-1, 0, 1, -1, 1]


def _min_image(d, L):
    return d - L * np.round(d / L)


def _away_from_edges(dist, w, tol=1e-4):
    f = dist / w
    return np.abs(f - np.round(f)) > tol


def test_distance_point_is_minimum_image():
    tr = synth.plain_config()
    L = tr.box.astype(np.float64)[:, None, :]
    w = 0.37
    for p in [(4.0, 4.0, 4.0), (0.5, 9.0, 2.25), (0.0, 0.0, 0.0)]:
        cell = orc.bins_distance(tr, 0, w, 100, point=p)
        r = np.stack([tr.x, tr.y, tr.z], -1).astype(np.float64)
        d = np.linalg.norm(_min_image(r - np.asarray(p, np.float64), L), axis=-1)
        ok = _away_from_edges(d, w)
        assert ok.mean() > 0.99
        assert np.array_equal(cell[ok], np.floor(d[ok] / w).astype(np.int32))
        # the reach cap: bins at or beyond n_bins are not binned
        capped = orc.bins_distance(tr, 0, w, 5, point=p)
        assert np.array_equal(capped, np.where(cell < 5, cell, -1))


def test_distance_plane_directions():
    tr = synth.plain_config()
    w = 0.29
    for axis in range(3):
        c = [tr.x, tr.y, tr.z][axis].astype(np.float64)
        L = tr.box[:, axis].astype(np.float64)[:, None]
        for pos in (0.0, 3.3, 9.0):
            d = c - pos
            want = {0: np.abs(_min_image(d, L)), 1: np.mod(d, L), 2: np.mod(-d, L)}
            got = {k: orc.bins_distance(tr, 1, w, 1000, point=(pos,), axis=axis, direction=k) for k in range(3)}
            for k in range(3):
                ok = _away_from_edges(want[k], w) & (np.abs(d) > 1e-3)
                assert ok.mean() > 0.99
                assert np.array_equal(got[k][ok], np.floor(want[k][ok] / w).astype(np.int32)), (axis, pos, k)
    # above and below sum to L: with w = L / n the two bins sum to n - 1
    n = 10
    Lx = float(tr.box[0, 0])
    a = orc.bins_distance(tr, 1, Lx / n, n, point=(3.3,), axis=0, direction=1)
    b = orc.bins_distance(tr, 1, Lx / n, n, point=(3.3,), axis=0, direction=2)
    d = np.mod(tr.x.astype(np.float64) - 3.3, Lx) / (Lx / n)
    ok = np.abs(d - np.round(d)) > 1e-3
    assert np.all(a[ok] + b[ok] == n - 1)


def test_distance_list_nearest_neighbour_frame0_box():
    # the box grows from 10 to 12; an atom 5.2 from its only neighbour along x
    # is 4.8 away through the frame-0 box (5.2 > 10/2) at every frame
    T, N = 3, 40
    rng = np.random.default_rng(9)
    box = np.array([[10.0] * 3, [11.0] * 3, [12.0] * 3], np.float32)
    p = rng.uniform(0, 10, (T, N, 3)).astype(np.float32)
    p[:, 0] = (1.0, 5.0, 5.0)
    p[:, 1] = (6.2, 5.0, 5.0)
    z = np.zeros((T, N), np.int16)
    tr = synth.Traj(p[..., 0].copy(), p[..., 1].copy(), p[..., 2].copy(), z, z, z, box)
    w = 0.1
    cell = orc.bins_distance(tr, 2, w, 1000, binned=np.array([0]), cluster=np.array([1]))
    assert np.all(cell[:, 0] == 48) and np.all(cell[:, 1:] == -1)
    # brute force over random lists
    binned, cluster = np.arange(0, 25), np.arange(15, 40)          # overlap: distance 0 for 15..24
    cell = orc.bins_distance(tr, 2, w, 1000, binned=binned, cluster=cluster)
    for t in range(T):
        r = p[t].astype(np.float64)
        d = np.linalg.norm(_min_image(r[binned][:, None] - r[cluster][None], box[0].astype(np.float64)), axis=-1)
        m = d.min(axis=1)
        ok = _away_from_edges(m, w) | (m == 0)
        assert np.array_equal(cell[t, binned][ok], np.floor(m[ok] / w).astype(np.int32))
        assert np.all(cell[t, 15:25] == 0)
        assert np.all(cell[t, 25:] == -1)                           # not in the binned list


def test_distance_time_varying_membership():
    tr = synth.plain_config()
    per = {t: np.arange(t % 7, tr.N, 3) for t in range(tr.T)}
    cell = orc.bins_distance(tr, 0, 0.5, 100, point=(4.0, 4.0, 4.0), binned=lambda t: per[t])
    full = orc.bins_distance(tr, 0, 0.5, 100, point=(4.0, 4.0, 4.0))
    for t in (0, 5, 100):
        mask = np.zeros(tr.N, bool)
        mask[per[t]] = True
        assert np.array_equal(cell[t][mask], full[t][mask]) and np.all(cell[t][~mask] == -1)
