"""Pins the oracle's rectangular binning (Trajectory_List_Bins::assign_bins)
by its defining properties: points well inside a bin map to
floor((c - lo) / w) on each axis, a point on an inner bin boundary goes to
the upper bin (the reference adds 10 FLT_EPSILON before dividing), points
on the box faces or outside the region are excluded, and the cell index is
(x ny + y) nz + z."""
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
