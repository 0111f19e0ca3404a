"""The drop-in driver's msd_2d, mean_displacement, vhs, u2dist and
displacement_dist commands (the reference's file formats) against the
oracle, to the 6 significant digits the writers print."""
import os
import subprocess

import numpy as np
import pytest

from amdat_amd import synth
from oracle import oracle as orc
from tests.conftest import ROOT

pytestmark = pytest.mark.gpu

RTOL = 2e-5


def rows_of(path, skip):
    lines = open(path).read().split("\n")
    return lines, np.array([[float(v) for v in l.strip("\t").split("\t")] for l in lines[skip:] if l.strip()])


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    d = tmp_path_factory.mktemp("disp")
    tr = synth.random_walk_box(400, (6.0, 7.0, 8.0), 21, 0.25, seed=17)
    # a drift so the mean displacement is not just noise
    u = tr.unwrapped64() + np.arange(tr.T)[:, None, None] * np.array([0.05, -0.02, 0.01])[None, None]
    (x, y, z), (ix, iy, iz) = synth._store(u, np.array([6.0, 7.0, 8.0]))
    tr = synth.Traj(x, y, z, ix, iy, iz, tr.box)
    synth.write_lammps_custom(tr, str(d / "traj.dump"))
    script = """system
custom
traj.dump
exponential 4 5 1.3 0 0 1
A 400
1
1

create_list all
all
msd_2d msd_xz.dat xz
list all
mean_displacement md.dat
list all
vhs vhs.dat 3.0 60
list all
u2dist u2.dat 30 1.5 6
list all
displacement_dist dd.dat 4 20 0.8 2
list all
"""
    (d / "in.amdat").write_text(script)
    print("case: amdat_hip msd_2d / mean_displacement / vhs, N=400 T=21", flush=True)
    p = subprocess.run([os.path.join(ROOT, "amdat_amd", "bin", "amdat_hip"), "-i", "in.amdat"], cwd=str(d),
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]
    return d, tr


def pairs():
    tg, t0, t1 = orc.ts_pairs(4, 5)
    return (tg, t0, t1), orc.ts_n_timegaps(4, 5)


def test_msd_2d(run):
    d, tr = run
    P, ntg = pairs()
    sd, sd2, cnt = orc.disp_f64(tr, P, ntg)
    lines, rows = rows_of(d / "msd_xz.dat", 1)
    assert lines[0].startswith("2-D mean square displacement data for xz plane created by AMDAT v.")
    want = (sd2[:, 0] + sd2[:, 2]) / cnt
    assert rows[0, 1] == 0.0
    assert np.allclose(rows[1:, 1], want[1:], rtol=RTOL)


def test_mean_displacement(run):
    d, tr = run
    P, ntg = pairs()
    sd, sd2, cnt = orc.disp_f64(tr, P, ntg)
    lines, rows = rows_of(d / "md.dat", 1)
    assert lines[0].startswith("Mean displacement data created by AMDAT v.")
    assert np.allclose(rows[:, 1:4], sd / cnt[:, None], rtol=RTOL, atol=1e-9)


def test_vhs(run):
    d, tr = run
    P, ntg = pairs()
    hist, w = orc.vhs_counts(tr, P, ntg, 60, 3.0)
    bs = np.float32(3.0) / np.float32(60)
    shell = np.array([np.float32((4.0 / 3.0) * np.pi * ((np.float32(b * bs) + bs) ** 3.0 - np.float32(b * bs) ** 3.0))
                      for b in range(60)], np.float64)
    want = hist / (w[:, None] * shell[None])
    lines, rows = rows_of(d / "vhs.dat", 5)
    assert lines[0].startswith("Correlation data created by AMDAT v.") and lines[1] == "60 bins"
    centers = [float(v) for v in lines[4].strip("\t").split("\t")]
    assert np.allclose(centers, bs / 2 + np.arange(60) * bs, rtol=1e-6)
    assert rows.shape == (ntg, 61)
    assert np.allclose(rows[:, 1:], want, rtol=RTOL, atol=1e-12)


def _dist_rows(path):
    lines = open(path).read().split("\n")
    return lines


def _gap_pairs(g):
    tg, t0, t1 = orc.ts_pairs(4, 5)
    sel = tg == g
    return tg[sel], t0[sel], t1[sel]


@pytest.mark.parametrize("name,g,bins,maxv", [("u2.dat", 6, 30, 1.5), ("dd.dat", 2, 20, 0.8)])
def test_squared_displacement_distributions(run, name, g, bins, maxv):
    d, tr = run
    counts, s2, s4, c = orc.displacement_dist(tr, _gap_pairs(g), bins, maxv)
    mean = s2 / c
    var = s4 / c - mean * mean
    lines = _dist_rows(d / name)
    tl = orc.ts_timelist(4, 5, np.float32(1.3), 0, 1.0)
    time = orc.ts_displacement_times(4, 5, 0, tl)[g]
    if name == "u2.dat":
        assert lines[0].startswith("MSD distribution created by AMDAT v.")
        assert np.isclose(float(lines[1].split(" = ")[1]), time, rtol=RTOL)
        assert lines[2].startswith("Mean_stiffness\t") and lines[3].startswith("Variance\t")
        body, tail = lines[5:5 + bins + 1], lines[5 + bins + 1]
        assert tail == "Last bin includes overflow."
    else:
        assert lines[0] == "Particle displacement distribution to the 4 power created by AMDAT v." + lines[0].split("v.")[1]
        assert lines[1].startswith("Last bin includes overflow.time = ")
        assert np.isclose(float(lines[1].split(" = ")[1]), time, rtol=RTOL)
        assert lines[2].startswith("Mean\t") and lines[3].startswith("Variance\t")
        body = lines[5:5 + bins + 1]
    assert np.isclose(float(lines[2].split("\t")[1]), mean, rtol=RTOL)
    assert np.isclose(float(lines[3].split("\t")[1]), var, rtol=1e-4)
    rows = np.array([[float(v) for v in l.split("\t")] for l in body])
    binsize = np.float32(maxv) / np.float32(bins)
    assert np.allclose(rows[:, 0], (np.arange(bins + 1) + 0.5) * binsize, rtol=RTOL)
    assert np.allclose(rows[:, 1], counts / c, rtol=RTOL, atol=1e-9)
    assert counts[-1] > 0                                        # the overflow bin is in the file
