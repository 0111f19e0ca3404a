"""The drop-in's vhs -> gaussian_comparison -> find_fast chain on a
two-population random walk (80% slow, 20% fast atoms), whose self van Hove
crosses its Gaussian approximation twice.  Checked against the oracle and
the closed forms: the written van Hove column (oracle counts x shell
volume), the Gaussian column's slow fraction against the Maxwell CDF, the
cutoffs against the crossing rule applied to the written columns, the time
index at the oracle NGP's maximum, the fast list bit-exact with the
oracle's selection at the fast boundary, and the MSD of that list."""
import math
import os
import subprocess

import numpy as np
import pytest

from amdat_amd import synth
from oracle import oracle as orc
from tests.conftest import ROOT, compose_pairs

pytestmark = pytest.mark.gpu

RTOL = 2e-5
NEXP, NES = 4, 5
NTG = NES + NEXP
L = 9.0
NBINS, RMAX = 150, 3.0


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    d = tmp_path_factory.mktemp("gauss")
    rng = np.random.default_rng(23)
    N, T = 1000, NEXP * NES + 1
    sig = np.where(np.arange(N) < 800, 0.04, 0.2)
    u = np.empty((T, N, 3))
    u[0] = rng.uniform(0, L, (N, 3))
    u[1:] = rng.normal(0.0, 1.0, (T - 1, N, 3)) * sig[None, :, None]
    u = np.cumsum(u, axis=0)
    (x, y, z), (ix, iy, iz) = synth._store(u, np.array([L, L, L]))
    tr = synth.Traj(x, y, z, ix, iy, iz, np.full((T, 3), L, np.float32))
    synth.write_lammps_custom(tr, str(d / "traj.dump"))
    script = """system
custom
traj.dump
exponential %d %d 1.3 0 0 1
A %d
1
1

create_list all
all
vhs vhs.dat %g %d
list all
gaussian_comparison gc.dat
list all
find_fast fast fast.count
list all
msd msd_fast.dat
list fast
""" % (NEXP, NES, N, RMAX, NBINS)
    (d / "in.amdat").write_text(script)
    print("case: amdat_hip vhs/gaussian_comparison/find_fast, N=1000 T=21", flush=True)
    p = subprocess.run([os.path.join(ROOT, "amdat_amd", "bin", "amdat_hip"), "-i", "in.amdat"], cwd=str(d),
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]
    return d, tr


def gc_file(d):
    lines = (d / "gc.dat").read_text().split("\n")
    assert lines[0].startswith("Gaussian comparison data created by AMDAT v.")
    out = {"t": float(lines[2]), "slow_cut": float(lines[3].split(": ")[1]),
           "fast_cut": float(lines[4].split(": ")[1])}
    out["gauss"] = [float(v) for v in lines[6].split("\t")[1:]]
    out["actual"] = [float(v) for v in lines[7].split("\t")[1:]]
    assert lines[10] == "Self_Van_Hove*shellvolume" and lines[11] == "radius\tgaussian\tactual"
    out["rows"] = np.array([[float(v) for v in l.split("\t")] for l in lines[12:12 + NBINS]])
    return out


def maxwell_cdf(r, alpha):
    return math.erf(math.sqrt(alpha) * r) - 2.0 * math.sqrt(alpha / math.pi) * r * math.exp(-alpha * r * r)


def test_gaussian_comparison(run):
    d, tr = run
    g = gc_file(d)
    pairs = orc.ts_pairs(NEXP, NES)
    m = orc.msd_ngp_f64(tr, pairs, NTG)
    gidx = int(np.argmax(m["ngp"][1:])) + 1                      # the NGP's maximum
    tl = orc.ts_timelist(NEXP, NES, np.float32(1.3), 0, 1.0)
    assert np.isclose(g["t"], orc.ts_displacement_times(NEXP, NES, 0, tl)[gidx], rtol=RTOL)
    bin_size = np.float32(RMAX) / np.float32(NBINS)
    rows = g["rows"]
    r_in, r_out = np.arange(NBINS) * float(bin_size), (np.arange(NBINS) + 1) * float(bin_size)
    shell = 4.0 / 3.0 * np.pi * (r_out ** 3 - r_in ** 3)
    hist, w = orc.vhs_counts(tr, pairs, NTG, NBINS, RMAX)
    actual = hist[gidx] / w[gidx]                                  # van Hove x shell volume
    assert np.allclose(rows[:, 2], actual, rtol=1e-4, atol=1e-6)
    alpha = 3.0 / (2.0 * m["msd"][gidx])
    gauss = (alpha / np.pi) ** 1.5 * np.exp(-alpha * ((np.arange(NBINS) + 0.5) * float(bin_size)) ** 2) * shell
    assert np.allclose(rows[:, 1], gauss, rtol=1e-4, atol=1e-7)
    # the crossing rule on the written columns
    ga, ac = rows[:, 1], rows[:, 2]
    s = next(b for b in range(1, NBINS) if ga[b] >= ac[b]) - 1
    f = next(b for b in range(s + 1, NBINS) if ga[b] <= ac[b])
    assert 0 < s < f
    assert np.isclose(g["slow_cut"], (s + 0.5) * float(bin_size), rtol=RTOL)
    assert np.isclose(g["fast_cut"], (f - 0.5) * float(bin_size), rtol=RTOL)
    # fractions: sums of the columns; the Gaussian one against the Maxwell CDF
    assert np.isclose(g["actual"][0], ac[1:s + 1].sum(), rtol=1e-4)
    assert np.isclose(g["gauss"][0], ga[1:s + 1].sum(), rtol=1e-4)
    cdf = maxwell_cdf((s + 1) * float(bin_size), alpha) - maxwell_cdf(float(bin_size), alpha)
    assert abs(g["gauss"][0] - cdf) < 2e-3
    assert np.isclose(g["actual"][1], 1.0 - ac[1:f].sum(), rtol=1e-4, atol=1e-6)


def test_find_fast_list_and_msd(run):
    d, tr = run
    g = gc_file(d)
    pairs = orc.ts_pairs(NEXP, NES)
    m = orc.msd_ngp_f64(tr, pairs, NTG)
    gidx = int(np.argmax(m["ngp"][1:])) + 1
    bin_size = np.float32(RMAX) / np.float32(NBINS)
    f = int(round(g["fast_cut"] / float(bin_size) + 0.5))
    fastboundary = np.float32(((np.float32(f) - 0.5) * np.float64(bin_size)) ** 2)
    lst = orc.select_displacement(tr, NEXP, NES, 0, 0, gidx, 0, fastboundary, 0.0)
    conv, off, idx = lst
    counts = np.diff(off)
    lines = (d / "fast.count").read_text().split("\n")
    got = np.array([int(l.split("\t")[1]) for l in lines[2:2 + NEXP]])
    assert np.array_equal(got, counts) and counts.sum() > 0
    # most fast atoms come from the fast population
    assert np.mean(idx >= 800) > 0.5
    s2, cnt = np.zeros(NTG), np.zeros(NTG)
    for mm in compose_pairs(lst, pairs, lambda p, at: orc.msd_ngp_f64(tr, p, NTG, atoms=at)):
        s2 += mm["s2"]
        cnt += mm["count"]
    rows = np.array([[float(v) for v in l.split("\t")] for l in (d / "msd_fast.dat").read_text().split("\n")[1:]
                     if l.strip()])
    assert np.allclose(rows[1:, 1], s2[1:] / cnt[1:], rtol=RTOL)
