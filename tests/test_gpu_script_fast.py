"""The C++ drop-in runs `find_fast_fixedthreshold` (a Fast_Particles list:
atoms whose squared displacement over a time index exceeds a threshold, per
exponential block) and then msd, ngp, isfs, isf and structure_factor on
that time-varying list.  The list is checked against the oracle's selection
through the counts file, and every analysis against the FP64 oracle
composed list time by list time, to the 6 significant digits the writers
print."""
import os
import subprocess

import numpy as np
import pytest

from amdat_amd import synth
from oracle import oracle as orc
from tests.conftest import ROOT, atoms_at, compose_pairs, ref_shells

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "amdat_amd", "bin", "amdat_hip")
RTOL = 2e-5
NEXP, NES = 3, 4
NTG = NES + NEXP
L = 7.0


def read_rows(path, skip):
    with open(path) as fh:
        lines = fh.read().split("\n")
    rows = [[float(v) for v in l.strip("\t").split("\t")] for l in lines[skip:] if l.strip()]
    return lines, np.array(rows)


@pytest.fixture(scope="module")
def run(tmp_path_factory, qvec):
    d = tmp_path_factory.mktemp("dropin_fast")
    off, n = qvec
    qdir = d / "qvectors3d"
    qdir.mkdir()
    for b in range(2, 101):
        with open(qdir / ("qvector.%03d" % b), "w", newline="") as fh:
            for v in n[off[b - 1]:off[b]]:
                fh.write("  %.7e   %.7e   %.7e\r\n" % tuple(v))
    tr = synth.plain_config(N=300, L=L, nexp=NEXP, nes=NES, sigma=0.08, seed=43)
    synth.write_lammps_custom(tr, str(d / "traj.dump"))
    u = tr.unwrapped64()
    thr = float(np.float32(np.median(((u[2] - u[0]) ** 2).sum(-1))))
    script = """system
custom
traj.dump
exponential %d %d 1.3 0 0 1
A 240 B 60
1 2
1 0
0 1

create_list all
all
find_fast_fixedthreshold fast fast.count 2 %.9g
list all
msd msd.dat
list fast
ngp ngp.dat
list fast
isfs isfs.dat 5 9 xyz 3.5 0
list fast
isf isf.dat auto symmetric xyz 0 0 20
list fast
structure_factor sk.dat symmetric xyz 0 -1
list fast
""" % (NEXP, NES, thr)
    (d / "in.amdat").write_text(script)
    env = dict(os.environ, AMDAT_WV3D=str(qdir / "qvector"))
    print("case: amdat_hip find_fast_fixedthreshold script, N=300 T=13", flush=True)
    p = subprocess.run([BIN, "-i", "in.amdat"], cwd=str(d), env=env, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]
    lst = orc.select_displacement(tr, NEXP, NES, 0, 0, 2, 0, np.float32(thr), 0.0)
    return d, tr, lst


def test_counts_file(run):
    d, tr, lst = run
    conv, off, idx = lst
    counts = np.diff(off)
    assert np.all(counts > 0) and np.all(counts < tr.N)
    lines = (d / "fast.count").read_text().split("\n")
    assert lines[0].startswith("Average_trajectories:\t")
    assert np.isclose(float(lines[0].split("\t")[1]), counts.mean(), rtol=RTOL)
    assert lines[1] == "Count_List"
    tl = orc.ts_timelist(NEXP, NES, np.float32(1.3), 0, 1.0)
    rows = np.array([[float(v) for v in l.split("\t")] for l in lines[2:2 + NEXP]])
    assert np.array_equal(rows[:, 1], counts)
    assert np.allclose(rows[:, 0], tl[np.arange(NEXP) * NES], rtol=RTOL)


def test_msd_ngp(run):
    d, tr, lst = run
    pairs = orc.ts_pairs(NEXP, NES)
    s2, s4, cnt = np.zeros(NTG), np.zeros(NTG), np.zeros(NTG)
    for m in compose_pairs(lst, pairs, lambda p, at: orc.msd_ngp_f64(tr, p, NTG, atoms=at)):
        s2 += m["s2"]
        s4 += m["s4"]
        cnt += m["count"]
    _, rows = read_rows(d / "msd.dat", 1)
    assert rows[0, 1] == 0.0
    assert np.allclose(rows[1:, 1], s2[1:] / cnt[1:], rtol=RTOL)
    _, rows = read_rows(d / "ngp.dat", 1)
    ngp = s4[1:] * 3.0 / (cnt[1:] * 5.0 * (s2[1:] / cnt[1:]) ** 2) - 1.0
    assert np.allclose(rows[1:, 1], ngp, rtol=1e-4, atol=1e-5)


def test_isfs(run, qvec):
    d, tr, lst = run
    off, n = ref_shells(qvec, 7, 11)
    pairs = orc.ts_pairs(NEXP, NES)
    ell = float(np.float32(3.5))
    SF, SC = np.zeros((NTG, 5)), np.zeros(NTG)
    for f, c in compose_pairs(lst, pairs, lambda p, at: orc.fs_f64(tr, p, NTG, off, n, ell, at)):
        SF += f * c[:, None]
        SC += c
    _, rows = read_rows(d / "isfs.dat", 2)
    assert np.allclose(rows[:, 1:], SF / SC[:, None], rtol=RTOL, atol=1e-6)


def test_isf_and_structure_factor(run, qvec):
    d, tr, lst = run
    off, n = ref_shells(qvec, 2, 100)
    Lf = float(np.float32(L))
    natoms = np.array([len(atoms_at(lst, t)) for t in range(tr.T)])
    assert natoms[-1] == 0  # the last frame lies past the block list times
    rho = np.zeros((tr.T, len(n)), complex)
    for t in range(tr.T):
        if natoms[t]:
            rho[t] = orc.rho_f64(tr.x[t], tr.y[t], tr.z[t], n, Lf, atoms_at(lst, t))
    S = orc.sk_f64(rho, rho, natoms, off)
    _, rows = read_rows(d / "sk.dat", 2)
    assert np.allclose(rows[:, 6], S, rtol=RTOL)
    pairs = orc.ts_pairs(NEXP, NES)
    F, cnt = orc.fkt_f64(rho[:, :off[21]], rho[:, :off[21]], natoms, pairs, NTG, off[:22])
    _, rows = read_rows(d / "isf.dat", 2)
    assert np.allclose(rows[:, 1:], F, rtol=RTOL, atol=1e-6)
