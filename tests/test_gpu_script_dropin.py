"""The C++ drop-in (bin/amdat_hip) runs an AMDAT input script with the
msd, ngp, isfs, isf and structure_factor commands on a LAMMPS custom dump
and writes the reference's file formats; values are checked against the
FP64 oracle to the 6 significant digits the writers print."""
import os
import subprocess

import numpy as np
import pytest

from amdat_amd import _abi, synth
from oracle import oracle as orc
from tests.conftest import ROOT, ref_shells

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "amdat_amd", "bin", "amdat_hip")
RTOL = 2e-5   # 6 significant digits in the text output


def read_rows(path, skip):
    rows = []
    with open(path) as fh:
        lines = fh.read().split("\n")
    for l in lines[skip:]:
        if l.strip():
            rows.append([float(v) for v in l.strip("\t").split("\t")])
    return lines, np.array(rows)


@pytest.fixture(scope="module")
def run(tmp_path_factory, qvec):
    d = tmp_path_factory.mktemp("dropin")
    off, n = qvec
    qdir = d / "qvectors3d"
    qdir.mkdir()
    for b in range(2, 101):
        with open(qdir / ("qvector.%03d" % b), "w", newline="") as fh:
            for v in n[off[b - 1]:off[b]]:
                fh.write("  %.7e   %.7e   %.7e\r\n" % tuple(v))
    tr = synth.plain_config(N=300, L=7.0, nexp=2, nes=3, sigma=0.08, seed=41)
    synth.write_lammps_custom(tr, str(d / "traj.dump"))
    script = """# drop-in check
system
custom
traj.dump
exponential 2 3 1.3 0 0 1
A 240 B 60
1 2
1 0
0 1

create_list all
all
create_list big
type_system 1
msd msd.dat
list all
ngp ngp.dat \\sstored
list all
isfs isfs.dat 5 9 xyz 3.5 0
list all
isf isf.dat auto symmetric xyz 0 0 20
list all
structure_factor sk.dat symmetric xyz 0 -1
list all
structure_factor sk_ab.dat asymmetric xyz 0 -1
list big
list all
structure_factor sk0.dat symmetric xyz 0
list all
isf isf_ab.dat auto asymmetric xyz 0 0 20
list big
list all
isf isf_man.dat manual symmetric xyz 1 0 sk_man.dat
list all
"""
    (d / "in.amdat").write_text(script)
    env = dict(os.environ, AMDAT_WV3D=str(qdir / "qvector"))
    print("case: amdat_hip drop-in script, N=300 T=7", flush=True)
    # isf manual reads its wavenumber bounds from standard input
    p = subprocess.run([BIN, "-i", "in.amdat"], cwd=str(d), env=env, capture_output=True, text=True, timeout=300,
                       input="3 8\n")
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]
    (d / "stdout.txt").write_text(p.stdout)
    return d, tr


def test_msd_ngp_files(run):
    d, tr = run
    tg, t0, t1 = orc.ts_pairs(2, 3)
    m = orc.msd_ngp_f64(tr, (tg, t0, t1), 5)
    tl = orc.ts_timelist(2, 3, np.float32(1.3), 0, 1.0)
    dt = orc.ts_displacement_times(2, 3, 0, tl)
    lines, rows = read_rows(d / "msd.dat", 1)
    assert lines[0].startswith("Mean square displacement data created by AMDAT v.")
    assert np.array_equal(rows[:, 0], dt)
    assert rows[0, 1] == 0.0
    assert np.allclose(rows[1:, 1], m["msd"][1:], rtol=RTOL)
    lines, rows = read_rows(d / "ngp.dat", 1)
    assert lines[0].startswith("Non-Gaussian parameter data created by AMDAT v.")
    assert np.isnan(rows[0, 1])
    assert np.allclose(rows[1:, 1], m["ngp"][1:], rtol=1e-4, atol=1e-5)


def test_isfs_file(run, qvec):
    d, tr = run
    off, n = ref_shells(qvec, 7, 11)          # wn 5..9
    tg, t0, t1 = orc.ts_pairs(2, 3)
    ell = float(np.float32(3.5))
    Fs, _ = orc.fs_f64(tr, (tg, t0, t1), 5, off, n, ell)
    lines, rows = read_rows(d / "isfs.dat", 2)
    assert lines[0].startswith("Correlation data created by AMDAT v.")
    k = [float(v) for v in lines[1].strip("\t").split("\t")]
    assert np.allclose(k, orc.approx_wavenumbers(np.float32(3.5), 99)[5:10], rtol=RTOL)
    assert np.allclose(rows[:, 1:], Fs, rtol=RTOL, atol=1e-6)


def test_isf_file(run, qvec):
    d, tr = run
    off, n = ref_shells(qvec, 2, 22)          # wn 0..20
    L = float(np.float32(7.0))
    rho = orc.rho_f64_traj(tr, n, L)
    tg, t0, t1 = orc.ts_pairs(2, 3)
    F, _ = orc.fkt_f64(rho, rho, np.full(tr.T, tr.N), (tg, t0, t1), 5, off)
    lines, rows = read_rows(d / "isf.dat", 2)
    assert rows.shape == (5, 22)
    assert np.allclose(rows[:, 1:], F, rtol=RTOL, atol=1e-6)


def test_structure_factor_files(run, qvec):
    d, tr = run
    off, n = ref_shells(qvec, 2, 100)
    L = float(np.float32(7.0))
    rho = orc.rho_f64_traj(tr, n, L)
    S = orc.sk_f64(rho, rho, np.full(tr.T, tr.N), off)
    lines, rows = read_rows(d / "sk.dat", 2)
    assert lines[1] == "Nominal_q\tMean_q\tStddev_q\tMean_qx\tMean_qy\tMean_qz\tS(q)"
    assert rows.shape == (99, 7)
    assert np.allclose(rows[:, 6], S, rtol=RTOL)
    assert np.allclose(rows[:, 0], orc.approx_wavenumbers(np.float32(7.0), 99), rtol=RTOL)
    kmag = np.array([np.mean(np.linalg.norm(n[off[s]:off[s + 1]] * 2 * np.pi / L, axis=1)) for s in range(99)])
    assert np.allclose(rows[:, 1], kmag, rtol=1e-5)
    # asymmetric: Re(rho_big conj rho_all) / N_big
    big = np.arange(240)
    rb = orc.rho_f64_traj(tr, n, L, atoms=big)
    Sab = orc.sk_f64(rb, rho, np.full(tr.T, 240), off)
    _, rows = read_rows(d / "sk_ab.dat", 2)
    assert np.allclose(rows[:, 6], Sab, rtol=RTOL, atol=1e-6)
    # 5-token structure_factor: timescheme 0 selects frame 0 only (block-count bound)
    r0 = rho[:1]
    S0 = orc.sk_f64(r0, r0, np.array([tr.N]), off)
    _, rows = read_rows(d / "sk0.dat", 2)
    assert np.allclose(rows[:, 6], S0, rtol=RTOL)


def test_isf_asymmetric_file(run, qvec):
    """isf auto asymmetric (src/control.cpp:3145-3152): rho of the first
    list at t0 against rho of the second at t1, normalised by the first
    list's atoms (src/intermediate_scattering_function.cpp:551-566)."""
    d, tr = run
    off, n = ref_shells(qvec, 2, 22)          # wn 0..20
    L = float(np.float32(7.0))
    big = np.arange(240)
    rb = orc.rho_f64_traj(tr, n, L, atoms=big)
    ra = orc.rho_f64_traj(tr, n, L)
    tg, t0, t1 = orc.ts_pairs(2, 3)
    F, _ = orc.fkt_f64(rb, ra, np.full(tr.T, 240), (tg, t0, t1), 5, off)
    S = orc.sk_f64(rb, rb, np.full(tr.T, 240), off)
    _, rows = read_rows(d / "isf_ab.dat", 2)
    assert rows.shape == (5, 22)
    assert np.allclose(rows[:, 1:], F, rtol=RTOL, atol=1e-6 * np.max(np.abs(S)))


def test_isf_manual_files(run, qvec):
    """isf manual (src/control.cpp:3154-3189): the S(k) file from the time
    gap 0 correlation over every shell (fullblock pairs of gap 0), then the
    bounds read from standard input ("3 8"); the reference would read the
    densities' slots 0..5 with the vector counts of shells 3..8, past those
    slots, so shells 3..8 are computed, with a warning."""
    d, tr = run
    off, n = ref_shells(qvec, 2, 100)
    L = float(np.float32(7.0))
    rho = orc.rho_f64_traj(tr, n, L)
    nat = np.full(tr.T, tr.N)
    tg, t0, t1 = orc.ts_pairs(2, 3, fullblock=1)
    sel = tg == 0
    G, _ = orc.fkt_f64(rho, rho, nat, (tg[sel], t0[sel], t1[sel]), 5, off)
    lines, rows = read_rows(d / "sk_man.dat", 2)
    assert lines[0].startswith("Correlation data created by AMDAT v.")
    assert rows.shape == (5, 100)
    assert np.allclose(rows[0, 1:], G[0], rtol=RTOL)
    assert np.all(rows[1:, 1:] == 0)
    F, _ = orc.fkt_f64(rho, rho, nat, orc.ts_pairs(2, 3, fullblock=1), 5, off)
    _, rows = read_rows(d / "isf_man.dat", 2)
    assert rows.shape == (5, 7)
    assert np.allclose(rows[:, 1:], F[:, 3:9], rtol=RTOL, atol=1e-6)
    out = (d / "stdout.txt").read_text()
    assert "Please input the lower and upper wavenumber index bounds" in out
    assert "Warning" in out
