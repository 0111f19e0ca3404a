"""The drop-in driver on the reference's 2D and 1D planes: isfs in xz,
isf in xy, structure_factor in yz and z, on an orthorhombic box so each
plane's default length scale (src/wave_vectors.cpp:85-140) is exercised.
Expected values come from the FP64 oracle on the integer vectors placed as
src/wave_vectors.cpp:266-424 places them (1D: 2n at twice the length scale),
to the 6 significant digits the writers print."""
import os
import subprocess

import numpy as np
import pytest

from amdat_amd import synth
from oracle import oracle as orc
from tests.conftest import ROOT, write_plane_files

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "amdat_amd", "bin", "amdat_hip")
RTOL = 2e-5
BOX = (6.0, 7.5, 9.0)


def read_rows(path, skip):
    with open(path) as fh:
        lines = fh.read().split("\n")
    rows = [[float(v) for v in l.strip("\t").split("\t")] for l in lines[skip:] if l.strip()]
    return lines, np.array(rows)


def place(raw, perm):
    out = np.zeros_like(raw)
    for col, axis in enumerate(perm):
        out[:, axis] = raw[:, col]
    return out


def shells2d(qvec2d, first_bin, last_bin, perm):
    off, n = qvec2d
    rows = n[off[first_bin]:off[last_bin + 1]]
    so = (off[first_bin:last_bin + 2] - off[first_bin]).astype(np.int32)
    return so, place(rows, perm).astype(np.int32)


@pytest.fixture(scope="module")
def run(tmp_path_factory, qvec2d, qvec1d):
    d = tmp_path_factory.mktemp("planes")
    stems = write_plane_files(str(d), qvec2d, qvec1d)
    tr = synth.random_walk_box(200, BOX, 7, 0.08, seed=44)
    synth.write_lammps_custom(tr, str(d / "traj.dump"))
    script = """# planes
system
custom
traj.dump
exponential 2 3 1.3 0 0 1
A 200
1
1

create_list all
all
isfs isfs_xz.dat 5 9 xz 3.5 0
list all
isf isf_xy.dat auto symmetric xy 0 0 20
list all
structure_factor sk_yz.dat symmetric yz 0 -1
list all
structure_factor sk_z.dat symmetric z 0 -1
list all
"""
    (d / "in.amdat").write_text(script)
    env = dict(os.environ, AMDAT_WV2D=stems["2d"], AMDAT_WV1D=stems["1d"])
    print("case: amdat_hip planes script, N=200 T=7 box", BOX, flush=True)
    p = subprocess.run([BIN, "-i", "in.amdat"], cwd=str(d), env=env, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]
    return d, tr


def test_isfs_xz(run, qvec2d):
    d, tr = run
    off, n = shells2d(qvec2d, 7, 11, (0, 2, 1))        # wn 5..9
    tg, t0, t1 = orc.ts_pairs(2, 3)
    Fs, _ = orc.fs_f64(tr, (tg, t0, t1), 5, off, n, float(np.float32(3.5)))
    _, rows = read_rows(d / "isfs_xz.dat", 2)
    assert rows.shape == (5, 6)
    assert np.allclose(rows[:, 1:], Fs, rtol=RTOL, atol=1e-6)


def test_isf_xy(run, qvec2d):
    d, tr = run
    off, n = shells2d(qvec2d, 2, 22, (0, 1, 2))        # wn 0..20
    L = float(np.float32(min(BOX[0], BOX[1])))
    rho = orc.rho_f64_traj(tr, n, L)
    tg, t0, t1 = orc.ts_pairs(2, 3)
    F, _ = orc.fkt_f64(rho, rho, np.full(tr.T, tr.N), (tg, t0, t1), 5, off)
    _, rows = read_rows(d / "isf_xy.dat", 2)
    assert rows.shape == (5, 22)
    assert np.allclose(rows[:, 1:], F, rtol=RTOL, atol=1e-6)


def test_structure_factor_yz(run, qvec2d):
    d, tr = run
    off, n = shells2d(qvec2d, 2, 100, (1, 2, 0))
    L = float(np.float32(min(BOX[1], BOX[2])))
    rho = orc.rho_f64_traj(tr, n, L)
    S = orc.sk_f64(rho, rho, np.full(tr.T, tr.N), off)
    _, rows = read_rows(d / "sk_yz.dat", 2)
    assert rows.shape == (99, 7)
    assert np.allclose(rows[:, 6], S, rtol=RTOL, atol=1e-6)
    empty = np.diff(off) == 0                          # shells without vectors print 0
    assert np.all(rows[empty, 6] == 0.0)
    assert np.allclose(rows[:, 0], orc.approx_wavenumbers(np.float32(L), 99), rtol=RTOL)


def test_structure_factor_z(run, qvec1d):
    d, tr = run
    n = np.zeros((99, 3), np.int32)
    n[:, 2] = qvec1d[1:100, 0]                         # 2 * (bin / 2)
    off = np.arange(100, dtype=np.int32)
    Lz = float(np.float32(BOX[2]))
    rho = orc.rho_f64_traj(tr, n, 2.0 * Lz)
    S = orc.sk_f64(rho, rho, np.full(tr.T, tr.N), off)
    _, rows = read_rows(d / "sk_z.dat", 2)
    assert np.allclose(rows[:, 6], S, rtol=RTOL, atol=1e-6)
    # Mean_q of a 1D shell is (bin / 2) * 2 pi / Lz
    assert np.allclose(rows[:, 1], np.arange(2, 101) / 2.0 * 2 * np.pi / Lz, rtol=1e-5)
