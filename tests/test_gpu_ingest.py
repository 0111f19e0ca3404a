"""Reading trajectory files straight into device memory (amdat_traj_read):
the device arrays equal an upload of the same data (the MSD/NGP moments and
rho match bit for bit), for custom and xyz files, in one chunk and in
chunks of a few frames (the xyz unwrapping then carries across chunk
boundaries); and the drop-in driver runs an xyz system block."""
import os
import subprocess

import numpy as np
import pytest

from amdat_amd import _abi, synth
from oracle import oracle as orc
from tests.conftest import ROOT, ref_shells
from tests.test_ingest_cpu import write_xyz

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def walk():
    return synth.random_walk_box(257, (6.0, 7.0, 8.0), 31, 0.3, seed=21)


def moments(ctx, d, T):
    tg, t0, t1 = _abi.ts_pairs(5, 6, 0, 1, 0)
    P = _abi.Pairs(ctx, tg, t0, t1, _abi.ts_n_timegaps(5, 6, 0))
    return _abi.msd_ngp(ctx, d, None, P)


@pytest.mark.parametrize("chunk", [None, "4"])
def test_custom_matches_upload(ctx, walk, tmp_path, qvec, chunk):
    print("case: amdat_traj_read custom, chunk", chunk, flush=True)
    synth.write_lammps_custom(walk, str(tmp_path / "a.dump"))
    os.environ.pop("AMDAT_READ_CHUNK_FRAMES", None)
    if chunk:
        os.environ["AMDAT_READ_CHUNK_FRAMES"] = chunk
    try:
        d, types = _abi.Trajectory.read(ctx, "custom", tmp_path / "a.dump", walk.N, walk.T)
    finally:
        os.environ.pop("AMDAT_READ_CHUNK_FRAMES", None)
    u = _abi.Trajectory(ctx, *walk.arrays)
    assert np.array_equal(d.box(), walk.box)
    a, b = moments(ctx, d, walk.T), moments(ctx, u, walk.T)
    for p, q in zip(a, b):
        assert np.array_equal(p, q)
    off, n = ref_shells(qvec, 2, 20)
    sh = _abi.Shells(ctx, off, n)
    r1 = _abi.Rho(ctx, d, None, sh, 0, 18, np.float32(6.0)).download()
    r2 = _abi.Rho(ctx, u, None, sh, 0, 18, np.float32(6.0)).download()
    assert np.array_equal(r1, r2)


@pytest.mark.parametrize("chunk", [None, "3"])
def test_xyz_matches_true_images(ctx, walk, tmp_path, chunk):
    print("case: amdat_traj_read xyz, chunk", chunk, flush=True)
    write_xyz(walk, str(tmp_path / "w.xyz"))
    os.environ.pop("AMDAT_READ_CHUNK_FRAMES", None)
    if chunk:
        os.environ["AMDAT_READ_CHUNK_FRAMES"] = chunk
    try:
        d, _ = _abi.Trajectory.read(ctx, "xyz", tmp_path / "w.xyz", walk.N, walk.T, box3=walk.box[0])
    finally:
        os.environ.pop("AMDAT_READ_CHUNK_FRAMES", None)
    # images relative to frame 0 are the true ones (small steps): same displacements
    rel = synth.Traj(walk.x, walk.y, walk.z, walk.ix - walk.ix[0:1], walk.iy - walk.iy[0:1],
                     walk.iz - walk.iz[0:1], walk.box)
    u = _abi.Trajectory(ctx, *rel.arrays)
    a, b = moments(ctx, d, walk.T), moments(ctx, u, walk.T)
    for p, q in zip(a, b):
        assert np.array_equal(p, q)


def test_dropin_xyz_system(ctx, walk, tmp_path, qvec):
    write_xyz(walk, str(tmp_path / "w.xyz"))
    script = """system
xyz
w.xyz
exponential 5 6 1.3 0 0 1
A 257
1
1
6.0 7.0 8.0
0 6.0 0 7.0 0 8.0

create_list all
all
msd msd.dat
list all
"""
    (tmp_path / "in.amdat").write_text(script)
    print("case: amdat_hip xyz system block", flush=True)
    p = subprocess.run([os.path.join(ROOT, "amdat_amd", "bin", "amdat_hip"), "-i", "in.amdat"], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    tg, t0, t1 = orc.ts_pairs(5, 6)
    m = orc.msd_ngp_f64(walk, (tg, t0, t1), orc.ts_n_timegaps(5, 6))
    rows = [l.split("\t") for l in (tmp_path / "msd.dat").read_text().split("\n")[1:] if l.strip()]
    got = np.array([float(r[1]) for r in rows])
    assert got[0] == 0.0
    assert np.allclose(got[1:], m["msd"][1:], rtol=2e-5)
