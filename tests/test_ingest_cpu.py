"""Trajectory readers (host_io.cpp): LAMMPS custom dumps and xyz files into
frame-major SoA.  Pins: exact round trip of the writer's %.9g floats and
the image columns; xu yu zu turned into wrapped + images by
floor((u - lo)/L); xyz image counts equal to the true images of a slow
walk and to a float32 replay of the reference's unwrapping rule
(src/trajectory.cpp:233-270); errors for short files and bad counts; the
same bytes for any thread count."""
import numpy as np
import pytest

from amdat_amd import _abi, synth


def write_xyz(tr, path, types=None):
    types = np.ones(tr.N, np.int32) if types is None else types
    with open(path, "w") as fh:
        for t in range(tr.T):
            fh.write("%d\natoms\n" % tr.N)
            fh.write("".join("%d %.9g %.9g %.9g\n" % r for r in zip(types, tr.x[t], tr.y[t], tr.z[t])))


def write_custom_unwrapped(tr, path):
    u = tr.unwrapped64().astype(np.float32)
    with open(path, "w") as fh:
        for t in range(tr.T):
            fh.write("ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n" % (t, tr.N))
            for c in range(3):
                fh.write("0 %.9g\n" % float(tr.box[t, c]))
            fh.write("ITEM: ATOMS id type xu yu zu\n")
            fh.write("".join("%d 1 %.9g %.9g %.9g\n" % (j + 1, *u[t, j]) for j in range(tr.N)))


@pytest.fixture(scope="module")
def walk():
    return synth.random_walk_box(300, (6.0, 7.0, 8.0), 23, 0.4, seed=12)


def test_custom_round_trip(tmp_path, walk):
    types = (np.arange(walk.N) % 3 + 1).astype(np.int32)
    synth.write_lammps_custom(walk, str(tmp_path / "a.dump"), types=types)
    for threads in (1, 5):
        d = _abi.read_custom(tmp_path / "a.dump", walk.N, walk.T, threads=threads)
        for k in ("x", "y", "z", "ix", "iy", "iz"):
            assert np.array_equal(d[k], getattr(walk, k)), k
        assert np.array_equal(d["box"], walk.box)
        assert np.array_equal(d["types"], types)


def test_custom_unwrapped_columns(tmp_path, walk):
    write_custom_unwrapped(walk, str(tmp_path / "u.dump"))
    d = _abi.read_custom(tmp_path / "u.dump", walk.N, walk.T)
    u = walk.unwrapped64().astype(np.float32).astype(np.float64)
    L = walk.box.astype(np.float64)[:, None, :]
    k = np.floor(u / L)
    assert np.array_equal(d["ix"], k[..., 0].astype(np.int16))
    assert np.array_equal(d["iz"], k[..., 2].astype(np.int16))
    assert np.array_equal(d["x"], (u[..., 0] - k[..., 0] * L[..., 0]).astype(np.float32))
    # the rebuilt unwrapped position is the file's value
    rebuilt = d["y"].astype(np.float64) + d["iy"] * L[..., 1]
    assert np.max(np.abs(rebuilt - u[..., 1])) < 1e-5


def reference_images(x, L):
    """float32 replay of Trajectory::unwrap_displacement: image(t) =
    image(t-1) - int((x(t) - x(t-1)) / (L * 0.55f))."""
    s = np.float32(L) * np.float32(0.55)
    d = (x[1:] - x[:-1]).astype(np.float32)
    jump = np.trunc((d / s).astype(np.float32)).astype(np.int64)
    return np.concatenate([np.zeros((1, x.shape[1]), np.int64), -np.cumsum(jump, axis=0)])


def test_xyz_images(tmp_path, walk):
    write_xyz(walk, str(tmp_path / "w.xyz"))
    box = walk.box[0]
    for threads in (1, 3):
        d = _abi.read_xyz(tmp_path / "w.xyz", walk.N, walk.T, box, threads=threads)
        for c, k in enumerate(("ix", "iy", "iz")):
            assert np.array_equal(d[k], reference_images(getattr(walk, "xyz"[c]), box[c]).astype(np.int16))
            # steps of sigma = 0.4 never look like a box crossing: the true images, relative to frame 0
            assert np.array_equal(d[k].astype(np.int64), getattr(walk, k) - getattr(walk, k)[0:1])
        assert np.array_equal(d["x"], walk.x)


def test_errors(tmp_path, walk):
    synth.write_lammps_custom(walk, str(tmp_path / "a.dump"))
    with pytest.raises(_abi.AmdatError) as e:
        _abi.read_custom(tmp_path / "a.dump", walk.N, walk.T + 1)      # too few frames
    assert e.value.code == -5
    with pytest.raises(_abi.AmdatError):
        _abi.read_custom(tmp_path / "a.dump", walk.N + 1, walk.T)      # wrong atom count
    text = (tmp_path / "a.dump").read_text()
    (tmp_path / "t.dump").write_text(text[: len(text) - 200])          # truncated last frame
    with pytest.raises(_abi.AmdatError):
        _abi.read_custom(tmp_path / "t.dump", walk.N, walk.T)
    with pytest.raises(_abi.AmdatError):
        _abi.read_custom(tmp_path / "missing.dump", walk.N, walk.T)
    write_xyz(walk, str(tmp_path / "w.xyz"))
    with pytest.raises(_abi.AmdatError):
        _abi.read_xyz(tmp_path / "w.xyz", walk.N + 2, walk.T, walk.box[0])
