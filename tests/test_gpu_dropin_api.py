"""The drop-in class API beyond the input-script driver (tests/dropin/
api_check.cpp): Wave_Density over every shell, its static S(q) form (block
starts only), set(), copies, and accumulation over two analyze() calls;
Intermediate_Scattering_Function for every shell (fullblock 1), one shell,
a shell range of a density that starts at shell 2 (physics-correct), the
reference's slot reading of a range past the density's first shell (slot
i over the vector count of shell wn1 + i), a slot read over fewer vectors,
a read past a slot refused, and one time gap (gap 0 over the block-start
density is S(k) over those frames; gap nes with fullblock 1), against the
FP64 oracle; densities of different shell ranges are refused; every class
driven through an Analysis* reports the reference's what_are_you kind and
writes its file.  N = 400, L = 7.0, exponential 3 5 (T = 16), all 99 shells
of the reference's q-vector files."""
import os
import subprocess

import numpy as np
import pytest

from amdat_amd import synth
from oracle import oracle as orc
from tests.conftest import ROOT, ref_shells

pytestmark = pytest.mark.gpu

NEXP, NES = 3, 5


@pytest.fixture(scope="module")
def out(tmp_path_factory, qvec):
    d = tmp_path_factory.mktemp("api")
    off, n = qvec
    qdir = d / "qvectors3d"
    qdir.mkdir()
    for b in range(2, 101):
        with open(qdir / ("qvector.%03d" % b), "w", newline="") as fh:
            for v in n[off[b - 1]:off[b]]:
                fh.write("  %.7e   %.7e   %.7e\r\n" % tuple(v))
    tr = synth.plain_config(N=400, L=7.0, nexp=NEXP, nes=NES, seed=21)
    raw = d / "traj.raw"
    with open(raw, "wb") as fh:
        for a in (tr.x, tr.y, tr.z):
            fh.write(np.ascontiguousarray(a, np.float32).tobytes())
        for a in (tr.ix, tr.iy, tr.iz):
            fh.write(np.ascontiguousarray(a, np.int16).tobytes())
        fh.write(np.ascontiguousarray(tr.box, np.float32).tobytes())
    print("case: api_check N=400 T=16 exponential 3 5", flush=True)
    p = subprocess.run([os.path.join(ROOT, "amdat_amd", "bin", "api_check"), str(raw), str(tr.N), str(tr.T),
                        str(NEXP), str(NES), "1.3"], env=dict(os.environ, AMDAT_WV3D=str(qdir / "qvector")),
                       capture_output=True, text=True, timeout=300, cwd=str(d))
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    res = {}
    for line in p.stdout.split("\n"):
        t = line.split()
        if t:
            res.setdefault(t[0], []).append(t[1:])
    res["_dir"] = d
    res["_stdout"] = p.stdout
    return tr, res


def table(rows, ntg, nsh):
    F = np.zeros((ntg, nsh))
    for t, s, v in rows:
        F[int(t), int(s)] = float(v)
    return F


def test_densities(out, qvec):
    tr, res = out
    off, n = ref_shells(qvec, 2, 100)
    L = float(np.float32(7.0))
    rho = orc.rho_f64_traj(tr, n, L)
    looped = res["LOOPED"][0]
    assert int(looped[0]) == tr.N * tr.T and int(looped[1]) == tr.N * NEXP and int(looped[2]) == 2
    for t, s, v, ar, ai, br, bi in res["RHO"]:
        t, s, v = int(t), int(s), int(v)
        want = rho[t, off[s] + v]
        got = complex(float(ar), float(ai))
        assert abs(got - want) <= 1e-6 * max(abs(want), np.sqrt(tr.N))
        blk = complex(float(br), float(bi))
        if t % NES == 0 and t // NES < NEXP:
            assert abs(blk - want) <= 1e-6 * max(abs(want), np.sqrt(tr.N))
        else:
            assert blk == 0
    assert res["MISMATCH"][0] == ["rejected"]


def test_isf_forms(out, qvec):
    tr, res = out
    off, n = ref_shells(qvec, 2, 100)
    L = float(np.float32(7.0))
    rho = orc.rho_f64_traj(tr, n, L)
    nat = np.full(tr.T, tr.N)
    ntg = orc.ts_n_timegaps(NEXP, NES, 0)
    S = orc.sk_f64(rho, rho, nat, off)
    scale = np.maximum(S, 1.0)

    def close(got, want, cols):
        assert np.all(np.abs(got - want) <= 1e-6 * scale[cols][None, :] + 1e-7 * np.abs(want))

    pf = orc.ts_pairs(NEXP, NES, fullblock=1)
    F1, _ = orc.fkt_f64(rho, rho, nat, pf, ntg, off)
    close(table(res["ALL"], ntg, 99), F1, np.arange(99))
    p0 = orc.ts_pairs(NEXP, NES)
    F0, _ = orc.fkt_f64(rho, rho, nat, p0, ntg, off)
    close(table(res["ONE"], ntg, 1), F0[:, 3:4], np.array([3]))
    close(table(res["RANGE"], ntg, 3), F0[:, 3:6], np.arange(3, 6))
    # gap 0 over the block-start density: S(k) over frames 0, 5, 10
    fr = [NES * b for b in range(NEXP)]
    Sb = orc.sk_f64(rho[fr], rho[fr], nat[fr], off)
    G0 = table(res["GAP0"], ntg, 99)
    close(G0[0:1], Sb[None, :], np.arange(99))
    assert np.all(G0[1:] == 0)
    # gap NES with fullblock 1: the fullblock pairs of that gap
    sel = pf[0] == NES
    FN, _ = orc.fkt_f64(rho, rho, nat, tuple(a[sel] for a in pf), ntg, off)
    GN = table(res["GAPN"], ntg, 99)
    close(GN[NES:NES + 1], FN[NES:NES + 1], np.arange(99))
    assert np.all(np.delete(GN, NES, axis=0) == 0)


def test_isf_slot_forms(out, qvec):
    """Intermediate_Scattering_Function(sys, wd1, wd2, wn1, wn2) reads the
    density's slot wn - wn1 with shell wn's vector count
    (src/intermediate_scattering_function.cpp:556-562): over a density of
    shells 20..30, shells 22..28 come out as shells 20..26 (300 vectors
    each); physics_correct reads shells 22..28; over a density of shells
    2..5, shell 3 is the first 12 of shell 2's 15 vectors, and shell 4 (33
    vectors from slot 0's 15) is refused."""
    tr, res = out
    off, n = ref_shells(qvec, 2, 100)
    L = float(np.float32(7.0))
    rho = orc.rho_f64_traj(tr, n, L)
    nat = np.full(tr.T, tr.N)
    ntg = orc.ts_n_timegaps(NEXP, NES, 0)
    S = orc.sk_f64(rho, rho, nat, off)
    p0 = orc.ts_pairs(NEXP, NES)
    F0, _ = orc.fkt_f64(rho, rho, nat, p0, ntg, off)

    def close(got, want, scale):
        assert np.all(np.abs(got - want) <= 1e-6 * np.maximum(scale, 1.0)[None, :] + 1e-7 * np.abs(want))

    close(table(res["SLOTS"], ntg, 7), F0[:, 20:27], S[20:27])
    close(table(res["ONESLOT"], ntg, 1), F0[:, 20:21], S[20:21])
    close(table(res["PHYS"], ntg, 7), F0[:, 22:29], S[22:29])
    assert not np.allclose(F0[:, 20:27], F0[:, 22:29])
    o2 = np.array([0, 12], np.int32)                   # shell 2, its first 12 vectors
    sub = rho[:, off[2]:off[2] + 12]
    F12, _ = orc.fkt_f64(sub, sub, nat, p0, ntg, o2)
    S12 = orc.sk_f64(sub, sub, nat, o2)
    close(table(res["SHORT"], ntg, 1), F12, S12)
    assert res["PASTSLOT"][0] == ["rejected"]


def test_density_accumulates(out, qvec):
    """Two analyze() calls over disjoint lists add densities and atom
    counts (src/wave_density.cpp:313-336); a copy taken between them keeps
    the first call's density."""
    tr, res = out
    off, n = ref_shells(qvec, 2, 11)
    L = float(np.float32(7.0))
    rho = orc.rho_f64_traj(tr, n, L)
    even = np.arange(0, tr.N, 3)
    rho_e = orc.rho_f64_traj(tr, n, L, atoms=even)
    looped = [int(v) for v in res["ACCLOOPED"][0]]
    assert looped == [tr.N * tr.T, len(even) * tr.T, tr.N]
    for t, s, v, ar, ai, br, bi in res["ACC"]:
        t, s, v = int(t), int(s), int(v)
        k = off[s] + v
        a, b = complex(float(ar), float(ai)), complex(float(br), float(bi))
        assert abs(a - rho[t, k]) <= 1e-6 * max(abs(rho[t, k]), np.sqrt(tr.N))
        assert abs(b - rho_e[t, k]) <= 1e-6 * max(abs(rho_e[t, k]), np.sqrt(len(even)))


def test_analysis_base_class(out, qvec):
    """Every class held as Analysis*: what_are_you() as the reference's
    classes report it (src/analysis.h:27-56; Wave_Density and
    Structure_Factor keep the base kind), show_system(), analyze() and
    write() through the base pointer; the base's messages for methods a
    class does not implement; asymmetric S(k) through analyze(l1, l2)."""
    tr, res = out
    kinds = {k: (int(v), int(s)) for k, v, s in res["KIND"]}
    want = {"msd": 12, "ngp": 15, "isfs": 9, "sk": 0, "msd2": 13, "md": 11, "vhs": 18, "wd": 0, "isf": 10}
    assert {k: v[0] for k, v in kinds.items()} == want
    assert all(v[1] == 1 for v in kinds.values())
    d = res["_dir"]
    for name in ("msd", "ngp", "isfs", "sk", "msd2", "md", "vhs", "isf"):
        assert (d / ("api_%s.dat" % name)).stat().st_size > 0
    assert not (d / "never.dat").exists()
    assert "No standard write method implemented" in res["_stdout"]
    assert "two lists not implemented" in res["_stdout"]
    off, n = ref_shells(qvec, 2, 100)
    L = float(np.float32(7.0))
    even = np.arange(0, tr.N, 3)
    re_ = orc.rho_f64_traj(tr, n, L, atoms=even)
    ra = orc.rho_f64_traj(tr, n, L)
    Sab = orc.sk_f64(re_, ra, np.full(tr.T, len(even)), off)
    got = np.array([float(v) for v in res["SKASYM"][0]])
    assert np.allclose(got, Sab[:5], rtol=2e-6, atol=1e-6)
