"""Rectangular bins (create_bin_list all/region, bin_list targets) on the
device: cell assignment bit-exact with the oracle, box bounds read from a
custom file, and the drop-in's per-bin msd, isfs and structure_factor files
(Bin_Static_Analysis names, quirk included) against the oracle composed
bin by bin.  The persistence form (Bin_Dynamics_Analysis: per frame pair,
the atoms in the bin at both frames) through pair-keyed lists, at the C ABI
and in the drop-in's files, against the oracle composed pair by pair.
Distance bins (point, plane, nearest atom of a second list): device cells
bit-exact with the oracle, errors, and the drop-in's create_bin_list
distance forms.  Plain form N = 1,000 in a box [-4.5, 4.91)^3,
exponential 5 20."""
import os
import subprocess

import numpy as np
import pytest

from amdat_amd import _abi, synth
from oracle import oracle as orc
from tests.conftest import ROOT, atoms_at, compose_pairs, ref_shells

pytestmark = pytest.mark.gpu

RTOL = 2e-5
NEXP, NES = 5, 20
NTG = NES + NEXP
LO = -4.5


def shifted():
    tr = synth.plain_config()
    lo = np.float32(LO)
    return synth.Traj((tr.x + lo).astype(np.float32), (tr.y + lo).astype(np.float32),
                      (tr.z + lo).astype(np.float32), tr.ix, tr.iy, tr.iz, tr.box)


def bounds(tr):
    lo = np.full((tr.T, 3), np.float32(LO), np.float32)
    return lo, (lo + tr.box).astype(np.float32)


def test_cells_match_oracle(ctx):
    print("case: bins_assign plain form", flush=True)
    tr = shifted()
    d = _abi.Trajectory(ctx, *tr.arrays)
    lo, hi = bounds(tr)
    for nb in [(2, 2, 2), (3, 1, 4), (7, 5, 3)]:
        got = _abi.bins_assign(ctx, d, nb, lo, hi)
        want = orc.bins_assign(tr, nb, lo, hi)
        assert np.array_equal(got, want)
        assert np.all(got >= 0)                                   # every atom is inside its box
    # a region: part of the atoms fall outside
    rlo = lo.copy()
    rhi = hi.copy()
    rlo[:, 0], rhi[:, 1] = -1.0, 2.0
    got = _abi.bins_assign(ctx, d, (4, 4, 4), rlo, rhi)
    assert np.array_equal(got, orc.bins_assign(tr, (4, 4, 4), rlo, rhi))
    assert 0 < np.mean(got < 0) < 1


def test_distance_cells_match_oracle(ctx):
    print("case: bins_distance point/plane/list, plain form", flush=True)
    tr = synth.plain_config()
    d = _abi.Trajectory(ctx, *tr.arrays)
    sub = np.arange(0, tr.N, 3)
    per = {t: np.arange(t % 5, tr.N, 4) for t in range(tr.T)}
    conv = np.arange(tr.T, dtype=np.int32)
    toff = np.r_[0, np.cumsum([len(per[t]) for t in range(tr.T)])].astype(np.int32)
    timed = _abi.TimedList(ctx, d, conv, toff, np.concatenate([per[t] for t in range(tr.T)]))
    static = _abi.AtomList(ctx, d, sub)
    lists = [(None, None), (static, sub), (timed, lambda t: per[t])]
    for dl, ol in lists:
        for p, nb in [((4.0, 4.0, 4.0), 100), ((0.5, 9.0, 2.25), 5)]:
            got = _abi.bins_distance(ctx, d, dl, _abi.DIST_POINT, 0.37, nb, point=p)
            assert np.array_equal(got, orc.bins_distance(tr, 0, 0.37, nb, point=p, binned=ol))
        for axis in range(3):
            for k in range(3):
                got = _abi.bins_distance(ctx, d, dl, _abi.DIST_PLANE, 0.29, 40, point=(3.3,), axis=axis, direction=k)
                want = orc.bins_distance(tr, 1, 0.29, 40, point=(3.3,), axis=axis, direction=k, binned=ol)
                assert np.array_equal(got, want), (axis, k)
    cl = np.arange(800, 1000)
    cluster = _abi.AtomList(ctx, d, cl)
    for dl, ol in lists:
        got = _abi.bins_distance(ctx, d, dl, _abi.DIST_LIST, 0.2, 8, cluster=cluster)
        want = orc.bins_distance(tr, 2, 0.2, 8, binned=ol, cluster=cl)
        assert np.array_equal(got, want)
        assert np.all(got[:, 800:][got[:, 800:] >= 0] == 0)
    got = _abi.bins_distance(ctx, d, None, _abi.DIST_LIST, 0.2, 8, cluster=timed)
    assert np.array_equal(got, orc.bins_distance(tr, 2, 0.2, 8, cluster=lambda t: per[t]))
    for kw in [dict(kind=_abi.DIST_POINT, point=(10.0, 1.0, 1.0)),            # beyond the box length
               dict(kind=_abi.DIST_PLANE, point=(9.5,), axis=1),
               dict(kind=3, point=(1.0, 1.0, 1.0))]:
        with pytest.raises(_abi.AmdatError):
            _abi.bins_distance(ctx, d, None, thickness=0.3, n_bins=4, **kw)


def bin_lists(tr, cells, nb):
    out = {}
    for b in range(int(np.prod(nb))):
        conv = np.arange(tr.T, dtype=np.int32)
        off, idx = [0], []
        for t in range(tr.T):
            a = np.nonzero(cells[t] == b)[0]
            idx.extend(a)
            off.append(len(idx))
        out[b] = (conv, np.array(off, np.int32), np.array(idx, np.int32))
    return out


def pair_bin_lists(cells, nb, pairs):
    """Per bin: over the distinct frame pairs (t0, t1), the atoms in the bin
    at t0 and at t1 (atom order), as (t0, t1, off, indices)."""
    _, t0, t1 = pairs
    keys = sorted(set(zip(np.asarray(t0).tolist(), np.asarray(t1).tolist())))
    k0 = np.array([k[0] for k in keys], np.int32)
    k1 = np.array([k[1] for k in keys], np.int32)
    out = {}
    for b in range(int(np.prod(nb))):
        off, idx = [0], []
        for a, c in keys:
            idx.extend(np.nonzero((cells[a] == b) & (cells[c] == b))[0])
            off.append(len(idx))
        out[b] = (k0, k1, np.array(off, np.int32), np.array(idx, np.int32))
    return out


def per_pair(plist, pairs, fn):
    """fn(one pair, its atoms) for every pair with atoms."""
    k0, k1, off, idx = plist
    key = {(a, c): i for i, (a, c) in enumerate(zip(k0.tolist(), k1.tolist()))}
    for g, a, c in zip(*(np.asarray(v).tolist() for v in pairs)):
        i = key[(a, c)]
        if off[i + 1] > off[i]:
            yield fn((np.array([g]), np.array([a]), np.array([c])), idx[off[i]:off[i + 1]])


def _list_atoms(kind, t, sub, per):
    """Atoms of the test list at frame t, ascending (None: every atom)."""
    if kind == "all":
        return None
    if kind == "static":
        return np.sort(sub)
    return np.sort(per[t]) if t in per else np.zeros(0, np.int64)


def _host_bin_lists(cell, nbins, kind, sub, per):
    T, N = cell.shape
    out = []
    for b in range(nbins):
        off, idx = [0], []
        for t in range(T):
            a = _list_atoms(kind, t, sub, per)
            a = np.nonzero(cell[t] == b)[0] if a is None else a[cell[t][a] == b]
            idx.extend(a.tolist())
            off.append(len(idx))
        out.append((np.array(off, np.int32), np.array(idx, np.int32)))
    return out


def test_device_bin_lists_match_host(ctx):
    """amdat_cells_bin_lists / amdat_cells_pair_lists (lists formed on the
    device) against the same lists formed from the downloaded cell table:
    rectangular and distance cells; every atom, a static subset in shuffled
    list order, a time-varying list (frames without atoms included); a bin
    range in the middle of the table."""
    print("case: device-formed bin lists vs host formation, plain form", flush=True)
    tr = shifted()
    d = _abi.Trajectory(ctx, *tr.arrays)
    lo, hi = bounds(tr)
    rng = np.random.default_rng(7)
    sub = rng.permutation(tr.N)[:371].astype(np.int32)      # list order is not index order
    per = {t: rng.permutation(tr.N)[:rng.integers(0, 300)] for t in range(tr.T) if t % 7 != 3}
    present = [t for t in range(tr.T) if t in per]
    lt = {t: i for i, t in enumerate(present)}
    conv = np.array([lt.get(t, -1) for t in range(tr.T)], np.int32)
    toff = np.r_[0, np.cumsum([len(per[t]) for t in present])].astype(np.int32)
    timed = _abi.TimedList(ctx, d, conv, toff, np.concatenate([per[t] for t in present]).astype(np.int32))
    static = _abi.AtomList(ctx, d, sub)
    tg, t0, t1 = _abi.ts_pairs(NEXP, NES, 0, 0, 0)
    keys = sorted(set(zip(t0.tolist(), t1.tolist())))
    k0 = np.array([k[0] for k in keys], np.int32)
    k1 = np.array([k[1] for k in keys], np.int32)
    tables = [("rect", _abi.Cells.assign(ctx, d, (3, 2, 2), lo, hi)),
              ("dist", _abi.Cells.distance(ctx, d, None, _abi.DIST_POINT, 0.45, 9, point=(4.0, 4.0, 4.0)))]
    for name, cells in tables:
        cell = cells.download()
        if name == "rect":
            assert np.array_equal(cell, _abi.bins_assign(ctx, d, (3, 2, 2), lo, hi))
        nb = cells.n_cells
        for kind, dl in (("all", None), ("static", static), ("timed", timed)):
            want = _host_bin_lists(cell, nb, kind, sub, per)
            got = cells.bin_lists(dl)
            assert len(got) == nb
            for b in range(nb):
                conv_b, off_b, idx_b = got[b].download()
                assert np.array_equal(conv_b, np.arange(tr.T))
                assert np.array_equal(off_b, want[b][0]), (name, kind, b)
                assert np.array_equal(idx_b, want[b][1]), (name, kind, b)
            part = cells.bin_lists(dl, first=1, n=nb - 2)
            for b in range(nb - 2):
                assert np.array_equal(part[b].download()[2], want[b + 1][1])
            # persistence form: the atoms of the list at t0 in the bin at t0 and t1
            pl = cells.pair_lists(k0, k1, dl)
            for b in range(nb):
                off_b, idx_b = pl[b].download()
                woff, widx = [0], []
                for a, c in keys:
                    at = _list_atoms(kind, a, sub, per)
                    m = (cell[a] == b) & (cell[c] == b)
                    sel = np.nonzero(m)[0] if at is None else at[m[at]]
                    widx.extend(sel.tolist())
                    woff.append(len(widx))
                assert np.array_equal(off_b, woff), (name, kind, b)
                assert np.array_equal(idx_b, widx), (name, kind, b)
        print("%s cells: %d bins, lists match" % (name, nb))
    cells = tables[0][1]
    with pytest.raises(_abi.AmdatError):
        cells.bin_lists(None, first=10, n=5)                       # past the table
    with pytest.raises(_abi.AmdatError):
        cells.pair_lists([0, 0], [1, 1])                           # a pair given twice


def test_pair_keyed_lists(ctx, qvec):
    print("case: pair-keyed bin lists at the C ABI, 2x2x2 bins", flush=True)
    tr = shifted()
    d = _abi.Trajectory(ctx, *tr.arrays)
    lo, hi = bounds(tr)
    cells = orc.bins_assign(tr, (2, 2, 2), lo, hi)
    pairs = orc.ts_pairs(NEXP, NES)
    P = _abi.Pairs(ctx, *pairs, NTG)
    o22, n22 = ref_shells(qvec, 22, 22)
    sh = _abi.Shells(ctx, o22, n22)
    ell = np.float32(9.41)
    plists = pair_bin_lists(cells, (2, 2, 2), pairs)
    static = bin_lists(tr, cells, (2, 2, 2))
    for b in (0, 5):
        pl = _abi.PairList(ctx, d, *plists[b])
        assert pl.n_keys == len(plists[b][0]) and pl.n == len(plists[b][3])
        s2, s4, cnt = _abi.msd_ngp(ctx, d, pl, P)
        R2, R4, RC = np.zeros(NTG), np.zeros(NTG), np.zeros(NTG)
        for m in per_pair(plists[b], pairs, lambda p, at: orc.msd_ngp_f64(tr, p, NTG, atoms=at)):
            R2 += m["s2"]
            R4 += m["s4"]
            RC += m["count"]
        assert np.array_equal(cnt, RC)
        assert np.allclose(s2, R2, rtol=1e-12, atol=1e-12) and np.allclose(s4, R4, rtol=1e-12, atol=1e-12)
        # atoms that stay in the bin are a subset of those in it at t0; at lag 0 they are the same
        sc = np.zeros(NTG)
        for m in compose_pairs(static[b], pairs, lambda p, at: orc.msd_ngp_f64(tr, p, NTG, atoms=at)):
            sc += m["count"]
        assert np.all(cnt <= sc) and cnt[0] == sc[0] and np.any(cnt < sc)
        F, fc = _abi.fs(ctx, d, pl, sh, 0, 0, ell, P)
        SF, SC = np.zeros((NTG, 1)), np.zeros(NTG)
        for f, c in per_pair(plists[b], pairs, lambda p, at: orc.fs_f64(tr, p, NTG, o22, n22, float(ell), at)):
            SF += f * c[:, None]
            SC += c
        assert np.array_equal(fc, SC)
        ok = SC > 0
        assert np.max(np.abs(F[ok] - SF[ok] / SC[ok, None])) < 1e-6
        sd, sd2, dc, hist = _abi.displacements(ctx, d, pl, P, n_bins=40, max_range=1.0)
        H, D2 = np.zeros((NTG, 40), np.uint64), np.zeros((NTG, 3))
        for h, _ in per_pair(plists[b], pairs, lambda p, at: orc.vhs_counts(tr, p, NTG, 40, 1.0, at)):
            H += h
        for _, b2, _ in per_pair(plists[b], pairs, lambda p, at: orc.disp_f64(tr, p, NTG, at)):
            D2 += b2
        assert np.array_equal(hist, H) and np.array_equal(dc, RC)
        assert np.allclose(sd2, D2, rtol=1e-12, atol=1e-12)
        # frame entry points reject a pair-keyed list
        with pytest.raises(_abi.AmdatError):
            _abi.Rho(ctx, d, pl, sh, 0, 0, ell)
        pl.close()
    k0, k1, off, idx = plists[0]
    with pytest.raises(_abi.AmdatError):                       # a frame pair given twice
        _abi.PairList(ctx, d, np.r_[k0, k0[:1]], np.r_[k1, k1[:1]], np.r_[off, off[-1]], idx)
    with pytest.raises(_abi.AmdatError):                       # frame out of range
        _abi.PairList(ctx, d, np.r_[k0[:-1], tr.T], k1, off, idx)


@pytest.fixture(scope="module")
def run(tmp_path_factory, qvec):
    d = tmp_path_factory.mktemp("bins")
    off, n = qvec
    qdir = d / "qvectors3d"
    qdir.mkdir()
    for b in range(2, 101):
        with open(qdir / ("qvector.%03d" % b), "w", newline="") as fh:
            for v in n[off[b - 1]:off[b]]:
                fh.write("  %.7e   %.7e   %.7e\r\n" % tuple(v))
    tr = shifted()
    synth.write_lammps_custom(tr, str(d / "traj.dump"), box_lo=LO)
    script = """system
custom
traj.dump
exponential %d %d 1.3 0 0 1
A 800 B 200
1 2
1 0
0 1

create_list all
all
create_bin_list grid
all 2 2 2
msd msd.dat
bin_list grid all
isfs isfs.dat 20 20 xyz 9.41 0
bin_list grid all
structure_factor sk.dat symmetric xyz 0 -1
bin_list grid all
msd msdp.dat
bin_list grid all 1
isfs isfsp.dat 20 20 xyz 9.41 1
bin_list grid all 1
structure_factor skp.dat symmetric xyz 0 -1
bin_list grid all 1
create_list bsp
species B
create_bin_list dpt
distance point all 0.5 0.5 0.5 1.0 4
create_bin_list dpl
distance plane above all z 0.2 1.5 3
create_bin_list dtr
distance trajectory all bsp 0.4 3
msd msd_dpt.dat
bin_list dpt all
msd msd_dpl.dat
bin_list dpl all 1
isfs isfs_dtr.dat 20 20 xyz 9.41 0
bin_list dtr all
""" % (NEXP, NES)
    (d / "in.amdat").write_text(script)
    env = dict(os.environ, AMDAT_WV3D=str(qdir / "qvector"))
    print("case: amdat_hip bin_list script, N=1000 T=101, 2x2x2 bins", flush=True)
    p = subprocess.run([os.path.join(ROOT, "amdat_amd", "bin", "amdat_hip"), "-i", "in.amdat"], cwd=str(d), env=env,
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]
    lo, hi = bounds(tr)
    cells = orc.bins_assign(tr, (2, 2, 2), lo, hi)
    return d, tr, bin_lists(tr, cells, (2, 2, 2)), cells


def name(stem, b, sep="_"):
    x, y, z = b // 4, (b // 2) % 2, b % 2
    return "%s_bin_%d%s%d%s%d.bindata" % (stem, x, sep, y, sep, z)


def test_bin_msd_files(run):
    d, tr, lists, _ = run
    pairs = orc.ts_pairs(NEXP, NES)
    for b, lst in lists.items():
        s2, cnt = np.zeros(NTG), np.zeros(NTG)
        for m in compose_pairs(lst, pairs, lambda p, at: orc.msd_ngp_f64(tr, p, NTG, atoms=at)):
            s2 += m["s2"]
            cnt += m["count"]
        rows = np.array([[float(v) for v in l.split("\t")] for l in
                         (d / name("msd.dat", b)).read_text().split("\n")[1:] if l.strip()])
        ok = cnt > 0
        assert np.allclose(rows[1:, 1][ok[1:]], (s2 / np.where(ok, cnt, 1))[1:][ok[1:]], rtol=RTOL)


def test_bin_isfs_and_sk_files(run, qvec):
    d, tr, lists, _ = run
    pairs = orc.ts_pairs(NEXP, NES)
    o22, n22 = ref_shells(qvec, 22, 22)
    offA, nA = ref_shells(qvec, 2, 100)
    ell = float(np.float32(9.41))
    for b, lst in lists.items():
        SF, SC = np.zeros((NTG, 1)), np.zeros(NTG)
        for f, c in compose_pairs(lst, pairs, lambda p, at: orc.fs_f64(tr, p, NTG, o22, n22, ell, at)):
            SF += f * c[:, None]
            SC += c
        lines = (d / name("isfs.dat", b)).read_text().split("\n")
        rows = np.array([[float(v) for v in l.strip("\t").split("\t")] for l in lines[2:] if l.strip()])
        ok = SC > 0
        assert np.allclose(rows[ok, 1], SF[ok, 0] / SC[ok], rtol=RTOL, atol=1e-6)
    # S(k) for two bins (every frame, the bin's atoms at that frame)
    L = float(np.float32(9.41))
    for b in (0, 7):
        lst = lists[b]
        natoms = np.array([len(atoms_at(lst, t)) for t in range(tr.T)])
        rho = np.stack([orc.rho_f64(tr.x[t], tr.y[t], tr.z[t], nA, L, atoms_at(lst, t)) if natoms[t] else
                        np.zeros(len(nA), complex) for t in range(tr.T)])
        S = orc.sk_f64(rho, rho, natoms, offA)
        lines = (d / name("sk.dat", b)).read_text().split("\n")
        rows = np.array([[float(v) for v in l.split("\t")] for l in lines[2:] if l.strip()])
        assert np.allclose(rows[:, 6], S, rtol=RTOL)


def test_bounds_from_custom_file(ctx, run):
    d, tr, lists, _ = run
    t, _ = _abi.Trajectory.read(ctx, "custom", str(d / "traj.dump"), tr.N, tr.T)
    lo = _abi.traj_bounds(t)
    assert np.all(lo == np.float32(LO))


def test_bin_persistence_files(run, qvec):
    """bin_list <bins> <list> 1: msd and isfs per bin over the atoms that
    stay in the bin (fullblock 0 even where isfs asks for 1); Bin_Dynamics_
    Analysis names; structure_factor is not dynamic and ignores the flag."""
    d, tr, lists, cells = run
    pairs = orc.ts_pairs(NEXP, NES)
    plists = pair_bin_lists(cells, (2, 2, 2), pairs)
    o22, n22 = ref_shells(qvec, 22, 22)
    ell = float(np.float32(9.41))
    for b, pl in plists.items():
        s2, cnt = np.zeros(NTG), np.zeros(NTG)
        SF, SC = np.zeros((NTG, 1)), np.zeros(NTG)
        for m in per_pair(pl, pairs, lambda p, at: orc.msd_ngp_f64(tr, p, NTG, atoms=at)):
            s2 += m["s2"]
            cnt += m["count"]
        for f, c in per_pair(pl, pairs, lambda p, at: orc.fs_f64(tr, p, NTG, o22, n22, ell, at)):
            SF += f * c[:, None]
            SC += c
        rows = np.array([[float(v) for v in l.split("\t")] for l in
                         (d / name("msdp.dat", b, ".")).read_text().split("\n")[1:] if l.strip()])
        ok = cnt > 0
        assert ok[0] and np.any(ok[1:])
        assert np.allclose(rows[1:, 1][ok[1:]], (s2 / np.where(ok, cnt, 1))[1:][ok[1:]], rtol=RTOL)
        lines = (d / name("isfsp.dat", b, ".")).read_text().split("\n")
        rows = np.array([[float(v) for v in l.strip("\t").split("\t")] for l in lines[2:] if l.strip()])
        ok = SC > 0
        assert np.allclose(rows[ok, 1], SF[ok, 0] / SC[ok], rtol=RTOL, atol=1e-6)
        assert (d / name("skp.dat", b)).read_text() == (d / name("sk.dat", b)).read_text()


def test_distance_bin_files(run, qvec):
    """create_bin_list distance point / plane above / trajectory: the bin
    count capped as the reference caps it (4, 3, 3 here, all below the cap),
    cells as the oracle forms them; msd per bin (static and persistence) and
    isfs per bin against the oracle composed bin by bin."""
    d, tr, _, _ = run
    pairs = orc.ts_pairs(NEXP, NES)
    bsp = np.arange(800, 1000)
    cases = [("msd_dpt.dat", orc.bins_distance(tr, 0, 1.0, 4, point=(0.5, 0.5, 0.5)), (1, 1, 4), False),
             ("msd_dpl.dat", orc.bins_distance(tr, 1, 1.5, 3, point=(0.2,), axis=2, direction=1), (1, 1, 3), True)]
    for stem, cells, nb, persist in cases:
        assert cells.max() == nb[2] - 1
        lists = pair_bin_lists(cells, nb, pairs) if persist else bin_lists(tr, cells, nb)
        for b, lst in lists.items():
            s2, cnt = np.zeros(NTG), np.zeros(NTG)
            it = per_pair(lst, pairs, lambda p, at: orc.msd_ngp_f64(tr, p, NTG, atoms=at)) if persist else \
                compose_pairs(lst, pairs, lambda p, at: orc.msd_ngp_f64(tr, p, NTG, atoms=at))
            for m in it:
                s2 += m["s2"]
                cnt += m["count"]
            rows = np.array([[float(v) for v in l.split("\t")] for l in
                             (d / ("%s_bin_0%s0%s%d.bindata" % (stem, *(".." if persist else "__"), b))).read_text()
                             .split("\n")[1:] if l.strip()])
            ok = cnt > 0
            assert ok[0]
            assert np.allclose(rows[1:, 1][ok[1:]], (s2 / np.where(ok, cnt, 1))[1:][ok[1:]], rtol=RTOL)
    cells = orc.bins_distance(tr, 2, 0.4, 3, cluster=bsp)
    assert np.all(cells[:, 800:] == 0)
    o22, n22 = ref_shells(qvec, 22, 22)
    ell = float(np.float32(9.41))
    for b, lst in bin_lists(tr, cells, (1, 1, 3)).items():
        SF, SC = np.zeros((NTG, 1)), np.zeros(NTG)
        for f, c in compose_pairs(lst, pairs, lambda p, at: orc.fs_f64(tr, p, NTG, o22, n22, ell, at)):
            SF += f * c[:, None]
            SC += c
        lines = (d / ("isfs_dtr.dat_bin_0_0_%d.bindata" % b)).read_text().split("\n")
        rows = np.array([[float(v) for v in l.strip("\t").split("\t")] for l in lines[2:] if l.strip()])
        ok = SC > 0
        assert ok[0] and np.allclose(rows[ok, 1], SF[ok, 0] / SC[ok], rtol=RTOL, atol=1e-6)
