"""Host layout of k-vector plans (plan_layout, host C++ behind the
amdat_internal_plan_layout hook): every vector lands exactly once with its
conjugation flag, a thread's R slots share nx, the per-chunk ranges cover
every slot the phase kernels read, and units past the active count are
padding only."""
import ctypes as C

import numpy as np
import pytest

from amdat_amd import _abi
from tests.conftest import ref_shells


def wave_codes(rng):
    """Group code of each chunk's 8 waves (2 bits per wave above bit 16 of
    the range word; group size 1 << code)."""
    w = rng[:, 3].astype(np.int64) & 0xFFFFFFFF
    return (w[:, None] >> (16 + 2 * np.arange(8))[None, :]) & 3


def layout(nvec, R, pair=0):
    L = _abi.lib()
    fn = L.amdat_internal_plan_layout
    fn.restype = C.c_int
    nv = np.ascontiguousarray(nvec, np.int32).reshape(-1)
    K = len(nv) // 3
    nc, cf = C.c_int(), C.c_long()
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
    assert fn(P(nv, C.c_int), K, R, pair, None, None, None, C.byref(nc), C.byref(cf)) == 0
    n, KB = nc.value, 512 * R
    pack = np.zeros(n * KB, np.uint32)
    dst = np.zeros(n * KB, np.int32)
    rng = np.zeros(n * 4, np.int32)
    assert fn(P(nv, C.c_int), K, R, pair, P(pack, C.c_uint32), P(dst, C.c_int), P(rng, C.c_int), C.byref(nc),
              C.byref(cf)) == 0
    nx = (pack & 0xFF).astype(np.int64) - 128
    ny = ((pack >> 8) & 0xFF).astype(np.int64) - 128
    nz = ((pack >> 16) & 0xFF).astype(np.int64)
    layout.swap = ((pack >> 24) & 1).astype(bool).reshape(n, R, 512)
    return n, nx.reshape(n, R, 512), ny.reshape(n, R, 512), nz.reshape(n, R, 512), dst.reshape(n, R, 512), \
        rng.reshape(n, 4), cf.value


@pytest.mark.parametrize("bins,R", [((2, 100), 8), ((2, 100), 4), ((2, 40), 2), ((40, 41), 8), ((2, 3), 2),
                                    ((95, 100), 4)])
def test_layout_invariants(qvec, bins, R):
    off, nvec = ref_shells(qvec, *bins)
    K = len(nvec)
    n, nx, ny, nz, dst, rng, conflicts = layout(nvec, R)
    assert n == max(1, -(-int(np.sum([-(-c // R) for c in np.unique(
        np.where(nvec[:, 2] < 0, -nvec[:, 0], nvec[:, 0]), return_counts=True)[1]])) // 512))
    real = dst >= 0
    idx = dst[real] & ~(1 << 30)
    assert np.array_equal(np.sort(idx), np.arange(K))          # each vector once
    flip = (dst[real] & (1 << 30)) != 0
    orig = nvec[idx]
    sign = np.where(flip, -1, 1)[:, None]
    got = np.stack([nx[real], ny[real], nz[real]], 1)
    assert np.array_equal(got, orig * sign)                      # canonical nz >= 0, flag on flips
    assert np.all(flip == (orig[:, 2] < 0))
    assert np.all(nz >= 0)
    assert np.all(nx == nx[:, :1, :])                            # a thread's slots share nx
    xlo, xhi = rng[:, 0][:, None, None], rng[:, 1][:, None, None]
    my, mz = (rng[:, 2] & 0xFFFF)[:, None, None], (rng[:, 2] >> 16)[:, None, None]
    assert np.all((nx >= xlo) & (nx <= xhi) & (np.abs(ny) <= my) & (nz <= mz))
    act = rng[:, 3] & 0xFFFF
    assert np.all(rng[:, 3] >> 16 == 0)                          # unpaired plan: no paired waves
    for c in range(n):
        unit_real = real[c].any(axis=0)                          # [512]
        assert np.all(unit_real[:act[c]]) and not np.any(unit_real[act[c]:])
    assert conflicts >= 0


def test_layout_lowers_bank_conflicts(qvec):
    """The slot permutation leaves fewer extra LDS passes than the plain
    sorted order of the same units would (estimate from the same metric)."""
    off, nvec = ref_shells(qvec, 2, 100)
    n, nx, ny, nz, dst, rng, conflicts = layout(nvec, 8)

    def extra(vals):
        worst = 0
        banks = {}
        for v in set(vals.tolist()):
            banks.setdefault(v % 32, set()).add(v)
            worst = max(worst, len(banks[v % 32]))
        return worst - 1

    # same units with slots re-sorted by (ny, nz): undoes the permutation
    total = 0
    for c in range(n):
        key = ny[c] * 1000 + nz[c]
        order = np.argsort(key, axis=0, kind="stable")
        sy = np.take_along_axis(ny[c], order, 0)
        sz = np.take_along_axis(nz[c], order, 0)
        for r in range(8):
            for h in range(0, 512, 32):
                total += extra(sy[r, h:h + 32]) + extra(sz[r, h:h + 32])
                if r == 0:
                    total += extra(nx[c, 0, h:h + 32])
    assert conflicts < total


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("bins,R", [((2, 100), 8), ((2, 100), 4), ((40, 41), 8), ((2, 3), 4), ((95, 100), 4)])
def test_paired_layout_invariants(qvec, bins, R, mode):
    """Paired plans: the same invariants; the units of the waves whose group
    code is 1 (pairs) hold pairs sharing ny, or nz with both slots
    flagged as swapped, and the other waves single vectors; chunk count from
    pairing on ny first, then on nz, and the vectors left padded into pairs
    (mode 1: every wave paired) or R per unit in unpaired waves of their
    own (mode 2: paired and unpaired waves of 64 units each, interleaved)."""
    off, nvec = ref_shells(qvec, *bins)
    K = len(nvec)
    n, nx, ny, nz, dst, rng, conflicts = layout(nvec, R, pair=mode)
    sw = layout.swap
    canon = np.where(nvec[:, 2:3] < 0, -nvec, nvec)
    up = us = 0
    for x in np.unique(canon[:, 0]):
        sel = canon[canon[:, 0] == x]
        sel = sel[np.lexsort((sel[:, 2], sel[:, 1]))]               # (ny, nz) order: the odd one is the last
        ys, cy = np.unique(sel[:, 1], return_counts=True)
        groups = int(np.sum(cy // 2))
        left = [sel[sel[:, 1] == y][-1] for y, c in zip(ys, cy) if c % 2]
        singles = 0
        if left:
            cz = np.unique(np.array(left)[:, 2], return_counts=True)[1]
            groups += int(np.sum(cz // 2))
            singles = int(np.sum(cz % 2))
        if mode == 1:
            up += -(-(groups + singles) // (R // 2))
        else:
            up += -(-groups // (R // 2))
            us += -(-singles // R)
    wave_p = (wave_codes(rng) == 1).astype(int)                          # [n, 8]: waves of pairs
    assert np.all(wave_codes(rng) <= 1)
    is_p = np.repeat(wave_p, 64, axis=1).astype(bool)                   # [n, 512] per unit
    act_w = -(-(rng[:, 3] & 0xFFFF) // 64)
    if mode == 1:
        assert n == max(1, -(-up // 512))
        for c in range(n):                                              # every active wave is paired
            assert np.array_equal(wave_p[c], (np.arange(8) < act_w[c]).astype(int))
    else:
        wp, ws = -(-up // 64), -(-us // 64)
        assert n == max(1, -(-(wp + ws) // 8))
        assert int(wave_p.sum()) == wp
        assert np.all(wave_p[np.arange(8)[None, :] >= act_w[:, None]] == 0)
    assert np.array_equal(sw[:, 0::2, :], sw[:, 1::2, :])       # both slots of a pair agree
    if bins == (2, 100):
        assert np.any(sw)                                         # nz pairs occur
    shared_y = ny[:, 0::2, :] == ny[:, 1::2, :]
    shared_z = nz[:, 0::2, :] == nz[:, 1::2, :]
    ok = np.where(sw[:, 0::2, :], shared_z, shared_y)
    assert np.all(ok | ~is_p[:, None, :])                          # paired units share a factor
    assert not np.any(sw & ~is_p[:, None, :])                     # single units are never swapped
    real = dst >= 0
    idx = dst[real] & ~(1 << 30)
    assert np.array_equal(np.sort(idx), np.arange(K))
    flip = (dst[real] & (1 << 30)) != 0
    got = np.stack([nx[real], ny[real], nz[real]], 1)
    assert np.array_equal(got, nvec[idx] * np.where(flip, -1, 1)[:, None])
    assert np.all(nz >= 0) and np.all(nx == nx[:, :1, :])
    xlo, xhi = rng[:, 0][:, None, None], rng[:, 1][:, None, None]
    my, mz = (rng[:, 2] & 0xFFFF)[:, None, None], (rng[:, 2] >> 16)[:, None, None]
    assert np.all((nx >= xlo) & (nx <= xhi) & (np.abs(ny) <= my) & (nz <= mz))
    act = rng[:, 3] & 0xFFFF
    for c in range(n):
        unit_real = real[c].any(axis=0)
        assert not np.any(unit_real[act[c]:])


@pytest.mark.parametrize("bins,R", [((2, 100), 8), ((2, 100), 4), ((40, 41), 8), ((2, 3), 4)])
def test_quad_layout_invariants(qvec, bins, R):
    """Quad plans (mode 3): every vector exactly once with its flip; the
    waves the mask marks hold units of R/4 quads whose four slots share ny
    (none swapped), the other waves pairs sharing ny or nz; the share of
    vectors in quads equals, per nx group, four times the quads its ny
    classes hold."""
    off, nvec = ref_shells(qvec, *bins)
    K = len(nvec)
    n, nx, ny, nz, dst, rng, conflicts = layout(nvec, R, pair=3)
    sw = layout.swap
    real = dst >= 0
    idx = dst[real] & ~(1 << 30)
    assert np.array_equal(np.sort(idx), np.arange(K))
    flip = (dst[real] & (1 << 30)) != 0
    assert np.array_equal(np.stack([nx[real], ny[real], nz[real]], 1), nvec[idx] * np.where(flip, -1, 1)[:, None])
    assert np.all(nx == nx[:, :1, :])
    codes = wave_codes(rng)
    assert np.all(codes[codes != 2] <= 1)                        # quads (code 2) and pairs
    isq = np.repeat(codes == 2, 64, axis=1)                      # [n, 512]
    for q in range(R // 4):
        g = slice(4 * q, 4 * q + 4)
        ok = np.all(ny[:, g, :] == ny[:, 4 * q:4 * q + 1, :], axis=1) & ~np.any(sw[:, g, :], axis=1)
        assert np.all(ok | ~isq)
    shared = np.where(sw[:, 0::2, :], nz[:, 0::2, :] == nz[:, 1::2, :], ny[:, 0::2, :] == ny[:, 1::2, :])
    assert np.all(shared | isq[:, None, :])
    canon = np.where(nvec[:, 2:3] < 0, -nvec, nvec)
    cls = np.unique(canon[:, :2], axis=0, return_counts=True)[1]
    assert int((real & isq[:, None, :]).sum()) == int(np.sum(cls // 4) * 4)


@pytest.mark.parametrize("mode", [4, 5])
@pytest.mark.parametrize("bins,R", [((2, 100), 8), ((2, 100), 4), ((40, 41), 8), ((2, 3), 4), ((95, 100), 8)])
def test_grouped_layout_invariants(qvec, bins, R, mode):
    """Grouped plans (modes 4 and 5): every vector exactly once with its
    flip; a thread's slots share nx; in a wave of group code g the slots of
    each group of 1 << g share the lead slot's roles and shared component
    (nz when swapped, else ny), which the kernel's product q is formed
    from; codes fit R; ranges cover every slot; units past the active count
    hold no vector.  Mode 5 with R = 8 marks its chunks (range bit 15) and
    keeps the slots its kernels skip empty: 4..7 in waves of singles, 6..7
    in waves of pairs."""
    off, nvec = ref_shells(qvec, *bins)
    K = len(nvec)
    n, nx, ny, nz, dst, rng, conflicts = layout(nvec, R, pair=mode)
    short = mode == 5 and R == 8
    assert np.all(((rng[:, 2] >> 15) & 1) == (1 if short else 0))
    sw = layout.swap
    real = dst >= 0
    idx = dst[real] & ~(1 << 30)
    assert np.array_equal(np.sort(idx), np.arange(K))
    flip = (dst[real] & (1 << 30)) != 0
    assert np.array_equal(np.stack([nx[real], ny[real], nz[real]], 1), nvec[idx] * np.where(flip, -1, 1)[:, None])
    assert np.all(nz >= 0) and np.all(nx == nx[:, :1, :])
    codes = wave_codes(rng)
    assert codes.max() <= {4: 2, 8: 3}[R]
    act = rng[:, 3] & 0xFFFF
    key = np.where(sw, nz, ny)
    for c in range(n):
        for w in range(8):
            if w * 64 >= act[c]:
                assert codes[c, w] == 0
                continue
            gs = 1 << codes[c, w]
            lanes = slice(w * 64, w * 64 + 64)
            if short and gs <= 2:
                ns = 4 if gs == 1 else 6
                assert not np.any(real[c, ns:, lanes])
            for g0 in range(0, R, gs):
                lead_sw, lead_key = sw[c, g0, lanes], key[c, g0, lanes]
                for r in range(g0, g0 + gs):
                    m = real[c, r, lanes]
                    assert np.all(sw[c, r, lanes][m] == lead_sw[m])
                    assert np.all(key[c, r, lanes][m] == lead_key[m])
    xlo, xhi = rng[:, 0][:, None, None], rng[:, 1][:, None, None]
    my, mz = (rng[:, 2] & 0x7FFF)[:, None, None], (rng[:, 2] >> 16)[:, None, None]
    assert np.all((nx >= xlo) & (nx <= xhi) & (np.abs(ny) <= my) & (nz <= mz))
    for c in range(n):
        unit_real = real[c].any(axis=0)
        assert not np.any(unit_real[act[c]:])


def test_grouped_layout_fewer_clashes_than_pairs(qvec):
    """On the reference shells 2..100, the grouped placement leaves fewer
    extra LDS passes than the paired layout."""
    off, nvec = ref_shells(qvec, 2, 100)
    c4 = layout(nvec, 8, pair=4)[-1]
    c1 = layout(nvec, 8, pair=1)[-1]
    assert c4 < c1
