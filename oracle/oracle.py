"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() as
the checker, and bench.py's cpu_baseline leg.  Never by the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_I, _F, _D, _P = C.c_int, C.c_float, C.c_double, C.c_void_p
_PI, _PF, _PD, _PS = C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int16)
_PL = C.POINTER(C.c_longlong)

SIGS = {
    "orc_set_threads": (None, [_I]),
    "orc_ts_n_timesteps": (_I, [_I, _I]),
    "orc_ts_n_timegaps": (_I, [_I, _I, _I]),
    "orc_ts_timelist": (None, [_I, _I, _F, _I, _F, _PF]),
    "orc_ts_displacement_times": (None, [_I, _I, _I, _PF, _PF]),
    "orc_ts_pairs": (_I, [_I, _I, _I, _I, _I, _PI, _PI, _PI, _I]),
    "orc_ts_weighting": (None, [_I, _I, _I, _I, _I, _PI]),
    "orc_shell_enumerate": (_I, [_I, _I, _PI, _I]),
    "orc_wave_spacing": (_F, [_F]),
    "orc_approx_wavenumbers": (None, [_F, _I, _PF]),
    "orc_rho_f64": (None, [_I, _PF, _PF, _PF, _PI, _I, _I, _PI, _D, _PD]),
    "orc_sk_f64": (None, [_I, _I, _PD, _PD, _PL, _I, _PI, _PD]),
    "orc_fkt_f64": (None, [_I, _I, _PD, _PD, _PL, _I, _PI, _PI, _PI, _I, _I, _PI, _PD, _PD]),
    "orc_fs_f64": (None, [_I, _PF, _PF, _PF, _PS, _PS, _PS, _PF, _PI, _I, _I, _PI, _PI, _PI, _I, _I, _PI, _PI,
                          _D, _PD, _PD]),
    "orc_msd_ngp_f64": (None, [_I, _PF, _PF, _PF, _PS, _PS, _PS, _PF, _PI, _I, _I, _PI, _PI, _PI, _I, _PD, _PD,
                               _PD, _PD, _PD]),
    "orc_ref32_rho": (None, [_I, _PF, _PF, _PF, _PI, _I, _I, _PI, _F, _PD]),
    "orc_ref32_sk": (None, [_I, _PF, _PF, _PF, _PI, _I, _I, _PI, _I, _PI, _PI, _F, _PF]),
    "orc_ref32_fs_pair": (_I, [_I, _PF, _PF, _PF, _PS, _PS, _PS, _PF, _PI, _I, _I, _I, _I, _PI, _PI, _F, _PF]),
    "orc_ref32_msd_pair": (_I, [_I, _PF, _PF, _PF, _PS, _PS, _PS, _PF, _PI, _I, _I, _I, _PF, _PF]),
    "orc_sk_loss_grad_f64": (_D, [_I, _PD, _PD, _PD, _I, _PI, _PI, _D, _PD, _PD, _PD, _PD, _PD, _PD]),
    "orc_disp_f64": (None, [_I, _PF, _PF, _PF, _PS, _PS, _PS, _PF, _PI, _I, _I, _PI, _PI, _PI, _I, _PD, _PD, _PD]),
    "orc_bins_distance": (None, [_I, _I, _PF, _PF, _PF, _PF, _I, _PF, _I, _I, _PI, _PI, _PI, _PI, _F, _I, _PI]),
    "orc_vhs_counts": (None, [_I, _PF, _PF, _PF, _PS, _PS, _PS, _PF, _PI, _I, _I, _PI, _PI, _PI, _I, _I, _F,
                              C.POINTER(C.c_ulonglong), _PD]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for k, (r, a) in SIGS.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dt)


def _p(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def set_threads(n):
    lib().orc_set_threads(n)


# ---- time scheme ----
def ts_timelist(nexp, nes, base, first_exp=0, dt=1.0):
    out = np.zeros(lib().orc_ts_n_timesteps(nexp, nes), np.float32)
    lib().orc_ts_timelist(nexp, nes, base, first_exp, dt, _p(out, C.c_float))
    return out


def ts_displacement_times(nexp, nes, frt, timelist):
    out = np.zeros(lib().orc_ts_n_timegaps(nexp, nes, frt), np.float32)
    tl = _c(timelist, np.float32)
    lib().orc_ts_displacement_times(nexp, nes, frt, _p(tl, C.c_float), _p(out, C.c_float))
    return out


def ts_pairs(nexp, nes, frt=0, fullblock=0, limit=0):
    n = lib().orc_ts_pairs(nexp, nes, frt, fullblock, limit, None, None, None, 0)
    tg, t0, t1 = (np.zeros(n, np.int32) for _ in range(3))
    lib().orc_ts_pairs(nexp, nes, frt, fullblock, limit, _p(tg, C.c_int), _p(t0, C.c_int), _p(t1, C.c_int), n)
    return tg, t0, t1


def ts_weighting(nexp, nes, frt=0, fullblock=0, limit=0):
    out = np.zeros(lib().orc_ts_n_timegaps(nexp, nes, frt), np.int32)
    lib().orc_ts_weighting(nexp, nes, frt, fullblock, limit, _p(out, C.c_int))
    return out


def ts_n_timegaps(nexp, nes, frt=0):
    return lib().orc_ts_n_timegaps(nexp, nes, frt)


# ---- shells ----
def shell_enumerate(b, maxc=100):
    n = lib().orc_shell_enumerate(b, maxc, None, 0)
    out = np.zeros(3 * n, np.int32)
    lib().orc_shell_enumerate(b, maxc, _p(out, C.c_int), n)
    return out.reshape(-1, 3)


def wave_spacing(length_scale):
    return lib().orc_wave_spacing(length_scale)


def approx_wavenumbers(length_scale, n=99):
    out = np.zeros(n, np.float32)
    lib().orc_approx_wavenumbers(length_scale, n, _p(out, C.c_float))
    return out


# ---- FP64 physics ----
def rho_f64(x, y, z, nvec, ell, atoms=None):
    """One frame: x, y, z [N] FP32; nvec [K,3]; returns complex [K]."""
    x, y, z = (_c(a, np.float32) for a in (x, y, z))
    nv = _c(nvec, np.int32).reshape(-1)
    K = len(nv) // 3
    at = None if atoms is None else _c(atoms, np.int32)
    out = np.zeros(2 * K)
    lib().orc_rho_f64(len(x), _p(x, C.c_float), _p(y, C.c_float), _p(z, C.c_float), _p(at, C.c_int),
                      0 if at is None else len(at), K, _p(nv, C.c_int), float(ell), _p(out, C.c_double))
    return out[0::2] + 1j * out[1::2]


def rho_f64_traj(tr, nvec, ell, frames=None, atoms=None):
    frames = range(tr.T) if frames is None else frames
    return np.stack([rho_f64(tr.x[f], tr.y[f], tr.z[f], nvec, ell, atoms) for f in frames])


def _cplx_to_pairs(r):
    a = np.empty(r.shape + (2,))
    a[..., 0] = r.real
    a[..., 1] = r.imag
    return np.ascontiguousarray(a)


def sk_f64(rho1, rho2, natoms, shell_off):
    """rho1/rho2 complex [F, K]; natoms [F]."""
    F, K = rho1.shape
    a, b = _cplx_to_pairs(rho1), _cplx_to_pairs(rho2)
    na = _c(natoms, np.int64)
    off = _c(shell_off, np.int32)
    S = np.zeros(len(off) - 1)
    lib().orc_sk_f64(F, K, _p(a, C.c_double), _p(b, C.c_double), _p(na, C.c_longlong), len(off) - 1,
                     _p(off, C.c_int), _p(S, C.c_double))
    return S


def fkt_f64(rho1, rho2, natoms, pairs, ntg, shell_off):
    T, K = rho1.shape
    a, b = _cplx_to_pairs(rho1), _cplx_to_pairs(rho2)
    na = _c(natoms, np.int64)
    tg, t0, t1 = (_c(v, np.int32) for v in pairs)
    off = _c(shell_off, np.int32)
    ns = len(off) - 1
    F = np.zeros(ntg * ns)
    cnt = np.zeros(ntg)
    lib().orc_fkt_f64(T, K, _p(a, C.c_double), _p(b, C.c_double), _p(na, C.c_longlong), len(tg), _p(tg, C.c_int),
                      _p(t0, C.c_int), _p(t1, C.c_int), ntg, ns, _p(off, C.c_int), _p(F, C.c_double),
                      _p(cnt, C.c_double))
    return F.reshape(ntg, ns), cnt


def _traj_ptrs(tr):
    x, y, z = (_c(a, np.float32) for a in (tr.x, tr.y, tr.z))
    ix, iy, iz = (_c(a, np.int16) for a in (tr.ix, tr.iy, tr.iz))
    box = _c(tr.box, np.float32)
    keep = (x, y, z, ix, iy, iz, box)
    return keep, [_p(x, C.c_float), _p(y, C.c_float), _p(z, C.c_float), _p(ix, C.c_int16), _p(iy, C.c_int16),
                  _p(iz, C.c_int16), _p(box, C.c_float)]


def fs_f64(tr, pairs, ntg, shell_off, nvec, ell, atoms=None):
    keep, ptrs = _traj_ptrs(tr)
    tg, t0, t1 = (_c(v, np.int32) for v in pairs)
    off = _c(shell_off, np.int32)
    nv = _c(nvec, np.int32).reshape(-1)
    ns = len(off) - 1
    at = None if atoms is None else _c(atoms, np.int32)
    Fs = np.zeros(ntg * ns)
    cnt = np.zeros(ntg)
    lib().orc_fs_f64(tr.N, *ptrs, _p(at, C.c_int), 0 if at is None else len(at), len(tg), _p(tg, C.c_int),
                     _p(t0, C.c_int), _p(t1, C.c_int), ntg, ns, _p(off, C.c_int), _p(nv, C.c_int), float(ell),
                     _p(Fs, C.c_double), _p(cnt, C.c_double))
    return Fs.reshape(ntg, ns), cnt


def msd_ngp_f64(tr, pairs, ntg, atoms=None):
    keep, ptrs = _traj_ptrs(tr)
    tg, t0, t1 = (_c(v, np.int32) for v in pairs)
    at = None if atoms is None else _c(atoms, np.int32)
    s2, s4, cnt, msd, ngp = (np.zeros(ntg) for _ in range(5))
    lib().orc_msd_ngp_f64(tr.N, *ptrs, _p(at, C.c_int), 0 if at is None else len(at), len(tg), _p(tg, C.c_int),
                          _p(t0, C.c_int), _p(t1, C.c_int), ntg, _p(s2, C.c_double), _p(s4, C.c_double),
                          _p(cnt, C.c_double), _p(msd, C.c_double), _p(ngp, C.c_double))
    return dict(s2=s2, s4=s4, count=cnt, msd=msd, ngp=ngp)


def disp_f64(tr, pairs, ntg, atoms=None):
    """Per gap: sums of d [ntg, 3], of d^2 [ntg, 3] and atom-pair counts."""
    keep, ptrs = _traj_ptrs(tr)
    tg, t0, t1 = (_c(v, np.int32) for v in pairs)
    at = None if atoms is None else _c(atoms, np.int32)
    sd, sd2, cnt = np.zeros(3 * ntg), np.zeros(3 * ntg), np.zeros(ntg)
    lib().orc_disp_f64(tr.N, *ptrs, _p(at, C.c_int), 0 if at is None else len(at), len(tg), _p(tg, C.c_int),
                       _p(t0, C.c_int), _p(t1, C.c_int), ntg, _p(sd, C.c_double), _p(sd2, C.c_double),
                       _p(cnt, C.c_double))
    return sd.reshape(ntg, 3), sd2.reshape(ntg, 3), cnt


def vhs_counts(tr, pairs, ntg, n_bins, max_range, atoms=None):
    """Reference-form self van Hove counts [ntg, n_bins] (uint64) and weighting [ntg]."""
    keep, ptrs = _traj_ptrs(tr)
    tg, t0, t1 = (_c(v, np.int32) for v in pairs)
    at = None if atoms is None else _c(atoms, np.int32)
    hist = np.zeros(ntg * n_bins, np.uint64)
    w = np.zeros(ntg)
    lib().orc_vhs_counts(tr.N, *ptrs, _p(at, C.c_int), 0 if at is None else len(at), len(tg), _p(tg, C.c_int),
                         _p(t0, C.c_int), _p(t1, C.c_int), ntg, n_bins, C.c_float(max_range),
                         hist.ctypes.data_as(C.POINTER(C.c_ulonglong)), _p(w, C.c_double))
    return hist.reshape(ntg, n_bins), w


def bins_assign(tr, nbins, lo, hi):
    """Rectangular bin of every (frame, atom): int32 [T, N], -1 outside."""
    x, y, z = (_c(a, np.float32) for a in (tr.x, tr.y, tr.z))
    lo = _c(lo, np.float32)
    hi = _c(hi, np.float32)
    cell = np.zeros((tr.T, tr.N), np.int32)
    lib().orc_bins_assign(tr.N, tr.T, _p(x, C.c_float), _p(y, C.c_float), _p(z, C.c_float), int(nbins[0]),
                          int(nbins[1]), int(nbins[2]), _p(lo, C.c_float), _p(hi, C.c_float), _p(cell, C.c_int))
    return cell


def _frame_lists(tr, lst):
    """(off [T+1], idx) of a list per frame: None = all atoms, an index
    array = the same atoms at every frame, a callable t -> atoms."""
    per = [np.arange(tr.N) if lst is None else (lst(t) if callable(lst) else lst) for t in range(tr.T)]
    off = np.zeros(tr.T + 1, np.int32)
    off[1:] = np.cumsum([len(a) for a in per])
    idx = np.concatenate([np.asarray(a, np.int32) for a in per]) if off[-1] else np.zeros(1, np.int32)
    return off, _c(idx, np.int32)


def bins_distance(tr, kind, thickness, n_bins, point=None, axis=0, direction=0, binned=None, cluster=None):
    """Distance bin of every (frame, atom) in the reference's FP32 form:
    int32 [T, N], -1 where not binned.  kind 0 point, 1 plane (point =
    (position,), direction 0 average / 1 above / 2 below), 2 nearest atom
    of `cluster`.  binned / cluster: None, an index array or t -> atoms."""
    x, y, z, box = (_c(a, np.float32) for a in (tr.x, tr.y, tr.z, tr.box))
    pt = _c(np.resize(np.asarray(point if point is not None else [0.0], np.float32), 3), np.float32)
    off, idx = _frame_lists(tr, binned)
    coff, cidx = _frame_lists(tr, cluster)
    cell = np.zeros((tr.T, tr.N), np.int32)
    lib().orc_bins_distance(tr.N, tr.T, _p(x, C.c_float), _p(y, C.c_float), _p(z, C.c_float), _p(box, C.c_float),
                            kind, _p(pt, C.c_float), axis, direction, _p(off, C.c_int), _p(idx, C.c_int),
                            _p(coff, C.c_int), _p(cidx, C.c_int), C.c_float(thickness), n_bins, _p(cell, C.c_int))
    return cell


def displacement_dist(tr, pairs, n_bins, max_value, atoms=None):
    """Reference-form squared-displacement histogram over the pairs:
    (counts [n_bins+1], sum d2, sum d2^2, count), d2 the FP32 msd."""
    keep, ptrs = _traj_ptrs(tr)
    tg, t0, t1 = (_c(v, np.int32) for v in pairs)
    at = None if atoms is None else _c(atoms, np.int32)
    counts = np.zeros(n_bins + 1, np.uint64)
    s2, s4, c = C.c_double(), C.c_double(), C.c_double()
    lib().orc_displacement_dist(tr.N, *ptrs, _p(at, C.c_int), 0 if at is None else len(at), len(t0),
                                _p(t0, C.c_int), _p(t1, C.c_int), n_bins, C.c_float(max_value),
                                counts.ctypes.data_as(C.POINTER(C.c_ulonglong)), C.byref(s2), C.byref(s4),
                                C.byref(c))
    return counts, s2.value, s4.value, c.value


def select_displacement(tr, nexp, nes, frt, limit, time_index, mode, lo, hi, in_list=None):
    """Fast (mode 0) / Slow (1) / Average (2) particle lists per block.
    in_list: None (all atoms), an index array (static list) or a tuple
    (conv, off, idx) (time-varying list).  Returns (conv, off, idx)."""
    keep, ptrs = _traj_ptrs(tr)
    n_times, conv_in, off_in, idx_in, n_in = 0, None, None, None, 0
    if isinstance(in_list, tuple):
        conv_in, off_in, idx_in = (_c(v, np.int32) for v in in_list)
        n_times = len(off_in) - 1
    elif in_list is not None:
        idx_in = _c(in_list, np.int32)
        n_in = len(idx_in)
    cap = tr.N * nexp + 1
    off = np.zeros(nexp + 1, np.int32)
    idx = np.zeros(cap, np.int32)
    conv = np.zeros(tr.T, np.int32)
    f = lib().orc_select_displacement
    f.restype = C.c_long
    n = f(tr.N, tr.T, *ptrs, n_times, _p(conv_in, C.c_int), _p(off_in, C.c_int), _p(idx_in, C.c_int), n_in, nexp,
          nes, frt, limit, time_index, mode, C.c_float(lo), C.c_float(hi), _p(off, C.c_int), _p(idx, C.c_int),
          C.c_long(cap), _p(conv, C.c_int))
    if n < 0:
        raise ValueError("time gap out of range")
    return conv, off, idx[:n]


# ---- reference FP32 arithmetic ----
def ref32_rho(x, y, z, nvec, length_scale, atoms=None):
    x, y, z = (_c(a, np.float32) for a in (x, y, z))
    nv = _c(nvec, np.int32).reshape(-1)
    K = len(nv) // 3
    at = None if atoms is None else _c(atoms, np.int32)
    out = np.zeros(2 * K)
    lib().orc_ref32_rho(len(x), _p(x, C.c_float), _p(y, C.c_float), _p(z, C.c_float), _p(at, C.c_int),
                        0 if at is None else len(at), K, _p(nv, C.c_int), length_scale, _p(out, C.c_double))
    return out[0::2] + 1j * out[1::2]


def ref32_sk(tr, frames, shell_off, nvec, length_scale, atoms=None):
    x, y, z = (_c(a, np.float32) for a in (tr.x, tr.y, tr.z))
    fr = _c(frames, np.int32)
    off = _c(shell_off, np.int32)
    nv = _c(nvec, np.int32).reshape(-1)
    at = None if atoms is None else _c(atoms, np.int32)
    S = np.zeros(len(off) - 1, np.float32)
    lib().orc_ref32_sk(tr.N, _p(x, C.c_float), _p(y, C.c_float), _p(z, C.c_float), _p(at, C.c_int),
                       0 if at is None else len(at), len(fr), _p(fr, C.c_int), len(off) - 1, _p(off, C.c_int),
                       _p(nv, C.c_int), length_scale, _p(S, C.c_float))
    return S


def ref32_fs(tr, pairs, ntg, shell_off, nvec, length_scale, atoms=None):
    keep, ptrs = _traj_ptrs(tr)
    off = _c(shell_off, np.int32)
    nv = _c(nvec, np.int32).reshape(-1)
    ns = len(off) - 1
    at = None if atoms is None else _c(atoms, np.int32)
    corr = np.zeros((ntg, ns), np.float32)
    natoms = np.zeros(ntg, np.float32)
    for g, a, b in zip(*pairs):
        row = np.ascontiguousarray(corr[g])
        n = lib().orc_ref32_fs_pair(tr.N, *ptrs, _p(at, C.c_int), 0 if at is None else len(at), int(a), int(b), ns,
                                    _p(off, C.c_int), _p(nv, C.c_int), length_scale, _p(row, C.c_float))
        corr[g] = row
        for _ in range(n):
            natoms[g] = np.float32(natoms[g] + np.float32(1.0))
    with np.errstate(invalid="ignore", divide="ignore"):   # time gaps without pairs: 0/0 as in the reference
        return corr / natoms[:, None], natoms


def ref32_msd_ngp(tr, pairs, ntg, atoms=None):
    """Reference-FP32 MSD and the non-Gaussian parameter's fourth-moment sum
    over the pairs in order (Mean_Square_Displacement::list_displacementkernel,
    listkernel and postprocess_list, src/mean_square_displacement.cpp:146-180;
    Non_Gaussian_Parameter's, src/non_gaussian_parameter.cpp:106-143):
    per time gap FP32 sums of d*d and pow(d, 4), the weighting (atoms per
    pair, an integer) and msd = sum / float(weighting)."""
    keep, ptrs = _traj_ptrs(tr)
    at = None if atoms is None else _c(atoms, np.int32)
    msd = np.zeros(ntg, np.float32)
    s4 = np.zeros(ntg, np.float32)
    w = np.zeros(ntg, np.int64)
    for g, a, b in zip(*pairs):
        m = np.array([msd[g]], np.float32)
        q = np.array([s4[g]], np.float32)
        n = lib().orc_ref32_msd_pair(tr.N, *ptrs, _p(at, C.c_int), 0 if at is None else len(at), int(a), int(b),
                                     _p(m, C.c_float), _p(q, C.c_float))
        msd[g], s4[g] = m[0], q[0]
        w[g] += n
    with np.errstate(invalid="ignore", divide="ignore"):   # time gaps without pairs: 0/0 as in the reference
        return dict(msd=msd / w.astype(np.float32), s2=msd, s4=s4, weighting=w)


# ---- S(k) misfit loss and gradient ----
def sk_loss_grad_f64(pos, shell_off, nvec, ell, w, target, grad=True):
    """pos [N,3] float64 (one frame).  Returns (loss, S[n_shells], grad [N,3] or None)."""
    pos = np.ascontiguousarray(pos, np.float64)
    x, y, z = (np.ascontiguousarray(pos[:, c]) for c in range(3))
    off = _c(shell_off, np.int32)
    nv = _c(nvec, np.int32).reshape(-1)
    ns = len(off) - 1
    w = _c(w, np.float64)
    t = _c(target, np.float64)
    S = np.zeros(ns)
    g = np.zeros((3, len(x))) if grad else None
    gp = [_p(np.ascontiguousarray(g[c]), C.c_double) for c in range(3)] if grad else [None, None, None]
    if grad:
        gx, gy, gz = (np.zeros(len(x)) for _ in range(3))
        gp = [_p(gx, C.c_double), _p(gy, C.c_double), _p(gz, C.c_double)]
    loss = lib().orc_sk_loss_grad_f64(len(x), _p(x, C.c_double), _p(y, C.c_double), _p(z, C.c_double), ns,
                                      _p(off, C.c_int), _p(nv, C.c_int), float(ell), _p(w, C.c_double),
                                      _p(t, C.c_double), _p(S, C.c_double), *gp)
    return loss, S, (np.stack([gx, gy, gz], 1) if grad else None)
