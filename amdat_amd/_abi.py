"""ctypes binding of the C ABI in include/amdat_hip.h.

The library is amdat_amd/lib/libamdat_hip.so, built in-tree by
``make -C amdat_amd`` (``__graft_entry__.build()``).  Loading fails loudly
when it is missing: there is no CPU fallback for any device entry point.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libamdat_hip.so")

AMDAT_OK = 0
PREC_FP32C = 0
PREC_FP64 = 1
PREC_FP32X = 2
PREC_FP64T = 3

_ERRNAMES = {-1: "EINVAL", -2: "ENOMEM", -3: "EDEVICE", -4: "ERANGE", -5: "EIO"}

# Every exported entry point: name -> (restype, argtypes).
_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_D = C.c_double
_PI = C.POINTER(C.c_int)
_PF = C.POINTER(C.c_float)
_PD = C.POINTER(C.c_double)
_PS = C.POINTER(C.c_int16)
_PP = C.POINTER(C.c_void_p)

SIGNATURES = {
    "amdat_last_error": (C.c_char_p, []),
    "amdat_version": (_I, []),
    "amdat_ctx_create": (_I, [_I, _PP]),
    "amdat_ctx_destroy": (_I, [_P]),
    "amdat_ctx_set_stream": (_I, [_P, _P]),
    "amdat_ctx_synchronize": (_I, [_P]),
    "amdat_ts_n_timesteps": (_I, [_I, _I]),
    "amdat_ts_n_timegaps": (_I, [_I, _I, _I]),
    "amdat_ts_timelist": (_I, [_I, _I, _F, _I, _F, _PF, _I]),
    "amdat_ts_displacement_times": (_I, [_I, _I, _I, _PF, _PF, _I]),
    "amdat_ts_pairs": (_I, [_I, _I, _I, _I, _I, _PI, _PI, _PI, _I, _PI]),
    "amdat_ts_weighting": (_I, [_I, _I, _I, _I, _I, _PI, _I]),
    "amdat_shells_read": (_I, [C.c_char_p, _I, _I, _PI, _PI, _I, _PI]),
    "amdat_shells_read_plane": (_I, [C.c_char_p, C.c_char_p, _I, _I, _PI, _PI, _I, _PI, _PI]),
    "amdat_shells_generate": (_I, [_I, _I, _I, C.c_uint64, _PI, _PI, _I, _PI]),
    "amdat_shells_upload": (_I, [_P, _I, _PI, _PI, _PP]),
    "amdat_shells_free": (_I, [_P, _P]),
    "amdat_shells_plan_hint": (_I, [_P, _I, _I]),
    "amdat_traj_upload": (_I, [_P, _I, _I, _PF, _PF, _PF, _PS, _PS, _PS, _PF, _PP]),
    "amdat_traj_wrap_device": (_I, [_P, _I, _I, _P, _P, _P, _P, _P, _P, _PF, _PP]),
    "amdat_traj_free": (_I, [_P, _P]),
    "amdat_read_custom": (_I, [C.c_char_p, _I, _I, _I, _PF, _PF, _PF, _PS, _PS, _PS, _PF, _PF, _PI]),
    "amdat_read_xyz": (_I, [C.c_char_p, _I, _I, _PF, _I, _PF, _PF, _PF, _PS, _PS, _PS, _PI]),
    "amdat_traj_read": (_I, [_P, C.c_char_p, C.c_char_p, _I, _I, _PF, _I, _PI, _PP]),
    "amdat_traj_shape": (_I, [_P, _PI, _PI, _PF]),
    "amdat_traj_bounds": (_I, [_P, _PF]),
    "amdat_bins_assign": (_I, [_P, _P, _I, _I, _I, _PF, _PF, _PI]),
    "amdat_bins_distance": (_I, [_P, _P, _P, _I, _PF, _I, _I, _P, _F, _I, _PI]),
    "amdat_cells_assign": (_I, [_P, _P, _I, _I, _I, _PF, _PF, _PP]),
    "amdat_cells_distance": (_I, [_P, _P, _P, _I, _PF, _I, _I, _P, _F, _I, _PP]),
    "amdat_cells_download": (_I, [_P, _P, _PI]),
    "amdat_cells_free": (_I, [_P, _P]),
    "amdat_cells_bin_lists": (_I, [_P, _P, _P, _I, _I, _PP]),
    "amdat_cells_pair_lists": (_I, [_P, _P, _P, _I, _PI, _PI, _I, _I, _PP]),
    "amdat_list_create": (_I, [_P, _P, _I, _PI, _PP]),
    "amdat_list_free": (_I, [_P, _P]),
    "amdat_list_create_timed": (_I, [_P, _P, _I, _PI, _PI, _PI, _PP]),
    "amdat_list_create_paired": (_I, [_P, _P, _I, _PI, _PI, _PI, _PI, _PP]),
    "amdat_list_shape": (_I, [_P, _PI, _PI]),
    "amdat_list_download": (_I, [_P, _P, _PI, _PI, _PI]),
    "amdat_list_select_displacement": (_I, [_P, _P, _P, _I, _I, _I, _I, _I, _I, _F, _F, _PP]),
    "amdat_pairs_upload": (_I, [_P, _I, _PI, _PI, _PI, _I, _PP]),
    "amdat_pairs_free": (_I, [_P, _P]),
    "amdat_rho_compute": (_I, [_P, _P, _P, _P, _I, _I, _F, _PI, _I, _I, _PP]),
    "amdat_rho_download": (_I, [_P, _P, _PD, C.c_size_t]),
    "amdat_rho_shape": (_I, [_P, _PI, _PI]),
    "amdat_rho_free": (_I, [_P, _P]),
    "amdat_sk": (_I, [_P, _P, _P, _PD]),
    "amdat_fkt": (_I, [_P, _P, _P, _P, _I, _I, _PD, _PD]),
    "amdat_fs": (_I, [_P, _P, _P, _P, _I, _I, _F, _P, _I, _PD, _PD]),
    "amdat_msd_ngp": (_I, [_P, _P, _P, _P, _PD, _PD, _PD]),
    "amdat_msd_ngp_noalloc": (_I, [_P, _P, _P, _P, _PD, _PD, _PD]),
    "amdat_fkt_vectors": (_I, [_P, _P, _P, _P, _I, _PI, _PI, _PD, _PD]),
    "amdat_rho_sum": (_I, [_P, _P, _P, C.POINTER(_P)]),
    "amdat_displacements": (_I, [_P, _P, _P, _P, _I, _F, _PD, _PD, C.POINTER(C.c_ulonglong), _PD]),
    "amdat_displacement_dist": (_I, [_P, _P, _P, _P, _I, _F, C.POINTER(C.c_ulonglong), _PD, _PD, _PD]),
    "amdat_ts_gap_pairs": (_I, [_I, _I, _I, _I, _I, _PI, _PI, _PI, _I, _PI]),
    "amdat_sk_loss_grad": (_I, [_P, _I, _P, _P, _P, _P, _I, _I, _F, _PD, _PD, _PD, _PD, _P, _P, _P]),
    "amdat_batch_create": (_I, [_P, _I, _PP, _PP, _PP]),
    "amdat_batch_free": (_I, [_P, _P]),
    "amdat_batch_sk": (_I, [_P, _P, _P, _I, _I, _PF, _PI, _I, _PD, _PD]),
    "amdat_batch_fs": (_I, [_P, _P, _P, _PI, _PI, _PF, _P, _PD, _PD]),
    "amdat_batch_msd_ngp": (_I, [_P, _P, _P, _PD, _PD, _PD]),
    "amdat_batch_run": (_I, [_P, _P, _PF, _P, _I, _I, _PI, _I, _P, _PI, _PI, _P, _I, _I,
                             _PD, _PD, _PD, _PD, _PD, _PD, _PD]),
    "amdat_batch_job_info": (_I, [_P, _PI, _PI, _PI, _PI, _PI]),
    "amdat_batch_job_settings": (_I, [_P, _P]),
    "amdat_batch_use_settings": (_I, [_P, _P]),
}


class JobSettings(C.Structure):
    """amdat_job_settings (include/amdat_hip.h)."""
    _fields_ = [(n, C.c_int) for n in ("sk_R", "sk_plan_mode", "sk_frames_per_item", "sk_frame_groups", "fs_R",
                                       "fs_pairs_per_item", "fs_atom_parts", "fs_pair_chunks", "atoms_per_tile",
                                       "table_half_width", "n_items", "grid", "graph", "tuned")] + \
               [("model_ms", C.c_float), ("tuned_ms", C.c_float)]

RUN_GRAPH = 1
RUN_TUNE = 2


class AmdatError(RuntimeError):
    """A C-ABI call returned a negative status."""

    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (_ERRNAMES.get(code, "?"), code, msg))
        self.code = code


_lib = None


def lib():
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libamdat_hip.so is not built (%s); run __graft_entry__.build() "
                               "or make -C amdat_amd" % LIB_PATH)
        # torch ships its own HIP runtime under the same soname.  Loading it
        # first makes this library bind to that one runtime, so torch tensors
        # and this library can share a process in either import order.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc < 0:
        raise AmdatError(rc, lib().amdat_last_error().decode())
    return rc


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


def ip(a):
    return _p(a, C.c_int)


def fp(a):
    return _p(a, C.c_float)


def dp(a):
    return _p(a, C.c_double)


def sp(a):
    return _p(a, C.c_int16)


# ---- host integer work --------------------------------------------------

def ts_n_timegaps(nexp, nes, frt=0):
    return check(lib().amdat_ts_n_timegaps(nexp, nes, frt))


def ts_timelist(nexp, nes, base, first_exp=0, dt=1.0):
    T = check(lib().amdat_ts_n_timesteps(nexp, nes))
    out = np.zeros(T, np.float32)
    check(lib().amdat_ts_timelist(nexp, nes, base, first_exp, dt, fp(out), T))
    return out


def ts_displacement_times(nexp, nes, frt, timelist):
    ntg = ts_n_timegaps(nexp, nes, frt)
    tl = np.ascontiguousarray(timelist, np.float32)
    out = np.zeros(ntg, np.float32)
    check(lib().amdat_ts_displacement_times(nexp, nes, frt, fp(tl), fp(out), ntg))
    return out


def ts_pairs(nexp, nes, frt=0, fullblock=0, limit=0):
    n = C.c_int(0)
    check(lib().amdat_ts_pairs(nexp, nes, frt, fullblock, limit, None, None, None, 0, C.byref(n)))
    tg = np.zeros(n.value, np.int32)
    t0 = np.zeros(n.value, np.int32)
    t1 = np.zeros(n.value, np.int32)
    check(lib().amdat_ts_pairs(nexp, nes, frt, fullblock, limit, ip(tg), ip(t0), ip(t1), n.value, C.byref(n)))
    return tg, t0, t1


def ts_weighting(nexp, nes, frt=0, fullblock=0, limit=0):
    ntg = ts_n_timegaps(nexp, nes, frt)
    out = np.zeros(ntg, np.int32)
    check(lib().amdat_ts_weighting(nexp, nes, frt, fullblock, limit, ip(out), ntg))
    return out


def shells_read(stem, first_bin, n_bins, cap=400000):
    off = np.zeros(n_bins + 1, np.int32)
    nvec = np.zeros(3 * cap, np.int32)
    n = C.c_int(0)
    check(lib().amdat_shells_read(stem.encode(), first_bin, n_bins, ip(off), ip(nvec), cap, C.byref(n)))
    return off, nvec[: 3 * n.value].reshape(-1, 3).copy()


def shells_read_plane(stem, plane, first_bin, n_bins, cap=400000):
    """(offsets, integer vectors on x, y, z, denominator): k = 2 pi n / (denominator * l)."""
    off = np.zeros(n_bins + 1, np.int32)
    nvec = np.zeros(3 * cap, np.int32)
    n, den = C.c_int(0), C.c_int(0)
    check(lib().amdat_shells_read_plane(stem.encode(), plane.encode(), first_bin, n_bins, ip(off), ip(nvec), cap,
                                        C.byref(n), C.byref(den)))
    return off, nvec[: 3 * n.value].reshape(-1, 3).copy(), den.value


def shells_generate(first_bin, n_bins, max_per_shell=300, seed=0xA11DA7, cap=2000000):
    off = np.zeros(n_bins + 1, np.int32)
    n = C.c_int(0)
    nvec = np.zeros(3 * cap, np.int32)
    check(lib().amdat_shells_generate(first_bin, n_bins, max_per_shell, seed, ip(off), ip(nvec), cap, C.byref(n)))
    return off, nvec[: 3 * n.value].reshape(-1, 3).copy()


# ---- device objects -------------------------------------------------------

class Context:
    """One device context (amdat_ctx)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().amdat_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().amdat_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(lib().amdat_ctx_synchronize(self.h))

    def set_stream(self, stream_ptr):
        check(lib().amdat_ctx_set_stream(self.h, C.c_void_p(stream_ptr) if stream_ptr else None))


class _Handle:
    _free = None

    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = h

    def close(self):
        if self.h and self.ctx is not None and self.ctx.h:
            getattr(lib(), self._free)(self.ctx.h, self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Shells(_Handle):
    _free = "amdat_shells_free"

    def __init__(self, ctx, shell_off, nvec):
        off = np.ascontiguousarray(shell_off, np.int32)
        nv = np.ascontiguousarray(nvec, np.int32).reshape(-1)
        h = C.c_void_p()
        check(lib().amdat_shells_upload(ctx.h, len(off) - 1, ip(off), ip(nv), C.byref(h)))
        super().__init__(ctx, h)
        self.off = off
        self.nvec = nv.reshape(-1, 3)

    def plan_hint(self, R=0, pair=0):
        """Fix vectors per thread and pairing of the phase kernels (R = 0: automatic)."""
        check(lib().amdat_shells_plan_hint(self.h, R, pair))


class Trajectory(_Handle):
    """Device-resident trajectory: SoA [T][N] FP32 wrapped + int16 images."""
    _free = "amdat_traj_free"

    def __init__(self, ctx, x, y, z, ix, iy, iz, box):
        x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
        ix, iy, iz = (np.ascontiguousarray(a, np.int16) for a in (ix, iy, iz))
        box = np.ascontiguousarray(box, np.float32).reshape(-1, 3)
        T, N = x.shape
        h = C.c_void_p()
        check(lib().amdat_traj_upload(ctx.h, N, T, fp(x), fp(y), fp(z), sp(ix), sp(iy), sp(iz), fp(box), C.byref(h)))
        super().__init__(ctx, h)
        self.N, self.T = N, T

    @classmethod
    def from_device(cls, ctx, N, T, ptrs, box):
        """Wrap device buffers (x, y, z, ix, iy, iz data pointers) without copying."""
        self = cls.__new__(cls)
        box = np.ascontiguousarray(box, np.float32).reshape(-1, 3)
        h = C.c_void_p()
        check(lib().amdat_traj_wrap_device(ctx.h, N, T, *[C.c_void_p(p) for p in ptrs], fp(box), C.byref(h)))
        _Handle.__init__(self, ctx, h)
        self.N, self.T = N, T
        return self


    def box(self):
        """Per-frame box lengths [T, 3] (host copy)."""
        b = np.zeros((self.T, 3), np.float32)
        check(lib().amdat_traj_shape(self.h, None, None, fp(b)))
        return b

    @classmethod
    def read(cls, ctx, fmt, path, N, T, box3=None, threads=0):
        """Read a "custom" or "xyz" trajectory file straight into device
        memory.  Returns (trajectory, types of frame 0)."""
        self = cls.__new__(cls)
        b = None if box3 is None else np.ascontiguousarray(box3, np.float32)
        types = np.zeros(N, np.int32)
        h = C.c_void_p()
        check(lib().amdat_traj_read(ctx.h, fmt.encode(), str(path).encode(), N, T, fp(b), threads, ip(types),
                                    C.byref(h)))
        _Handle.__init__(self, ctx, h)
        self.N, self.T = N, T
        return self, types


def read_custom(path, N, T, threads=0):
    """Host SoA arrays of a LAMMPS custom dump: dict of x y z ix iy iz [T, N],
    box [T, 3], lo [T, 3], types [N]."""
    out = {k: np.zeros((T, N), np.float32) for k in ("x", "y", "z")}
    out.update({k: np.zeros((T, N), np.int16) for k in ("ix", "iy", "iz")})
    out["box"] = np.zeros((T, 3), np.float32)
    out["lo"] = np.zeros((T, 3), np.float32)
    out["types"] = np.zeros(N, np.int32)
    check(lib().amdat_read_custom(str(path).encode(), N, T, threads, fp(out["x"]), fp(out["y"]), fp(out["z"]),
                                  sp(out["ix"]), sp(out["iy"]), sp(out["iz"]), fp(out["box"]), fp(out["lo"]),
                                  ip(out["types"])))
    return out


def read_xyz(path, N, T, box3, threads=0):
    """Host SoA arrays of an xyz trajectory with image counts from the
    reference's unwrapping rule: dict of x y z ix iy iz [T, N], types [N]."""
    out = {k: np.zeros((T, N), np.float32) for k in ("x", "y", "z")}
    out.update({k: np.zeros((T, N), np.int16) for k in ("ix", "iy", "iz")})
    out["types"] = np.zeros(N, np.int32)
    b = np.ascontiguousarray(box3, np.float32)
    check(lib().amdat_read_xyz(str(path).encode(), N, T, fp(b), threads, fp(out["x"]), fp(out["y"]), fp(out["z"]),
                               sp(out["ix"]), sp(out["iy"]), sp(out["iz"]), ip(out["types"])))
    return out


class AtomList(_Handle):
    _free = "amdat_list_free"

    def __init__(self, ctx, traj, indices=None):
        h = C.c_void_p()
        if indices is None:
            check(lib().amdat_list_create(ctx.h, traj.h, traj.N, None, C.byref(h)))
            self.n = traj.N
        else:
            idx = np.ascontiguousarray(indices, np.int32)
            check(lib().amdat_list_create(ctx.h, traj.h, len(idx), ip(idx), C.byref(h)))
            self.n = len(idx)
        super().__init__(ctx, h)


SELECT_FAST, SELECT_SLOW, SELECT_BETWEEN = 0, 1, 2


class TimedList(_Handle):
    """Time-varying atom list: list time c holds indices[off[c]:off[c+1]];
    conv[t] is the list time of frame t (outside [0, n_times): no atoms)."""
    _free = "amdat_list_free"

    def __init__(self, ctx, traj, conv, off, indices, _h=None):
        if _h is None:
            conv = np.ascontiguousarray(conv, np.int32)
            off = np.ascontiguousarray(off, np.int32)
            idx = np.ascontiguousarray(indices, np.int32)
            assert len(conv) == traj.T
            h = C.c_void_p()
            check(lib().amdat_list_create_timed(ctx.h, traj.h, len(off) - 1, ip(conv), ip(off), ip(idx),
                                                C.byref(h)))
        else:
            h = _h
        super().__init__(ctx, h)
        self.T = traj.T
        nt, n = C.c_int(), C.c_int()
        check(lib().amdat_list_shape(self.h, C.byref(nt), C.byref(n)))
        self.n_times, self.n = nt.value, n.value

    @classmethod
    def select_displacement(cls, ctx, traj, in_list, nexp, nes, frt, limit, time_index, mode, lo, hi):
        """Fast/Slow/Average_Particles selection per exponential block."""
        h = C.c_void_p()
        check(lib().amdat_list_select_displacement(ctx.h, traj.h, in_list.h if in_list is not None else None, nexp,
                                                   nes, frt, limit, time_index, mode, float(lo), float(hi),
                                                   C.byref(h)))
        return cls(ctx, traj, None, None, None, _h=h)

    def download(self):
        """(conv [T], off [n_times+1], indices [n])."""
        conv = np.zeros(self.T, np.int32)
        off = np.zeros(self.n_times + 1, np.int32)
        idx = np.zeros(max(self.n, 1), np.int32)
        check(lib().amdat_list_download(self.ctx.h, self.h, ip(conv), ip(off), ip(idx)))
        return conv, off, idx[:self.n]


class PairList(_Handle):
    """Pair-keyed atom list: the atoms indices[off[c]:off[c+1]] apply to
    the frame pair (t0[c], t1[c]); pairs without a key have no atoms."""
    _free = "amdat_list_free"

    def __init__(self, ctx, traj, t0, t1, off, indices, _h=None):
        if _h is None:
            t0, t1, off, idx = (np.ascontiguousarray(a, np.int32) for a in (t0, t1, off, indices))
            assert len(t0) == len(t1) == len(off) - 1
            h = C.c_void_p()
            check(lib().amdat_list_create_paired(ctx.h, traj.h, len(t0), ip(t0), ip(t1), ip(off), ip(idx),
                                                 C.byref(h)))
        else:
            h = _h
        super().__init__(ctx, h)
        nt, n = C.c_int(), C.c_int()
        check(lib().amdat_list_shape(self.h, C.byref(nt), C.byref(n)))
        self.n_keys, self.n = nt.value, n.value

    def download(self):
        """(off [n_keys+1], indices [n])."""
        off = np.zeros(self.n_keys + 1, np.int32)
        idx = np.zeros(max(self.n, 1), np.int32)
        check(lib().amdat_list_download(self.ctx.h, self.h, None, ip(off), ip(idx)))
        return off, idx[:self.n]


class Cells(_Handle):
    """A cell table [T][N] kept on the device (amdat_cells_assign /
    amdat_cells_distance), and the per-bin lists formed from it."""
    _free = "amdat_cells_free"

    def __init__(self, ctx, traj, h, n_cells):
        super().__init__(ctx, h)
        self.traj, self.n_cells = traj, n_cells

    @classmethod
    def assign(cls, ctx, traj, nbins, lo, hi):
        lo = np.ascontiguousarray(lo, np.float32)
        hi = np.ascontiguousarray(hi, np.float32)
        h = C.c_void_p()
        check(lib().amdat_cells_assign(ctx.h, traj.h, int(nbins[0]), int(nbins[1]), int(nbins[2]), fp(lo), fp(hi),
                                       C.byref(h)))
        return cls(ctx, traj, h, int(nbins[0]) * int(nbins[1]) * int(nbins[2]))

    @classmethod
    def distance(cls, ctx, traj, binned, kind, thickness, n_bins, point=None, axis=0, direction=0, cluster=None):
        pt = np.zeros(3, np.float32)
        if point is not None:
            pt[:len(point)] = point
        h = C.c_void_p()
        check(lib().amdat_cells_distance(ctx.h, traj.h, binned.h if binned is not None else None, kind, fp(pt), axis,
                                         direction, cluster.h if cluster is not None else None, float(thickness),
                                         int(n_bins), C.byref(h)))
        return cls(ctx, traj, h, int(n_bins))

    def download(self):
        cell = np.zeros((self.traj.T, self.traj.N), np.int32)
        check(lib().amdat_cells_download(self.ctx.h, self.h, ip(cell)))
        return cell

    def bin_lists(self, alist=None, first=0, n=None):
        """TimedList per bin first .. first+n-1 (one list time per frame)."""
        n = self.n_cells - first if n is None else n
        hs = (C.c_void_p * n)()
        check(lib().amdat_cells_bin_lists(self.ctx.h, self.h, alist.h if alist is not None else None, first, n,
                                          C.cast(hs, _PP)))
        return [TimedList(self.ctx, self.traj, None, None, None, _h=C.c_void_p(h)) for h in hs]

    def pair_lists(self, t0, t1, alist=None, first=0, n=None):
        """PairList per bin, keys (t0[k], t1[k])."""
        n = self.n_cells - first if n is None else n
        t0, t1 = np.ascontiguousarray(t0, np.int32), np.ascontiguousarray(t1, np.int32)
        hs = (C.c_void_p * n)()
        check(lib().amdat_cells_pair_lists(self.ctx.h, self.h, alist.h if alist is not None else None, len(t0),
                                           ip(t0), ip(t1), first, n, C.cast(hs, _PP)))
        return [PairList(self.ctx, self.traj, None, None, None, None, _h=C.c_void_p(h)) for h in hs]


class Pairs(_Handle):
    _free = "amdat_pairs_free"

    def __init__(self, ctx, tg, t0, t1, n_timegaps):
        tg, t0, t1 = (np.ascontiguousarray(a, np.int32) for a in (tg, t0, t1))
        h = C.c_void_p()
        check(lib().amdat_pairs_upload(ctx.h, len(tg), ip(tg), ip(t0), ip(t1), n_timegaps, C.byref(h)))
        super().__init__(ctx, h)
        self.n = len(tg)
        self.ntg = n_timegaps


class Rho(_Handle):
    _free = "amdat_rho_free"

    def __init__(self, ctx, traj, alist, shells, first, last, length_scale, frames=None, precision=PREC_FP32C):
        h = C.c_void_p()
        fr = None if frames is None else np.ascontiguousarray(frames, np.int32)
        check(lib().amdat_rho_compute(ctx.h, traj.h, alist.h if alist is not None else None, shells.h, first, last,
                                      length_scale, ip(fr), 0 if fr is None else len(fr), precision, C.byref(h)))
        super().__init__(ctx, h)
        nf, nv = C.c_int(), C.c_int()
        check(lib().amdat_rho_shape(h, C.byref(nf), C.byref(nv)))
        self.n_frames, self.K = nf.value, nv.value
        self.n_shells = last - first + 1

    def download(self):
        out = np.zeros((self.n_frames, self.K, 2), np.float64)
        check(lib().amdat_rho_download(self.ctx.h, self.h, dp(out), out.size))
        return out[..., 0] + 1j * out[..., 1]


def sk(ctx, rho1, rho2=None):
    S = np.zeros(rho1.n_shells, np.float64)
    check(lib().amdat_sk(ctx.h, rho1.h, (rho2 or rho1).h, dp(S)))
    return S


def fkt(ctx, rho1, rho2, pairs, wn1=0, wn2=None):
    wn2 = rho1.n_shells - 1 if wn2 is None else wn2
    F = np.zeros((pairs.ntg, wn2 - wn1 + 1), np.float64)
    cnt = np.zeros(pairs.ntg, np.float64)
    check(lib().amdat_fkt(ctx.h, rho1.h, rho2.h, pairs.h, wn1, wn2, dp(F), dp(cnt)))
    return F, cnt


def fkt_vectors(ctx, rho1, rho2, pairs, v_first, v_count):
    """F over arbitrary vector ranges: column c sums vectors
    v_first[c] .. v_first[c] + v_count[c] - 1 of the densities."""
    vf = np.ascontiguousarray(v_first, np.int32)
    vc = np.ascontiguousarray(v_count, np.int32)
    F = np.zeros((pairs.ntg, len(vf)), np.float64)
    cnt = np.zeros(pairs.ntg, np.float64)
    check(lib().amdat_fkt_vectors(ctx.h, rho1.h, rho2.h, pairs.h, len(vf), ip(vf), ip(vc), dp(F), dp(cnt)))
    return F, cnt


def rho_sum(ctx, a, b):
    """A new density a + b (same frames and vectors; atom counts added)."""
    h = C.c_void_p()
    check(lib().amdat_rho_sum(ctx.h, a.h, b.h, C.byref(h)))
    r = Rho.__new__(Rho)
    _Handle.__init__(r, ctx, h)
    r.n_frames, r.K, r.n_shells = a.n_frames, a.K, a.n_shells
    return r


def fs(ctx, traj, alist, shells, first, last, length_scale, pairs, precision=PREC_FP32C):
    F = np.zeros((pairs.ntg, last - first + 1), np.float64)
    cnt = np.zeros(pairs.ntg, np.float64)
    check(lib().amdat_fs(ctx.h, traj.h, alist.h if alist is not None else None, shells.h, first, last,
                         length_scale, pairs.h, precision, dp(F), dp(cnt)))
    return F, cnt


def msd_ngp(ctx, traj, alist, pairs):
    s2 = np.zeros(pairs.ntg, np.float64)
    s4 = np.zeros(pairs.ntg, np.float64)
    cnt = np.zeros(pairs.ntg, np.float64)
    check(lib().amdat_msd_ngp(ctx.h, traj.h, alist.h if alist is not None else None, pairs.h, dp(s2), dp(s4),
                              dp(cnt)))
    return s2, s4, cnt


def msd_ngp_noalloc(ctx, traj, alist, pairs):
    """msd_ngp through the entry point that allocates nothing (static lists,
    bounded by the library's per-GPU scratch)."""
    s2 = np.zeros(pairs.ntg, np.float64)
    s4 = np.zeros(pairs.ntg, np.float64)
    cnt = np.zeros(pairs.ntg, np.float64)
    check(lib().amdat_msd_ngp_noalloc(ctx.h, traj.h, alist.h if alist is not None else None, pairs.h, dp(s2), dp(s4),
                                      dp(cnt)))
    return s2, s4, cnt


def displacements(ctx, traj, alist, pairs, n_bins=0, max_range=0.0):
    """Per time gap: sums of d [ntg, 3], of d^2 [ntg, 3], counts [ntg] and
    (n_bins > 0) the van Hove histogram [ntg, n_bins] (uint64)."""
    sd = np.zeros((pairs.ntg, 3))
    sd2 = np.zeros((pairs.ntg, 3))
    cnt = np.zeros(pairs.ntg)
    hist = np.zeros((pairs.ntg, max(n_bins, 1)), np.uint64)
    check(lib().amdat_displacements(ctx.h, traj.h, alist.h if alist is not None else None, pairs.h, n_bins,
                                    max_range, dp(sd), dp(sd2), hist.ctypes.data_as(C.POINTER(C.c_ulonglong)),
                                    dp(cnt)))
    return sd, sd2, cnt, (hist if n_bins > 0 else None)


def sk_loss_grad(ctx, n_atoms, x_ptr, y_ptr, z_ptr, shells, first, last, length_scale, weights, target,
                 grad_ptrs=None):
    """S(k)-misfit loss on one frame held in device memory (pointers to FP32
    [n_atoms] arrays).  grad_ptrs: three device FP64 [n_atoms] outputs, or
    None for the loss only.  Returns (loss, S per shell)."""
    nsh = last - first + 1
    w = np.ascontiguousarray(weights, np.float64)
    t = np.ascontiguousarray(target, np.float64)
    assert len(w) == nsh and len(t) == nsh
    S = np.zeros(nsh)
    loss = C.c_double()
    g = [C.c_void_p(p) for p in grad_ptrs] if grad_ptrs else [None, None, None]
    check(lib().amdat_sk_loss_grad(ctx.h, n_atoms, C.c_void_p(x_ptr), C.c_void_p(y_ptr), C.c_void_p(z_ptr),
                                   shells.h, first, last, length_scale, dp(w), dp(t), C.byref(loss), dp(S), *g))
    return loss.value, S


class Batch(_Handle):
    """Many trajectories (same frame count) analysed by single persistent
    launches.  Keeps references to its members so they outlive the handle."""
    _free = "amdat_batch_free"

    def __init__(self, ctx, trajs, lists=None):
        n = len(trajs)
        th = (C.c_void_p * n)(*[t.h for t in trajs])
        lh = None
        if lists is not None:
            lh = (C.c_void_p * n)(*[(l.h if l is not None else None) for l in lists])
        h = C.c_void_p()
        check(lib().amdat_batch_create(ctx.h, n, th, lh, C.byref(h)))
        super().__init__(ctx, h)
        self.members = list(trajs)
        self.lists = list(lists) if lists is not None else None
        self.n = n

    def sk(self, shells, first, last, length_scales, frames=None):
        """S[n_traj][n_shells] and count[n_traj] (= frames x atoms)."""
        nsh = last - first + 1
        ell = np.ascontiguousarray(length_scales, np.float32)
        assert len(ell) == self.n
        fr = None if frames is None else np.ascontiguousarray(frames, np.int32)
        S = np.zeros((self.n, nsh), np.float64)
        cnt = np.zeros(self.n, np.float64)
        check(lib().amdat_batch_sk(self.ctx.h, self.h, shells.h, first, last, fp(ell), ip(fr),
                                   0 if fr is None else len(fr), dp(S), dp(cnt)))
        return S, cnt

    def fs(self, shells, first, last, length_scales, pairs):
        """Per member i: Fs_i [n_timegaps][last[i]-first[i]+1]; count [n_traj][n_timegaps]."""
        fi = np.ascontiguousarray(first, np.int32)
        la = np.ascontiguousarray(last, np.int32)
        ell = np.ascontiguousarray(length_scales, np.float32)
        assert len(fi) == len(la) == len(ell) == self.n
        nsh = la - fi + 1
        out = np.zeros(int(np.sum(nsh)) * pairs.ntg, np.float64)
        cnt = np.zeros((self.n, pairs.ntg), np.float64)
        check(lib().amdat_batch_fs(self.ctx.h, self.h, shells.h, ip(fi), ip(la), fp(ell), pairs.h, dp(out),
                                   dp(cnt)))
        res, o = [], 0
        for k in nsh:
            res.append(out[o:o + pairs.ntg * k].reshape(pairs.ntg, k))
            o += pairs.ntg * k
        return res, cnt

    def msd_ngp(self, pairs):
        """s2, s4, count: [n_traj][n_timegaps]."""
        s2 = np.zeros((self.n, pairs.ntg), np.float64)
        s4 = np.zeros((self.n, pairs.ntg), np.float64)
        cnt = np.zeros((self.n, pairs.ntg), np.float64)
        check(lib().amdat_batch_msd_ngp(self.ctx.h, self.h, pairs.h, dp(s2), dp(s4), dp(cnt)))
        return s2, s4, cnt

    def run(self, length_scales, sk=None, frames=None, fs=None, pairs=None, msd=False, graph=False, tune=False):
        """One job in one persistent launch.  sk = (shells, first, last);
        fs = (shells, first[], last[]); pairs serves Fs and MSD.  Returns a
        dict with the outputs of sk(), fs() and msd_ngp() for the parts asked."""
        n = self.n
        ell = np.ascontiguousarray(length_scales, np.float32)
        assert len(ell) == n
        out = {}
        S = Sc = Fs = Fc = s2 = s4 = mc = None
        sk_h, f0, f1, fr = None, 0, -1, None
        if sk is not None:
            sk_h, f0, f1 = sk[0].h, sk[1], sk[2]
            S = np.zeros((n, f1 - f0 + 1), np.float64)
            Sc = np.zeros(n, np.float64)
            fr = None if frames is None else np.ascontiguousarray(frames, np.int32)
        fs_h, fi, la, nsh = None, None, None, None
        ntg = pairs.ntg if pairs is not None else 0
        if fs is not None:
            fs_h = fs[0].h
            fi = np.ascontiguousarray(fs[1], np.int32)
            la = np.ascontiguousarray(fs[2], np.int32)
            nsh = la - fi + 1
            Fs = np.zeros(int(np.sum(nsh)) * ntg, np.float64)
            Fc = np.zeros((n, ntg), np.float64)
        if msd:
            s2 = np.zeros((n, ntg), np.float64)
            s4 = np.zeros((n, ntg), np.float64)
            mc = np.zeros((n, ntg), np.float64)
        flags = (RUN_GRAPH if graph else 0) | (RUN_TUNE if tune else 0)
        check(lib().amdat_batch_run(self.ctx.h, self.h, fp(ell), sk_h, f0, f1, ip(fr), 0 if fr is None else len(fr),
                                    fs_h, ip(fi), ip(la), pairs.h if pairs is not None else None, int(bool(msd)),
                                    flags, dp(S), dp(Sc), dp(Fs), dp(Fc), dp(s2), dp(s4), dp(mc)))
        if sk is not None:
            out["S"], out["S_count"] = S, Sc
        if fs is not None:
            res, o = [], 0
            for k in nsh:
                res.append(Fs[o:o + ntg * k].reshape(ntg, k))
                o += ntg * k
            out["Fs"], out["Fs_count"] = res, Fc
        if msd:
            out["s2"], out["s4"], out["msd_count"] = s2, s4, mc
        return out

    def job_info(self):
        rs, rf, ps, ni, g = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().amdat_batch_job_info(self.h, C.byref(rs), C.byref(rf), C.byref(ps), C.byref(ni), C.byref(g)))
        return dict(RS=rs.value, RF=rf.value, paired=bool(ps.value), n_items=ni.value, graph=bool(g.value))

    def job_settings(self):
        """Every launch setting of the last job (amdat_batch_job_settings)."""
        st = JobSettings()
        check(lib().amdat_batch_job_settings(self.h, C.byref(st)))
        out = {n: getattr(st, n) for n, _ in JobSettings._fields_}
        out["graph"], out["tuned"] = bool(out["graph"]), bool(out["tuned"])
        return out

    def use_settings(self, **kw):
        """Fix launch settings of the next jobs (amdat_batch_use_settings):
        sk_R, sk_plan_mode, sk_frames_per_item, fs_R, fs_pairs_per_item,
        fs_atom_parts (omitted: the cost model's).  No keywords: back to the
        cost model and tuner."""
        if not kw:
            check(lib().amdat_batch_use_settings(self.h, None))
            return
        st = JobSettings()
        for k, v in kw.items():
            if k not in ("sk_R", "sk_plan_mode", "sk_frames_per_item", "fs_R", "fs_pairs_per_item", "fs_atom_parts"):
                raise TypeError("unknown setting %r" % k)
            setattr(st, k, int(v))
        check(lib().amdat_batch_use_settings(self.h, C.byref(st)))


def ts_gap_pairs(nexp, nes, frt, limit, time_index):
    """Pairs (tg, t0, t1) of one time gap, first frame of each block only."""
    n = C.c_int()
    check(lib().amdat_ts_gap_pairs(nexp, nes, frt, limit, time_index, None, None, None, 0, C.byref(n)))
    tg, t0, t1 = (np.zeros(max(n.value, 1), np.int32) for _ in range(3))
    check(lib().amdat_ts_gap_pairs(nexp, nes, frt, limit, time_index, ip(tg), ip(t0), ip(t1), n.value, C.byref(n)))
    return tg[:n.value], t0[:n.value], t1[:n.value]


def displacement_dist(ctx, traj, alist, pairs, n_bins, max_value):
    """Squared-displacement distribution: (counts [n_bins+1] uint64,
    sum |dr|^2, sum |dr|^4, atom-pair count)."""
    counts = np.zeros(n_bins + 1, np.uint64)
    s2, s4, c = C.c_double(), C.c_double(), C.c_double()
    check(lib().amdat_displacement_dist(ctx.h, traj.h, alist.h if alist is not None else None, pairs.h, n_bins,
                                        float(max_value), counts.ctypes.data_as(C.POINTER(C.c_ulonglong)),
                                        C.byref(s2), C.byref(s4), C.byref(c)))
    return counts, s2.value, s4.value, c.value


def bins_assign(ctx, traj, nbins, lo, hi):
    """Cell of every (frame, atom) in a rectangular grid of nbins = (nx, ny, nz)
    over (lo, hi) [T, 3] per frame: int32 [T, N], -1 outside."""
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    cell = np.zeros((traj.T, traj.N), np.int32)
    check(lib().amdat_bins_assign(ctx.h, traj.h, int(nbins[0]), int(nbins[1]), int(nbins[2]), fp(lo), fp(hi),
                                  ip(cell)))
    return cell


DIST_POINT, DIST_PLANE, DIST_LIST = 0, 1, 2


def bins_distance(ctx, traj, binned, kind, thickness, n_bins, point=None, axis=0, direction=0, cluster=None):
    """Distance bin of every (frame, atom) of `binned` (None = all atoms):
    int32 [T, N], -1 where not binned.  kind DIST_POINT (point = (x, y, z)),
    DIST_PLANE (point = (position,), axis 0-2, direction 0 average / 1 above
    / 2 below) or DIST_LIST (distance to the atoms of `cluster`)."""
    pt = None if point is None else np.ascontiguousarray(np.resize(np.asarray(point, np.float32), 3), np.float32)
    cell = np.zeros((traj.T, traj.N), np.int32)
    check(lib().amdat_bins_distance(ctx.h, traj.h, binned.h if binned is not None else None, kind, fp(pt), axis,
                                    direction, cluster.h if cluster is not None else None, float(thickness),
                                    int(n_bins), ip(cell)))
    return cell


def traj_bounds(traj):
    lo = np.zeros((traj.T, 3), np.float32)
    check(lib().amdat_traj_bounds(traj.h, fp(lo)))
    return lo
