"""CPU pins of the device phase helpers (amdat_amd/csrc/phase_math.h),
built for the host from the same header: sin/cos from exact fixed-point
turns, the FP64 -> fixed-point fraction, and the TwoSum accumulator."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("shim") / "libphase_shim.so")
    src = os.path.join(ROOT, "tests", "phase_math_shim.cpp")
    # FMA contraction off, as in the reference build; no -ffast-math
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-o", out, src], check=True)
    L = C.CDLL(out)
    P = C.POINTER
    L.shim_sincos_turns.argtypes = [P(C.c_uint32), C.c_int, P(C.c_float), P(C.c_float)]
    L.shim_turns_fixed.argtypes = [P(C.c_double), C.c_int, P(C.c_uint32)]
    L.shim_two_sum.argtypes = [P(C.c_float), C.c_int, P(C.c_float), P(C.c_float)]
    L.shim_fast_two_sum.argtypes = [P(C.c_float), C.c_int, P(C.c_float), P(C.c_float)]
    return L


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def test_sincos_turns_accuracy(shim):
    rng = np.random.default_rng(0)
    P = np.concatenate([rng.integers(0, 2 ** 32, 2_000_000, dtype=np.uint64),
                        np.arange(0, 2 ** 32, 2 ** 20, dtype=np.uint64),         # grid incl. quadrant edges
                        np.array([0, 1, 2 ** 29, 2 ** 29 - 1, 2 ** 30, 2 ** 31, 3 * 2 ** 30, 2 ** 32 - 1], np.uint64)]
                       ).astype(np.uint32)
    c = np.zeros(len(P), np.float32)
    s = np.zeros(len(P), np.float32)
    shim.shim_sincos_turns(_ptr(P, C.c_uint32), len(P), _ptr(c, C.c_float), _ptr(s, C.c_float))
    th = 2.0 * np.pi * P.astype(np.float64) / 2.0 ** 32
    ec = np.abs(c - np.cos(th)).max()
    es = np.abs(s - np.sin(th)).max()
    # about one ulp of 1.0 (2^-24 = 5.96e-8); a wrong quadrant or sign is O(1)
    assert ec < 1.2e-7 and es < 1.2e-7, (ec, es)
    assert c[len(P) - 8] == 1.0 and s[len(P) - 8] == 0.0


def test_turns_fixed_exact(shim):
    rng = np.random.default_rng(1)
    u = np.concatenate([rng.uniform(-50, 50, 100000), [0.0, 0.5, -0.25, 1.0 - 2.0 ** -40, 3.0]])
    out = np.zeros(len(u), np.uint32)
    shim.shim_turns_fixed(_ptr(u, C.c_double), len(u), _ptr(out, C.c_uint32))
    f = u - np.floor(u)
    exp = (np.rint(f * 2.0 ** 32).astype(np.uint64) % 2 ** 32).astype(np.uint32)
    assert np.array_equal(out, exp)
    # m * P(u) mod 2^32 equals P(m u) up to the rounding of P(u) times m
    m = 50
    pm = (out.astype(np.uint64) * m % 2 ** 32).astype(np.int64)
    um = np.zeros(len(u), np.uint32)
    um_in = np.ascontiguousarray(u * m)
    shim.shim_turns_fixed(_ptr(um_in, C.c_double), len(u), _ptr(um, C.c_uint32))
    d = (pm - um.astype(np.int64) + 2 ** 31) % 2 ** 32 - 2 ** 31
    assert np.abs(d).max() <= m


def test_two_sum_compensation(shim):
    rng = np.random.default_rng(2)
    x = rng.normal(0, 1, 200000).astype(np.float32)
    x[::1000] = 1e6
    hi = C.c_float()
    lo = C.c_float()
    shim.shim_two_sum(_ptr(x, C.c_float), len(x), C.byref(hi), C.byref(lo))
    exact = np.sum(x.astype(np.float64))
    naive = float(np.cumsum(x, dtype=np.float32)[-1])
    comp = float(hi.value) + float(lo.value)
    assert abs(comp - exact) < 1e-6 * abs(exact) < abs(naive - exact)


@pytest.mark.parametrize("case", ["random", "coherent", "through_zero"])
def test_fast_two_sum_folds(shim, case):
    """Folding tile partials (sums of 32 unit phases) into (hi, lo) with
    Fast2Sum: within 1e-7 of the exact FP64 sum on the scale
    max(|sum|, sqrt(n) * spread) of the folds, and far closer than plain
    FP32 accumulation, for random-sign partials, same-sign (coherent)
    partials, and a running sum that crosses zero many times."""
    rng = np.random.default_rng(7)
    n = 4000
    if case == "random":
        x = rng.normal(0.0, 5.7, n)
    elif case == "coherent":
        x = 32.0 - rng.uniform(0.0, 1e-3, n)
    else:
        x = np.sin(np.arange(n) * 0.05) * 30.0 + rng.normal(0.0, 1.0, n)
    x = x.astype(np.float32)
    hi, lo = C.c_float(), C.c_float()
    shim.shim_fast_two_sum(_ptr(x, C.c_float), n, C.byref(hi), C.byref(lo))
    exact = float(np.sum(x.astype(np.float64)))
    comp = float(hi.value) + float(lo.value)
    naive = float(np.cumsum(x, dtype=np.float32)[-1])
    scale = max(abs(exact), np.sqrt(n) * float(np.std(x)))
    assert abs(comp - exact) < 1e-7 * scale
    assert abs(comp - exact) <= abs(naive - exact)
