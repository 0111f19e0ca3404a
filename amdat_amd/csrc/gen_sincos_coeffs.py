"""Fit the FP32 polynomial coefficients used by sincos_turns() in
phase_math.h.

sin(t) ~ t + t^3 * (s1 + s2 t^2 + s3 t^4)
cos(t) ~ 1 + t^2 * (c1 + c2 t^2 + c3 t^4 + c4 t^6)
on |t| <= pi/4.  The fit is a discrete minimax (iteratively reweighted
least squares on a dense grid) of the relative error of the correction
term, done in float64; the coefficients are then rounded to float32 and
the FP32 evaluation error is measured against float64 sin/cos.

Usage: python amdat_amd/csrc/gen_sincos_coeffs.py
"""
import numpy as np


def fit(f, basis, x, iters=60):
    A = np.stack([b(x) for b in basis], axis=1)
    w = np.ones_like(x)
    for _ in range(iters):
        coef, *_ = np.linalg.lstsq(A * w[:, None], f * w, rcond=None)
        err = np.abs(A @ coef - f)
        w = w * (1.0 + 4.0 * err / err.max())
        w /= w.mean()
    return coef


def main():
    t = np.linspace(1e-6, np.pi / 4, 200001)
    # sin: (sin t - t)/t^3 = s1 + s2 t^2 + s3 t^4, weight relative to sin t
    fs = (np.sin(t) - t) / t**3
    s = fit(fs * t**3 / np.sin(t), [lambda u: u**3 / np.sin(u), lambda u: u**5 / np.sin(u), lambda u: u**7 / np.sin(u)], t)
    fc = (np.cos(t) - 1.0)
    c = fit(fc, [lambda u: u**2, lambda u: u**4, lambda u: u**6, lambda u: u**8], t)
    s32 = s.astype(np.float32)
    c32 = c.astype(np.float32)
    tt = np.linspace(-np.pi / 4, np.pi / 4, 2000001).astype(np.float32)
    t2 = tt * tt
    ps = np.float32(s32[2])
    ps = ps * t2 + s32[1]
    ps = ps * t2 + s32[0]
    sv = tt + (tt * t2) * ps
    pc = np.float32(c32[3])
    pc = pc * t2 + c32[2]
    pc = pc * t2 + c32[1]
    pc = pc * t2 + c32[0]
    cv = np.float32(1.0) + t2 * pc
    t64 = tt.astype(np.float64)
    es = np.abs(sv.astype(np.float64) - np.sin(t64)).max()
    ec = np.abs(cv.astype(np.float64) - np.cos(t64)).max()
    for i, v in enumerate(s32):
        print("S%d = %.9e" % (i + 1, v))
    for i, v in enumerate(c32):
        print("C%d = %.9e" % (i + 1, v))
    print("max abs err sin %.3e cos %.3e (2^-24 = %.3e)" % (es, ec, 2.0**-24))


if __name__ == "__main__":
    main()
