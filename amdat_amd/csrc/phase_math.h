// phase_math.h -- device helpers shared by the phase kernels.
//
// Phases are carried as exact 32-bit fixed-point fractions of a turn:
// P = frac(n . r / l) * 2^32, so that P(m) = m * P(1) mod 2^32 is exact
// integer arithmetic and FP32 only enters when a table entry
// exp(2*pi*i*P/2^32) is evaluated.  That keeps the phase error at the
// FP32 rounding of sin/cos (about 1 ulp) however large |k.r| is, where a
// plain FP32 dot product k.r carries an error of |k.r| * 2^-24 radians.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define AMDAT_HD __host__ __device__
#else
// Host-only build (the CPU self-test of these helpers): plain C++.
#include <cmath>
#define AMDAT_HD
using std::floor;
using std::fmaf;
using std::rint;
struct float2 {
    float x, y;
};
inline float2 make_float2(float a, float b) { return float2{a, b}; }
#endif

namespace amdat {

// Fixed-point fraction of u turns, in units of 2^-32 turn (u in FP64).
AMDAT_HD inline uint32_t turns_fixed(double u)
{
    double f = u - floor(u);                 // [0, 1]
    double s = rint(f * 4294967296.0);       // [0, 2^32]
    return (uint32_t)(unsigned long long)s;  // 2^32 wraps to 0
}

// Polynomial coefficients fitted by amdat_amd/csrc/gen_sincos_coeffs.py
// (discrete minimax on |t| <= pi/4; max error about 1 ulp of 1.0).
constexpr float kS1 = -1.666665524e-01f;
constexpr float kS2 = 8.332161233e-03f;
constexpr float kS3 = -1.951534214e-04f;
constexpr float kC1 = -5.000000000e-01f;
constexpr float kC2 = 4.166662320e-02f;
constexpr float kC3 = -1.388676465e-03f;
constexpr float kC4 = 2.439049968e-05f;
// 2*pi/2^32 split into a float and its float remainder.
constexpr float kTurnHi = 1.462918120e-09f;
constexpr float kTurnLo = -4.070940268e-17f;

// (cos, sin) of 2*pi*P/2^32.  Quadrant from the top bits of P; the
// residual r in [-2^29, 2^29) is split into a multiple of 4096 and a
// remainder so both convert to float exactly, and the angle is formed
// with the split constant.
AMDAT_HD inline float2 sincos_turns(uint32_t P)
{
    uint32_t q = (P + 0x20000000u) >> 30;
    int32_t r = (int32_t)(P - (q << 30));
    int32_t rh = r & ~0xFFF;
    int32_t rl = r - rh;
    float fh = (float)rh, fl = (float)rl;
    float t = fmaf(fh, kTurnHi, fmaf(fh, kTurnLo, fl * kTurnHi));
    float t2 = t * t;
    float ps = fmaf(fmaf(kS3, t2, kS2), t2, kS1);
    float s = fmaf(t * t2, ps, t);
    float pc = fmaf(fmaf(fmaf(kC4, t2, kC3), t2, kC2), t2, kC1);
    float c = fmaf(t2, pc, 1.0f);
    float sn = (q & 1u) ? c : s;
    float cs = (q & 1u) ? s : c;
    if (q & 2u) sn = -sn;
    if ((q + 1u) & 2u) cs = -cs;
    return make_float2(cs, sn);
}

// Error-free transformation a + b = s + e (Knuth TwoSum); the error is
// folded into a running FP32 compensation term.
AMDAT_HD inline void two_sum_acc(float &hi, float &lo, float x)
{
    float s = hi + x;
    float bp = s - hi;
    float e = (hi - (s - bp)) + (x - bp);
    hi = s;
    lo += e;
}

// Compensated accumulation by Dekker's Fast2Sum: hi + x = s + e exactly
// when |hi| >= |x| (the usual case when tile partials are folded into a
// running sum), in 4 operations against TwoSum's 7; a step with |hi| < |x|
// (the first folds, or the sum passing through zero) keeps the error of one
// plain FP32 addition.
AMDAT_HD inline void fast_two_sum_acc(float &hi, float &lo, float x)
{
    float s = hi + x;
    float z = s - hi;
    lo += x - z;
    hi = s;
}

}  // namespace amdat
