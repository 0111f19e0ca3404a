// Host build of the device phase helpers (amdat_amd/csrc/phase_math.h) for
// the CPU tests: exports them with C linkage.
#include "../amdat_amd/csrc/phase_math.h"

extern "C" {
void shim_sincos_turns(const uint32_t *P, int n, float *c, float *s)
{
    for (int i = 0; i < n; i++) {
        float2 v = amdat::sincos_turns(P[i]);
        c[i] = v.x;
        s[i] = v.y;
    }
}
void shim_turns_fixed(const double *u, int n, uint32_t *out)
{
    for (int i = 0; i < n; i++) out[i] = amdat::turns_fixed(u[i]);
}
void shim_two_sum(const float *x, int n, float *hi, float *lo)
{
    float h = 0.0f, l = 0.0f;
    for (int i = 0; i < n; i++) amdat::two_sum_acc(h, l, x[i]);
    *hi = h;
    *lo = l;
}
void shim_fast_two_sum(const float *x, int n, float *hi, float *lo)
{
    float h = 0.0f, l = 0.0f;
    for (int i = 0; i < n; i++) amdat::fast_two_sum_acc(h, l, x[i]);
    *hi = h;
    *lo = l;
}
}
