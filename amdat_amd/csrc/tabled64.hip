// tabled64.hip -- tabled-FP64 form of the phase kernels (AMDAT_PREC_FP64T):
// the compensated-FP32 kernels' plan (chunks, slot groups, LDS tables of
// separable phase factors, hot loop of complex products), with every table
// entry, product and sum in FP64.  It answers how much of the FP32 form's
// speed comes from FP32 arithmetic rather than from the tables.
//
//   * u_c is a 64-bit fixed-point fraction of r_c / l (of the FP64 rebuilt
//     displacement for Fs), so m u_c mod 1 stays exact integer arithmetic;
//   * anchors e(d), d < 8, and e(8j) come from the FP64 sincospi of the
//     exact turn; entries e(8j + d) = e(8j) e(d) are one FP64 complex
//     product;
//   * tables hold 16-byte entries, so a tile holds half the atoms of the
//     FP32 form (or fewer), within 76 KiB for two 512-thread blocks per CU;
//   * sums are plain FP64 (no compensation needed at these sizes).
#include <algorithm>

#include "kernels.h"
#include "phase_device.h"

namespace amdat {

namespace {

// Atoms per tile: tables of at most 76 KiB and the reads of all but the
// last three atoms of a tile within the 64 KiB reach of an LDS immediate
// offset.
template <int MM>
struct Geom64 {
    using G = Geom<MM>;
    static constexpr int W = G::W, OX = G::OX, OY = G::OY, OZ = G::OZ;
    static constexpr int A = (16 * W * 16 <= 77824 && 13 * W * 16 < 65536) ? 16 : ((8 * W * 16 <= 77824) ? 8 : 4);
    static constexpr int LOG2A = A == 16 ? 4 : (A == 8 ? 3 : 2);
    static constexpr size_t tab_bytes = (size_t)A * W * sizeof(double2);
};

// Fraction of a turn as a 64-bit fixed-point number: frac(u) 2^64.
__device__ __forceinline__ uint64_t turns_fixed64(double u)
{
    const double f = u - floor(u);  // [0, 1]
    return f >= 1.0 ? 0ull : (uint64_t)(f * 18446744073709551616.0);
}

// exp(2 pi i P / 2^64) in FP64: the signed turn P / 2^64 in [-1/2, 1/2).
__device__ __forceinline__ double2 sincos_turns64(uint64_t P)
{
    const double t = (double)(int64_t)P * 0x1p-63;  // 2 x turns, in [-1, 1)
    double s, c;
    sincospi(t, &s, &c);
    return make_double2(c, s);
}

__device__ __forceinline__ double2 cmul64(double2 a, double2 b)
{
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}

struct CoordFrac64 {
    const float *X, *Y, *Z;
    const int *list;
    size_t foff;
    double inv_ell;
    __device__ __forceinline__ uint64_t operator()(int a, int c) const
    {
        const size_t o = foff + (list ? list[a] : a);
        const float v = c == 0 ? X[o] : (c == 1 ? Y[o] : Z[o]);
        return turns_fixed64((double)v * inv_ell);
    }
};

struct DispFrac64 {
    const float *X, *Y, *Z;
    const int16_t *IX, *IY, *IZ;
    const float *L0, *L1;
    const int *list;
    size_t o0, o1;
    double inv_ell;
    __device__ __forceinline__ uint64_t operator()(int a, int c) const
    {
        const int j = list ? list[a] : a;
        const size_t p0 = o0 + j, p1 = o1 + j;
        float x0, x1, l0, l1;
        int i0, i1;
        if (c == 0) {
            x0 = X[p0], x1 = X[p1], i0 = IX[p0], i1 = IX[p1], l0 = L0[0], l1 = L1[0];
        } else if (c == 1) {
            x0 = Y[p0], x1 = Y[p1], i0 = IY[p0], i1 = IY[p1], l0 = L0[1], l1 = L1[1];
        } else {
            x0 = Z[p0], x1 = Z[p1], i0 = IZ[p0], i1 = IZ[p1], l0 = L0[2], l1 = L1[2];
        }
        const double d = ((double)x1 - (double)x0) + ((double)i1 * (double)l1 - (double)i0 * (double)l0);
        return turns_fixed64(d * inv_ell);
    }
};

// One tile of tables, as build_tables of the FP32 form: anchors e(d) into
// anc[d][3A], then per (axis, j, atom) the anchor e(8j) and the eight
// entries e(8j + d) at +m and, on x and y, their conjugates at -m.
template <int MM>
__device__ __forceinline__ void build_tables64(double2 *tab, double2 *anc, const uint64_t *sU, const int *sValid,
                                               int4 rng)
{
    using Gm = Geom64<MM>;
    constexpr int NR = 3 * Gm::A;
    const int tid = threadIdx.x;
    const int xlo = rng.x, xhi = rng.y, my = range_my(rng), mz = range_mz(rng);
    for (int e = tid; e < NR * 8; e += kPhaseBlock) {
        const int d = e / NR, row = e - d * NR;
        anc[e] = sincos_turns64(sU[row] * (uint64_t)d);
    }
    __syncthreads();
    const int xa = xlo >= 0 ? xlo : (xhi <= 0 ? -xhi : 0);
    const int xb = max(abs(xlo), abs(xhi));
    const int jx0 = xa >> 3, njx = (xb >> 3) - jx0 + 1;
    const int njy = (my >> 3) + 1, njz = (mz >> 3) + 1;
    const int nj = njx + njy + njz;
    for (int it = tid; it < nj << Gm::LOG2A; it += kPhaseBlock) {
        const int i = it & (Gm::A - 1), k = it >> Gm::LOG2A;
        const int c = k < njx ? 0 : (k < njx + njy ? 1 : 2);
        const int j = c == 0 ? jx0 + k : (c == 1 ? k - njx : k - njx - njy);
        const int off = c == 0 ? Gm::OX : (c == 1 ? Gm::OY : Gm::OZ);
        const int row = 3 * i + c;
        double2 a8 = sincos_turns64(sU[row] * (uint64_t)(8 * j));
        if (!sValid[i]) a8 = make_double2(0.0, 0.0);
        double2 v[8];
#pragma unroll
        for (int d = 0; d < 8; ++d) v[d] = cmul64(a8, anc[d * NR + row]);
        double2 *up = tab + i * Gm::W + off + 8 * j;
        if (c != 2) {
            double2 *dn = tab + i * Gm::W + off - 8 * j - 7;
#pragma unroll
            for (int d = 0; d < 8; ++d) dn[7 - d] = make_double2(v[d].x, -v[d].y);
        }
#pragma unroll
        for (int d = 0; d < 8; ++d) up[d] = v[d];
    }
}

// Updates in flight in the FP64 hot loop: its 16-byte entries and FP64
// sums double the registers of the FP32 form's pipeline, which spills at
// R = 8 with 4 in flight under the 128-register limit of two 512-thread
// blocks per CU.
constexpr int kDepth64 = 2;

// One tile of the hot loop in FP64, as hot_tile of the FP32 form (same
// software pipeline, kDepth64 updates in flight, same slot groups).
template <int R, int MM, bool FS, int GS>
__device__ __forceinline__ void hot_tile64(const double2 *px, const double2 *const (&py)[R],
                                           const double2 *const (&pz)[R], double (&p0)[R], double (&p1)[R])
{
    using Gm = Geom64<MM>;
    constexpr int kDepth = kDepth64;
    static_assert(R % GS == 0, "slots per thread must hold whole groups");
    double2 by[kDepth], bz[kDepth];
#pragma unroll
    for (int u = 0; u < kDepth; ++u) {
        if (u % GS == 0) by[u] = py[u % R][(u / R) * Gm::W];
        else by[u] = make_double2(0.0, 0.0);
        bz[u] = pz[u % R][(u / R) * Gm::W];
    }
    double2 ex = px[0], ex_next = ex;
    double qr = 0.0, qi = 0.0;
#pragma clang loop unroll(full)
    for (int u = 0; u < Gm::A * R; ++u) {
        const int a = u / R, r = u % R, slot = u % kDepth;
        const bool lead = r % GS == 0;
        const double2 ey = by[slot], ez = bz[slot];
        if (u + kDepth < Gm::A * R) {
            const int v = u + kDepth;
            if (v % GS == 0) by[slot] = py[v % R][(v / R) * Gm::W];
            bz[slot] = pz[v % R][(v / R) * Gm::W];
        }
        if (r == 0 && a + 1 < Gm::A) ex_next = px[(a + 1) * Gm::W];
        __builtin_amdgcn_sched_barrier(0);
        if (lead) {
            qr = fma(ex.x, ey.x, -ex.y * ey.y);
            qi = fma(ex.x, ey.y, ex.y * ey.x);
        }
        p0[r] = fma(qr, ez.x, fma(-qi, ez.y, p0[r]));
        if constexpr (!FS) p1[r] = fma(qr, ez.y, fma(qi, ez.x, p1[r]));
        __builtin_amdgcn_sched_barrier(0);
        if (r == R - 1) ex = ex_next;
    }
}

// The tile loop over one block's atoms (FP64 sums s0 = Re or Fs, s1 = Im).
template <int R, int MM, bool FS, bool GROUPED, typename UF>
__device__ __forceinline__ void phase_tiles64(const uint32_t *__restrict__ kp, int4 rng, int a_begin, int a_end,
                                              double2 *tab, double2 *anc, uint64_t *sU, int *sValid, const UF uf,
                                              double (&s0)[R], double (&s1)[R])
{
    using Gm = Geom64<MM>;
    const int tid = threadIdx.x;
    const int wave0 = __builtin_amdgcn_readfirstlane(tid >> 6) * 64;
    const bool active = wave0 < range_units(rng);
    const int gs_code = GROUPED ? range_gs_code(rng, wave0 >> 6) : 0;
    const double2 *px = nullptr, *py[R], *pz[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int nx, ny, nz;
        const uint32_t k = kp[r * kPhaseBlock + tid];
        unpack_k(k, nx, ny, nz);
        if (r == 0) px = tab + Gm::OX + nx;
        const bool sw = (k >> 24) & 1u;
        py[r] = sw ? tab + Gm::OZ + nz : tab + Gm::OY + ny;
        pz[r] = sw ? tab + Gm::OY + ny : tab + Gm::OZ + nz;
        s0[r] = s1[r] = 0.0;
    }
    for (int a0 = a_begin; a0 < a_end; a0 += Gm::A) {
        __syncthreads();
        if (tid < 3 * Gm::A) {
            const int i = tid / 3, c = tid - 3 * (tid / 3);
            const int a = a0 + i;
            sU[tid] = a < a_end ? uf(a, c) : 0ull;
            if (c == 0) sValid[i] = a < a_end;
        }
        __syncthreads();
        build_tables64<MM>(tab, anc, sU, sValid, rng);
        __syncthreads();
        if (!active) continue;
        if constexpr (GROUPED) {
            if constexpr (R >= 8) {
                if (gs_code == 3) {
                    hot_tile64<R, MM, FS, 8>(px, py, pz, s0, s1);
                    continue;
                }
            }
            if constexpr (R >= 4) {
                if (gs_code == 2) {
                    hot_tile64<R, MM, FS, 4>(px, py, pz, s0, s1);
                    continue;
                }
            }
            if (gs_code == 1) hot_tile64<R, MM, FS, 2>(px, py, pz, s0, s1);
            else hot_tile64<R, MM, FS, 1>(px, py, pz, s0, s1);
        } else {
            hot_tile64<R, MM, FS, 1>(px, py, pz, s0, s1);
        }
    }
}

#define AMDAT_PHASE_SHARED64(MM)                      \
    extern __shared__ double2 tab64[];                \
    __shared__ double2 anc64[3 * Geom64<MM>::A * 8];  \
    __shared__ uint64_t sU64[3 * Geom64<MM>::A];      \
    __shared__ int sValid64[Geom64<MM>::A]

// Waves per SIMD the register budget is set for: R = 8 grouped rho needs
// more than the 128 registers of 4 waves (its complex FP64 sums and group
// factors), so it runs one 512-thread block per CU.
template <int R, bool GROUPED>
constexpr int rho_t64_waves() { return (R == 8 && GROUPED) ? 2 : 4; }

// rho_k(t), tabled FP64: work items as rho_c32_kernel.
template <int R, int MM, bool GROUPED>
__global__ __launch_bounds__(kPhaseBlock, (rho_t64_waves<R, GROUPED>())) void rho_t64_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z, int N,
    const int *__restrict__ list, int n_list, const int *__restrict__ frames, int n_frames, int n_chunks,
    const uint32_t *__restrict__ kpack, const int *__restrict__ kdst, const int4 *__restrict__ range,
    int K, int KB, double inv_ell, int per_split, int n_splits, double2 *__restrict__ out)
{
    AMDAT_PHASE_SHARED64(MM);
    int chunk, row;
    if (!xcd_item(n_chunks, n_frames * n_splits, chunk, row)) return;
    const int fslot = row % n_frames, split = row / n_frames;
    const int a_begin = split * per_split;
    const int a_end = min(n_list, a_begin + per_split);
    const CoordFrac64 uf{X, Y, Z, list, (size_t)frames[fslot] * (size_t)N, inv_ell};
    double re[R], im[R];
    phase_tiles64<R, MM, false, GROUPED>(kpack + (size_t)chunk * KB, range[chunk], a_begin, a_end, tab64, anc64,
                                         sU64, sValid64, uf, re, im);
    const int *kd = kdst + (size_t)chunk * KB;
    double2 *dst = out + ((size_t)fslot * n_splits + split) * (size_t)K;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int d = kd[r * kPhaseBlock + threadIdx.x];
        if (d >= 0) dst[d & ~(1 << 30)] = make_double2(re[r], (d & (1 << 30)) ? -im[r] : im[r]);
    }
}

// Fs per (pair, vector), tabled FP64: work items as fs_c32_kernel.
template <int R, int MM, bool GROUPED>
__global__ __launch_bounds__(kPhaseBlock, 4) void fs_t64_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z,
    const int16_t *__restrict__ IX, const int16_t *__restrict__ IY, const int16_t *__restrict__ IZ,
    const float *__restrict__ box, int N, const int *__restrict__ list, int n_list, const int2 *__restrict__ seg,
    const int *__restrict__ T0, const int *__restrict__ T1, int n_pairs, int n_chunks,
    const uint32_t *__restrict__ kpack, const int *__restrict__ kdst, const int4 *__restrict__ range, int K, int KB,
    double inv_ell, int per_split, int n_splits, double *__restrict__ out)
{
    AMDAT_PHASE_SHARED64(MM);
    int chunk, row;
    if (!xcd_item(n_chunks, n_pairs * n_splits, chunk, row)) return;
    const int pslot = row % n_pairs, split = row / n_pairs;
    const int t0 = T0[pslot], t1 = T1[pslot];
    if (seg) {  // the pair's own atoms: list[seg.x .. seg.x + seg.y)
        const int2 sg = seg[pslot];
        list += sg.x;
        n_list = sg.y;
    }
    const int a_begin = split * per_split;
    const int a_end = min(n_list, a_begin + per_split);
    const DispFrac64 uf{X, Y, Z, IX, IY, IZ, box + 3 * t0, box + 3 * t1, list, (size_t)t0 * N, (size_t)t1 * N,
                        inv_ell};
    double s[R], unused[R];
    phase_tiles64<R, MM, true, GROUPED>(kpack + (size_t)chunk * KB, range[chunk], a_begin, a_end, tab64, anc64,
                                        sU64, sValid64, uf, s, unused);
    const int *kd = kdst + (size_t)chunk * KB;
    double *dst = out + ((size_t)pslot * n_splits + split) * (size_t)K;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int d = kd[r * kPhaseBlock + threadIdx.x];
        if (d >= 0) dst[d & ~(1 << 30)] = s[r];
    }
}

template <int R, int MM, bool GROUPED>
hipError_t launch_rho_t64_t(const TrajDev &tr, const int *list, int n_list, const int *frames, int n_frames,
                            const KPlan &kp, double inv_ell, int n_splits, double2 *out, hipStream_t st)
{
    const int per_split = (n_list + n_splits - 1) / n_splits;
    const size_t lds = Geom64<MM>::tab_bytes;
    auto kfn = rho_t64_kernel<R, MM, GROUPED>;
    hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const long rows = (long)n_frames * n_splits;
    dim3 grid((unsigned)(kXcds * kp.n_chunks * ((rows + kXcds - 1) / kXcds)));
    (void)hipGetLastError();
    hipLaunchKernelGGL(kfn, grid, dim3(kPhaseBlock), lds, st, tr.x, tr.y, tr.z, tr.N, list, n_list, frames,
                       n_frames, kp.n_chunks, kp.kpack, kp.kdst, kp.range, kp.K, kp.KB, inv_ell, per_split, n_splits,
                       out);
    return hipGetLastError();
}

template <int R, int MM, bool GROUPED>
hipError_t launch_fs_t64_t(const TrajDev &tr, const int *list, int n_list, const int2 *seg, const int *t0,
                           const int *t1, int n_pairs, const KPlan &kp, double inv_ell, int n_splits, double *out,
                           hipStream_t st)
{
    const int per_split = (n_list + n_splits - 1) / n_splits;
    const size_t lds = Geom64<MM>::tab_bytes;
    auto kfn = fs_t64_kernel<R, MM, GROUPED>;
    hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const long rows = (long)n_pairs * n_splits;
    dim3 grid((unsigned)(kXcds * kp.n_chunks * ((rows + kXcds - 1) / kXcds)));
    (void)hipGetLastError();
    hipLaunchKernelGGL(kfn, grid, dim3(kPhaseBlock), lds, st, tr.x, tr.y, tr.z, tr.ix, tr.iy, tr.iz, tr.box,
                       tr.N, list, n_list, seg, t0, t1, n_pairs, kp.n_chunks, kp.kpack, kp.kdst, kp.range, kp.K,
                       kp.KB, inv_ell, per_split, n_splits, out);
    return hipGetLastError();
}

template <int R, bool GROUPED>
hipError_t rho_t64_mm(const TrajDev &tr, const int *list, int n_list, const int *frames, int n_frames,
                      const KPlan &kp, double inv_ell, int n_splits, double2 *out, hipStream_t st)
{
    switch (kp.MM) {
    case 24: return launch_rho_t64_t<R, 24, GROUPED>(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    case 50: return launch_rho_t64_t<R, 50, GROUPED>(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    case 100: return launch_rho_t64_t<R, 100, GROUPED>(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    case 127: return launch_rho_t64_t<R, 127, GROUPED>(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    default: return hipErrorInvalidValue;
    }
}

template <int R, bool GROUPED>
hipError_t fs_t64_mm(const TrajDev &tr, const int *list, int n_list, const int2 *seg, const int *t0, const int *t1,
                     int n_pairs, const KPlan &kp, double inv_ell, int n_splits, double *out, hipStream_t st)
{
    switch (kp.MM) {
    case 24: return launch_fs_t64_t<R, 24, GROUPED>(tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    case 50: return launch_fs_t64_t<R, 50, GROUPED>(tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    case 100: return launch_fs_t64_t<R, 100, GROUPED>(tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    case 127: return launch_fs_t64_t<R, 127, GROUPED>(tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_rho_t64(const TrajDev &tr, const int *list, int n_list, const int *frames, int n_frames,
                          const KPlan &kp, double inv_ell, int n_splits, double2 *out, hipStream_t st)
{
    switch (kp.R) {
    case 2:
        if (kp.pair) return hipErrorInvalidValue;
        return rho_t64_mm<2, false>(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    case 4:
        if (kp.pair) return rho_t64_mm<4, true>(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
        return rho_t64_mm<4, false>(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    case 8:
        if (kp.pair) return rho_t64_mm<8, true>(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
        return rho_t64_mm<8, false>(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fs_t64(const TrajDev &tr, const int *list, int n_list, const int2 *seg, const int *t0,
                         const int *t1, int n_pairs, const KPlan &kp, double inv_ell, int n_splits, double *out,
                         hipStream_t st)
{
    switch (kp.R) {
    case 2:
        if (kp.pair) return hipErrorInvalidValue;
        return fs_t64_mm<2, false>(tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    case 4:
        if (kp.pair) return fs_t64_mm<4, true>(tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
        return fs_t64_mm<4, false>(tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    case 8:
        if (kp.pair) return fs_t64_mm<8, true>(tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
        return fs_t64_mm<8, false>(tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace amdat
