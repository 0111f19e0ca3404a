// kernels.hip -- gfx950 kernels of the reciprocal-space and time-correlation
// path: rho_k(t), Fs(k,t), MSD/NGP moments, and the FP64 reductions that turn
// per-vector sums into per-shell S(k), F(k,t) and Fs(k,t).
//
// Phase kernels (rho, Fs), compensated-FP32 form:
//   * a block of 512 threads owns one chunk of KB = 512*R k-vectors (R per
//     thread, kept in registers as LDS addresses) and one frame (rho) or
//     time-origin pair (Fs), and streams the atoms of its split through LDS
//     in tiles of A atoms;
//   * per tile, every atom gets separable phase tables
//         e_c(m) = exp(2*pi*i * m * u_c),  u_c = frac(r_c / l),
//     for m in [-M, M] on x and y and [0, M] on z (vectors are stored with
//     nz >= 0; a vector with nz < 0 is flipped and its result conjugated).
//     u_c is an exact 32-bit fixed-point fraction formed in FP64 from the
//     FP32 coordinate (for Fs from the FP64 displacement rebuilt from image
//     counts), so m*u_c mod 1 is exact integer arithmetic.  Entries are
//     e(8j + d) = e(8j) * e(d) with both factors from sincos_turns(), so
//     every entry carries one FP32 complex product over exact anchors;
//   * the hot loop is per (atom, vector): three LDS reads, all with a
//     per-lane base and a compile-time offset (the tile loop is fully
//     unrolled), and two complex products,
//         exp(i k.r) = e_x(nx) e_y(ny) e_z(nz),
//     accumulated with FMA into an FP32 partial per tile; tile partials are
//     folded into a TwoSum-compensated FP32 (hi, lo) pair; the block writes
//     hi + lo as FP64.
// The plain-FP64 form evaluates exp(i k.r) with FP64 sincos per update and
// accumulates in FP64; it is the reference the FP32 form is measured against.
#include <algorithm>

#include "kernels.h"
#include "phase_math.h"

namespace amdat {

namespace {

constexpr int kWaves = kPhaseBlock / 64;
// Prefetch depth (updates in flight) of the hot loops.
constexpr int kDepth = 4;

// Table geometry for half-width MM: row of W float2 per atom,
// x at [0, 2MM], y at [2MM+1, 4MM+1], z (m >= 0 only) at [4MM+2, 5MM+2].
// A atoms per tile: (A-1)*W*8 stays below 64 KiB so every read is one
// ds_read_b64 with an immediate offset.  NA anchor slots per table row:
// e(0..7) and e(8j), j = 0..MM/8.
template <int MM>
struct Geom {
    // at least 257 float2 per row: reads of neighbouring atoms then never
    // pair into one ds_read2 (8-bit offsets), which keeps register use low
    static constexpr int W = (5 * MM + 3) > 257 ? (5 * MM + 3) : 257;
    static constexpr int OX = MM;
    static constexpr int OY = 2 * MM + 1 + MM;
    static constexpr int OZ = 2 * (2 * MM + 1);
    static constexpr int A = (15 * W * 8 < 65536) ? 16 : 8;
    static constexpr int NA = 8 + MM / 8 + 1;
    static constexpr size_t tab_bytes = (size_t)A * W * sizeof(float2);
};

__device__ __forceinline__ void unpack_k(uint32_t kp, int &nx, int &ny, int &nz)
{
    nx = (int)(kp & 0xFFu) - 128;
    ny = (int)((kp >> 8) & 0xFFu) - 128;
    nz = (int)((kp >> 16) & 0xFFu);
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

// Build one tile of tables from the fixed-point fractions sU[3A] (row
// 3i+c = atom i, axis c; invalid atoms carry sValid[i] == 0 and get zero
// rows).  Two phases: exact anchors into anc[3A][NA], then products.
template <int MM>
__device__ __forceinline__ void build_tables(float2 *tab, float2 *anc, const uint32_t *sU, const int *sValid, int Mc)
{
    using Gm = Geom<MM>;
    const int tid = threadIdx.x;
    const int nj = (Mc >> 3) + 1;           // anchors e(8j), j < nj
    const int nslot = 8 + nj;
    for (int e = tid; e < 3 * Gm::A * Gm::NA; e += kPhaseBlock) {
        const int row = e / Gm::NA;
        const int s = e - row * Gm::NA;
        if (s < nslot) {
            const uint32_t m = s < 8 ? (uint32_t)s : 8u * (uint32_t)(s - 8);
            anc[e] = sincos_turns(sU[row] * m);
        }
    }
    __syncthreads();
    const int lane = tid & 63;
    const int wave = tid >> 6;
    for (int row = wave; row < 3 * Gm::A; row += kWaves) {
        const int i = row / 3;
        const int c = row - 3 * i;
        const bool valid = sValid[i] != 0;
        const float2 *a = anc + row * Gm::NA;
        float2 *dst = tab + i * Gm::W;
        for (int m = lane; m <= Mc; m += 64) {
            float2 v = valid ? cmul(a[8 + (m >> 3)], a[m & 7]) : make_float2(0.0f, 0.0f);
            if (c == 0) {
                dst[Gm::OX + m] = v;
                if (m) dst[Gm::OX - m] = make_float2(v.x, -v.y);
            } else if (c == 1) {
                dst[Gm::OY + m] = v;
                if (m) dst[Gm::OY - m] = make_float2(v.x, -v.y);
            } else {
                dst[Gm::OZ + m] = v;
            }
        }
    }
}

// rho_k(t), compensated FP32.  grid = (chunks, frames, splits).
template <int R, int MM>
__global__ __launch_bounds__(kPhaseBlock, 4) void rho_c32_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z, int N,
    const int *__restrict__ list, int n_list, const int *__restrict__ frames,
    const uint32_t *__restrict__ kpack, const int *__restrict__ kdst, const int *__restrict__ chunkM,
    int K, int KB, double inv_ell, int per_split, int n_splits, double2 *__restrict__ out)
{
    using Gm = Geom<MM>;
    extern __shared__ float2 tab[];
    __shared__ float2 anc[3 * Gm::A * Gm::NA];
    __shared__ uint32_t sU[3 * Gm::A];
    __shared__ int sValid[Gm::A];

    const int chunk = blockIdx.x, fslot = blockIdx.y, split = blockIdx.z;
    const int tid = threadIdx.x;
    const int Mc = chunkM[chunk];
    const size_t foff = (size_t)frames[fslot] * (size_t)N;
    const int a_begin = split * per_split;
    const int a_end = min(n_list, a_begin + per_split);

    const float2 *px[R], *py[R], *pz[R];
    float hr[R], hi[R], lr[R], li[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int nx, ny, nz;
        unpack_k(kpack[(size_t)chunk * KB + r * kPhaseBlock + tid], nx, ny, nz);
        px[r] = tab + Gm::OX + nx;
        py[r] = tab + Gm::OY + ny;
        pz[r] = tab + Gm::OZ + nz;
        hr[r] = hi[r] = lr[r] = li[r] = 0.0f;
    }

    for (int a0 = a_begin; a0 < a_end; a0 += Gm::A) {
        __syncthreads();
        if (tid < 3 * Gm::A) {
            const int i = tid / 3, c = tid - 3 * (tid / 3);
            const int a = a0 + i;
            uint32_t U = 0;
            if (a < a_end) {
                const int j = list ? list[a] : a;
                const float *src = c == 0 ? X : (c == 1 ? Y : Z);
                U = turns_fixed((double)src[foff + j] * inv_ell);
            }
            sU[tid] = U;
            if (c == 0) sValid[i] = a < a_end;
        }
        __syncthreads();
        build_tables<MM>(tab, anc, sU, sValid, Mc);
        __syncthreads();

        float pr[R], pi[R];
#pragma unroll
        for (int r = 0; r < R; ++r) pr[r] = pi[r] = 0.0f;
        // Software pipeline over the flattened (atom, vector) sequence: the
        // reads of update u + kDepth are issued before update u is computed.
        float2 bx[kDepth], by[kDepth], bz[kDepth];
#pragma unroll
        for (int u = 0; u < kDepth; ++u) {
            bx[u] = px[u % R][(u / R) * Gm::W];
            by[u] = py[u % R][(u / R) * Gm::W];
            bz[u] = pz[u % R][(u / R) * Gm::W];
        }
#pragma unroll
        for (int u = 0; u < Gm::A * R; ++u) {
            const int r = u % R, slot = u % kDepth;
            const float2 ex = bx[slot], ey = by[slot], ez = bz[slot];
            if (u + kDepth < Gm::A * R) {
                const int v = u + kDepth;
                bx[slot] = px[v % R][(v / R) * Gm::W];
                by[slot] = py[v % R][(v / R) * Gm::W];
                bz[slot] = pz[v % R][(v / R) * Gm::W];
            }
            __builtin_amdgcn_sched_barrier(0);
            const float qr = fmaf(ex.x, ey.x, -ex.y * ey.y);
            const float qi = fmaf(ex.x, ey.y, ex.y * ey.x);
            pr[r] = fmaf(qr, ez.x, fmaf(-qi, ez.y, pr[r]));
            pi[r] = fmaf(qr, ez.y, fmaf(qi, ez.x, pi[r]));
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            two_sum_acc(hr[r], lr[r], pr[r]);
            two_sum_acc(hi[r], li[r], pi[r]);
        }
    }

    double2 *dst = out + ((size_t)fslot * n_splits + split) * (size_t)K;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int d = kdst[(size_t)chunk * KB + r * kPhaseBlock + tid];
        if (d >= 0) {
            const double re = (double)hr[r] + (double)lr[r];
            double im = (double)hi[r] + (double)li[r];
            if (d & (1 << 30)) im = -im;
            dst[d & ~(1 << 30)] = make_double2(re, im);
        }
    }
}

// Fs per (pair, vector), compensated FP32: sum over the split's atoms of
// cos(k.dr).  grid = (chunks, pairs, splits).
template <int R, int MM>
__global__ __launch_bounds__(kPhaseBlock, 4) void fs_c32_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z,
    const int16_t *__restrict__ IX, const int16_t *__restrict__ IY, const int16_t *__restrict__ IZ,
    const float *__restrict__ box, int N, const int *__restrict__ list, int n_list,
    const int *__restrict__ T0, const int *__restrict__ T1, const uint32_t *__restrict__ kpack,
    const int *__restrict__ kdst, const int *__restrict__ chunkM, int K, int KB, double inv_ell,
    int per_split, int n_splits, double *__restrict__ out)
{
    using Gm = Geom<MM>;
    extern __shared__ float2 tab[];
    __shared__ float2 anc[3 * Gm::A * Gm::NA];
    __shared__ uint32_t sU[3 * Gm::A];
    __shared__ int sValid[Gm::A];

    const int chunk = blockIdx.x, pslot = blockIdx.y, split = blockIdx.z;
    const int tid = threadIdx.x;
    const int Mc = chunkM[chunk];
    const int t0 = T0[pslot], t1 = T1[pslot];
    const size_t o0 = (size_t)t0 * (size_t)N, o1 = (size_t)t1 * (size_t)N;
    const int a_begin = split * per_split;
    const int a_end = min(n_list, a_begin + per_split);

    const float2 *px[R], *py[R], *pz[R];
    float hs[R], ls[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int nx, ny, nz;
        unpack_k(kpack[(size_t)chunk * KB + r * kPhaseBlock + tid], nx, ny, nz);
        px[r] = tab + Gm::OX + nx;
        py[r] = tab + Gm::OY + ny;
        pz[r] = tab + Gm::OZ + nz;
        hs[r] = ls[r] = 0.0f;
    }

    for (int a0 = a_begin; a0 < a_end; a0 += Gm::A) {
        __syncthreads();
        if (tid < 3 * Gm::A) {
            const int i = tid / 3, c = tid - 3 * (tid / 3);
            const int a = a0 + i;
            uint32_t U = 0;
            if (a < a_end) {
                const int j = list ? list[a] : a;
                const float *src = c == 0 ? X : (c == 1 ? Y : Z);
                const int16_t *img = c == 0 ? IX : (c == 1 ? IY : IZ);
                // dr = (x1 + i1*L1) - (x0 + i0*L0), exact in FP64.
                const double d = ((double)src[o1 + j] - (double)src[o0 + j]) +
                                 ((double)img[o1 + j] * (double)box[3 * t1 + c] -
                                  (double)img[o0 + j] * (double)box[3 * t0 + c]);
                U = turns_fixed(d * inv_ell);
            }
            sU[tid] = U;
            if (c == 0) sValid[i] = a < a_end;
        }
        __syncthreads();
        build_tables<MM>(tab, anc, sU, sValid, Mc);
        __syncthreads();

        float ps[R];
#pragma unroll
        for (int r = 0; r < R; ++r) ps[r] = 0.0f;
        float2 bx[kDepth], by[kDepth], bz[kDepth];
#pragma unroll
        for (int u = 0; u < kDepth; ++u) {
            bx[u] = px[u % R][(u / R) * Gm::W];
            by[u] = py[u % R][(u / R) * Gm::W];
            bz[u] = pz[u % R][(u / R) * Gm::W];
        }
#pragma unroll
        for (int u = 0; u < Gm::A * R; ++u) {
            const int r = u % R, slot = u % kDepth;
            const float2 ex = bx[slot], ey = by[slot], ez = bz[slot];
            if (u + kDepth < Gm::A * R) {
                const int v = u + kDepth;
                bx[slot] = px[v % R][(v / R) * Gm::W];
                by[slot] = py[v % R][(v / R) * Gm::W];
                bz[slot] = pz[v % R][(v / R) * Gm::W];
            }
            __builtin_amdgcn_sched_barrier(0);
            const float qr = fmaf(ex.x, ey.x, -ex.y * ey.y);
            const float qi = fmaf(ex.x, ey.y, ex.y * ey.x);
            ps[r] = fmaf(qr, ez.x, fmaf(-qi, ez.y, ps[r]));
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) two_sum_acc(hs[r], ls[r], ps[r]);
    }

    double *dst = out + ((size_t)pslot * n_splits + split) * (size_t)K;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int d = kdst[(size_t)chunk * KB + r * kPhaseBlock + tid];
        if (d >= 0) dst[d & ~(1 << 30)] = (double)hs[r] + (double)ls[r];
    }
}

// Plain FP64 rho: FP64 phase, FP64 sincos, FP64 sums.
template <int R>
__global__ __launch_bounds__(kPhaseBlock) void rho_f64_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z, int N,
    const int *__restrict__ list, int n_list, const int *__restrict__ frames,
    const uint32_t *__restrict__ kpack, const int *__restrict__ kdst, int K, int KB, double kscale,
    int per_split, int n_splits, double2 *__restrict__ out)
{
    const int chunk = blockIdx.x, fslot = blockIdx.y, split = blockIdx.z;
    const int tid = threadIdx.x;
    const size_t foff = (size_t)frames[fslot] * (size_t)N;
    const int a_begin = split * per_split;
    const int a_end = min(n_list, a_begin + per_split);
    double kx[R], ky[R], kz[R], ar[R], ai[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int nx, ny, nz;
        unpack_k(kpack[(size_t)chunk * KB + r * kPhaseBlock + tid], nx, ny, nz);
        kx[r] = kscale * nx; ky[r] = kscale * ny; kz[r] = kscale * nz;
        ar[r] = ai[r] = 0.0;
    }
    for (int a = a_begin; a < a_end; ++a) {
        const int j = list ? list[a] : a;
        const double x = X[foff + j], y = Y[foff + j], z = Z[foff + j];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            double s, c;
            sincos(kx[r] * x + ky[r] * y + kz[r] * z, &s, &c);
            ar[r] += c;
            ai[r] += s;
        }
    }
    double2 *dst = out + ((size_t)fslot * n_splits + split) * (size_t)K;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int d = kdst[(size_t)chunk * KB + r * kPhaseBlock + tid];
        if (d >= 0) dst[d & ~(1 << 30)] = make_double2(ar[r], (d & (1 << 30)) ? -ai[r] : ai[r]);
    }
}

// Plain FP64 Fs.
template <int R>
__global__ __launch_bounds__(kPhaseBlock) void fs_f64_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z,
    const int16_t *__restrict__ IX, const int16_t *__restrict__ IY, const int16_t *__restrict__ IZ,
    const float *__restrict__ box, int N, const int *__restrict__ list, int n_list,
    const int *__restrict__ T0, const int *__restrict__ T1, const uint32_t *__restrict__ kpack,
    const int *__restrict__ kdst, int K, int KB, double kscale, int per_split, int n_splits,
    double *__restrict__ out)
{
    const int chunk = blockIdx.x, pslot = blockIdx.y, split = blockIdx.z;
    const int tid = threadIdx.x;
    const int t0 = T0[pslot], t1 = T1[pslot];
    const size_t o0 = (size_t)t0 * (size_t)N, o1 = (size_t)t1 * (size_t)N;
    const double L0x = box[3 * t0], L0y = box[3 * t0 + 1], L0z = box[3 * t0 + 2];
    const double L1x = box[3 * t1], L1y = box[3 * t1 + 1], L1z = box[3 * t1 + 2];
    const int a_begin = split * per_split;
    const int a_end = min(n_list, a_begin + per_split);
    double kx[R], ky[R], kz[R], as[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int nx, ny, nz;
        unpack_k(kpack[(size_t)chunk * KB + r * kPhaseBlock + tid], nx, ny, nz);
        kx[r] = kscale * nx; ky[r] = kscale * ny; kz[r] = kscale * nz;
        as[r] = 0.0;
    }
    for (int a = a_begin; a < a_end; ++a) {
        const int j = list ? list[a] : a;
        const double dx = ((double)X[o1 + j] - (double)X[o0 + j]) + ((double)IX[o1 + j] * L1x - (double)IX[o0 + j] * L0x);
        const double dy = ((double)Y[o1 + j] - (double)Y[o0 + j]) + ((double)IY[o1 + j] * L1y - (double)IY[o0 + j] * L0y);
        const double dz = ((double)Z[o1 + j] - (double)Z[o0 + j]) + ((double)IZ[o1 + j] * L1z - (double)IZ[o0 + j] * L0z);
#pragma unroll
        for (int r = 0; r < R; ++r) as[r] += cos(kx[r] * dx + ky[r] * dy + kz[r] * dz);
    }
    double *dst = out + ((size_t)pslot * n_splits + split) * (size_t)K;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int d = kdst[(size_t)chunk * KB + r * kPhaseBlock + tid];
        if (d >= 0) dst[d & ~(1 << 30)] = as[r];
    }
}

__device__ __forceinline__ double block_sum(double v, double *red)
{
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kBlock) void split_reduce2_kernel(const double2 *__restrict__ in, int n_rows,
                                                                int n_splits, int K, double2 *__restrict__ out)
{
    const size_t n = (size_t)n_rows * K;
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += (size_t)gridDim.x * kBlock) {
        const size_t f = e / K, k = e - f * K;
        double re = 0.0, im = 0.0;
        for (int s = 0; s < n_splits; ++s) {
            const double2 v = in[(f * n_splits + s) * K + k];
            re += v.x;
            im += v.y;
        }
        out[e] = make_double2(re, im);
    }
}

// grid = (shells, groups).
__global__ __launch_bounds__(kBlock) void corr_shell_kernel(
    const double2 *__restrict__ rho1, const double2 *__restrict__ rho2, int K,
    const int *__restrict__ group_off, const int *__restrict__ row0, const int *__restrict__ row1,
    const int *__restrict__ shell_off, int n_shells, const double *__restrict__ norm,
    double *__restrict__ out)
{
    __shared__ double red[kBlock];
    const int s = blockIdx.x, g = blockIdx.y;
    const int v0 = shell_off[s], H = shell_off[s + 1] - v0;
    const int p0 = group_off[g], np = group_off[g + 1] - p0;
    const long items = (long)np * H;
    double acc = 0.0;
    for (long it = threadIdx.x; it < items; it += kBlock) {
        const int p = p0 + (int)(it / H);
        const int v = v0 + (int)(it % H);
        const double2 a = rho1[(size_t)row0[p] * K + v];
        const double2 b = rho2[(size_t)row1[p] * K + v];
        acc += a.x * b.x + a.y * b.y;
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) {
        const double nrm = norm[g];
        out[(size_t)g * n_shells + s] = (H > 0 && nrm != 0.0) ? acc / (double)H / nrm : 0.0;
    }
}

__global__ __launch_bounds__(kBlock) void fs_shell_kernel(
    const double *__restrict__ V, int n_splits, int K, const int *__restrict__ group_off,
    const int *__restrict__ order, const int *__restrict__ shell_off, int n_shells,
    const double *__restrict__ norm, double *__restrict__ out)
{
    __shared__ double red[kBlock];
    const int s = blockIdx.x, g = blockIdx.y;
    const int v0 = shell_off[s], H = shell_off[s + 1] - v0;
    const int p0 = group_off[g], np = group_off[g + 1] - p0;
    const long per_pair = (long)n_splits * H;
    const long items = (long)np * per_pair;
    double acc = 0.0;
    for (long it = threadIdx.x; it < items; it += kBlock) {
        const int pl = (int)(it / per_pair);
        const long rem = it - (long)pl * per_pair;
        const int sp = (int)(rem / H);
        const int v = v0 + (int)(rem - (long)sp * H);
        acc += V[((size_t)order[p0 + pl] * n_splits + sp) * K + v];
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) {
        const double nrm = norm[g];
        out[(size_t)g * n_shells + s] = (H > 0 && nrm != 0.0) ? acc / (double)H / nrm : 0.0;
    }
}

// grid = (atom blocks, pairs).
__global__ __launch_bounds__(kBlock) void msd_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z,
    const int16_t *__restrict__ IX, const int16_t *__restrict__ IY, const int16_t *__restrict__ IZ,
    const float *__restrict__ box, int N, const int *__restrict__ list, int n_list,
    const int *__restrict__ T0, const int *__restrict__ T1, double *__restrict__ partial)
{
    __shared__ double red[kBlock];
    const int p = blockIdx.y;
    const int t0 = T0[p], t1 = T1[p];
    const size_t o0 = (size_t)t0 * (size_t)N, o1 = (size_t)t1 * (size_t)N;
    const double L0x = box[3 * t0], L0y = box[3 * t0 + 1], L0z = box[3 * t0 + 2];
    const double L1x = box[3 * t1], L1y = box[3 * t1 + 1], L1z = box[3 * t1 + 2];
    double a2 = 0.0, a4 = 0.0;
    for (int a = blockIdx.x * kBlock + threadIdx.x; a < n_list; a += gridDim.x * kBlock) {
        const int j = list ? list[a] : a;
        const double dx = ((double)X[o1 + j] - (double)X[o0 + j]) + ((double)IX[o1 + j] * L1x - (double)IX[o0 + j] * L0x);
        const double dy = ((double)Y[o1 + j] - (double)Y[o0 + j]) + ((double)IY[o1 + j] * L1y - (double)IY[o0 + j] * L0y);
        const double dz = ((double)Z[o1 + j] - (double)Z[o0 + j]) + ((double)IZ[o1 + j] * L1z - (double)IZ[o0 + j] * L0z);
        const double r2 = dx * dx + dy * dy + dz * dz;
        a2 += r2;
        a4 += r2 * r2;
    }
    a2 = block_sum(a2, red);
    a4 = block_sum(a4, red);
    if (threadIdx.x == 0) {
        partial[((size_t)p * gridDim.x + blockIdx.x) * 2] = a2;
        partial[((size_t)p * gridDim.x + blockIdx.x) * 2 + 1] = a4;
    }
}

__global__ __launch_bounds__(kBlock) void msd_reduce_kernel(const double *__restrict__ partial, int n_blocks,
                                                             const int *__restrict__ group_off,
                                                             const int *__restrict__ order,
                                                             double *__restrict__ s2, double *__restrict__ s4)
{
    __shared__ double red[kBlock];
    const int g = blockIdx.x;
    const int p0 = group_off[g], np = group_off[g + 1] - p0;
    const long items = (long)np * n_blocks;
    double a2 = 0.0, a4 = 0.0;
    for (long it = threadIdx.x; it < items; it += kBlock) {
        const int pl = (int)(it / n_blocks);
        const int b = (int)(it - (long)pl * n_blocks);
        const size_t e = ((size_t)order[p0 + pl] * n_blocks + b) * 2;
        a2 += partial[e];
        a4 += partial[e + 1];
    }
    a2 = block_sum(a2, red);
    a4 = block_sum(a4, red);
    if (threadIdx.x == 0) {
        s2[g] = a2;
        s4[g] = a4;
    }
}

// rho_k by direct evaluation: one exact fixed-point phase
// P = nx*ux + ny*uy + nz*uz (mod 2^32) and one sincos_turns() per
// (atom, vector), about 1 ulp per term against about 5 ulp for the table
// products; compensated FP32 sums.  Used where a single frame with few
// vectors per shell leaves no averaging (the S(k)-misfit entry point).
// grid = (chunks, frames, splits).
constexpr int kDirectTile = 256;

template <int R>
__global__ __launch_bounds__(kPhaseBlock) void rho_direct_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z, int N,
    const int *__restrict__ list, int n_list, const int *__restrict__ frames,
    const uint32_t *__restrict__ kpack, const int *__restrict__ kdst, int K, int KB, double inv_ell,
    int per_split, int n_splits, double2 *__restrict__ out)
{
    __shared__ uint32_t sU[3][kDirectTile];
    const int chunk = blockIdx.x, fslot = blockIdx.y, split = blockIdx.z;
    const int tid = threadIdx.x;
    const size_t foff = (size_t)frames[fslot] * (size_t)N;
    const int a_begin = split * per_split;
    const int a_end = min(n_list, a_begin + per_split);
    uint32_t nx[R], ny[R], nz[R];
    float hr[R], hi[R], lr[R], li[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int a, b, c;
        unpack_k(kpack[(size_t)chunk * KB + r * kPhaseBlock + tid], a, b, c);
        nx[r] = (uint32_t)a; ny[r] = (uint32_t)b; nz[r] = (uint32_t)c;
        hr[r] = hi[r] = lr[r] = li[r] = 0.0f;
    }
    for (int a0 = a_begin; a0 < a_end; a0 += kDirectTile) {
        const int cnt = min(kDirectTile, a_end - a0);
        __syncthreads();
        if (tid < cnt) {
            const int j = list ? list[a0 + tid] : a0 + tid;
            sU[0][tid] = turns_fixed((double)X[foff + j] * inv_ell);
            sU[1][tid] = turns_fixed((double)Y[foff + j] * inv_ell);
            sU[2][tid] = turns_fixed((double)Z[foff + j] * inv_ell);
        }
        __syncthreads();
        float pr[R], pi[R];
#pragma unroll
        for (int r = 0; r < R; ++r) pr[r] = pi[r] = 0.0f;
        for (int i = 0; i < cnt; ++i) {
            const uint32_t ux = sU[0][i], uy = sU[1][i], uz = sU[2][i];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float2 e = sincos_turns(nx[r] * ux + ny[r] * uy + nz[r] * uz);
                pr[r] += e.x;
                pi[r] += e.y;
            }
            if ((i & 15) == 15) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    two_sum_acc(hr[r], lr[r], pr[r]);
                    two_sum_acc(hi[r], li[r], pi[r]);
                    pr[r] = pi[r] = 0.0f;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            two_sum_acc(hr[r], lr[r], pr[r]);
            two_sum_acc(hi[r], li[r], pi[r]);
        }
    }
    double2 *dst = out + ((size_t)fslot * n_splits + split) * (size_t)K;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int d = kdst[(size_t)chunk * KB + r * kPhaseBlock + tid];
        if (d >= 0) {
            const double re = (double)hr[r] + (double)lr[r];
            double im = (double)hi[r] + (double)li[r];
            if (d & (1 << 30)) im = -im;
            dst[d & ~(1 << 30)] = make_double2(re, im);
        }
    }
}

// Adjoint of the S(k) misfit on one frame: for each atom j,
//   g_j = (2 pi / l) sum_v n_v Im(conj(q_v) e^{i k_v.r_j}),  q_v = c_v rho_v,
// with the phase recomputed per (atom, vector) from the exact fixed-point
// fractions (no N x K table is stored).  Lanes are atoms, vectors are
// wave-uniform; sums are FP64.
__global__ __launch_bounds__(kBlock) void sk_adjoint_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z, int N, double inv_ell,
    const int *__restrict__ nvec, const float2 *__restrict__ q, int K, double kscale, double *__restrict__ gx,
    double *__restrict__ gy, double *__restrict__ gz)
{
    const int j = blockIdx.x * kBlock + threadIdx.x;
    const int jj = j < N ? j : N - 1;
    const uint32_t ux = turns_fixed((double)X[jj] * inv_ell);
    const uint32_t uy = turns_fixed((double)Y[jj] * inv_ell);
    const uint32_t uz = turns_fixed((double)Z[jj] * inv_ell);
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int v = 0; v < K; ++v) {
        const int nx = nvec[3 * v], ny = nvec[3 * v + 1], nz = nvec[3 * v + 2];
        const uint32_t P = (uint32_t)nx * ux + (uint32_t)ny * uy + (uint32_t)nz * uz;
        const float2 e = sincos_turns(P);
        const float2 qq = q[v];
        const double t = (double)fmaf(qq.x, e.y, -qq.y * e.x);
        ax = fma((double)nx, t, ax);
        ay = fma((double)ny, t, ay);
        az = fma((double)nz, t, az);
    }
    if (j < N) {
        gx[j] = ax * kscale;
        gy[j] = ay * kscale;
        gz[j] = az * kscale;
    }
}

template <int R, int MM>
hipError_t launch_rho_c32_t(const TrajDev &tr, const int *list, int n_list, const int *frames, int n_frames,
                            const KPlan &kp, double inv_ell, int n_splits, double2 *out, hipStream_t st)
{
    const int per_split = (n_list + n_splits - 1) / n_splits;
    const size_t lds = Geom<MM>::tab_bytes;
    auto kfn = rho_c32_kernel<R, MM>;
    hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    dim3 grid(kp.n_chunks, n_frames, n_splits);
    (void)hipGetLastError();

    hipLaunchKernelGGL(kfn, grid, dim3(kPhaseBlock), lds, st, tr.x, tr.y, tr.z, tr.N, list, n_list, frames,
                       kp.kpack, kp.kdst, kp.chunkM, kp.K, kp.KB, inv_ell, per_split, n_splits, out);
    return hipGetLastError();
}

template <int R, int MM>
hipError_t launch_fs_c32_t(const TrajDev &tr, const int *list, int n_list, const int *t0, const int *t1,
                           int n_pairs, const KPlan &kp, double inv_ell, int n_splits, double *out,
                           hipStream_t st)
{
    const int per_split = (n_list + n_splits - 1) / n_splits;
    const size_t lds = Geom<MM>::tab_bytes;
    auto kfn = fs_c32_kernel<R, MM>;
    hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    dim3 grid(kp.n_chunks, n_pairs, n_splits);
    (void)hipGetLastError();

    hipLaunchKernelGGL(kfn, grid, dim3(kPhaseBlock), lds, st, tr.x, tr.y, tr.z, tr.ix, tr.iy, tr.iz, tr.box,
                       tr.N, list, n_list, t0, t1, kp.kpack, kp.kdst, kp.chunkM, kp.K, kp.KB, inv_ell,
                       per_split, n_splits, out);
    return hipGetLastError();
}

template <int R>
hipError_t dispatch_rho_c32(int MM, const TrajDev &tr, const int *list, int n_list, const int *frames,
                            int n_frames, const KPlan &kp, double inv_ell, int n_splits, double2 *out,
                            hipStream_t st)
{
    switch (MM) {
    case 24: return launch_rho_c32_t<R, 24>(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    case 50: return launch_rho_c32_t<R, 50>(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    case 100: return launch_rho_c32_t<R, 100>(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    case 127: return launch_rho_c32_t<R, 127>(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    default: return hipErrorInvalidValue;
    }
}

template <int R>
hipError_t dispatch_fs_c32(int MM, const TrajDev &tr, const int *list, int n_list, const int *t0,
                           const int *t1, int n_pairs, const KPlan &kp, double inv_ell, int n_splits,
                           double *out, hipStream_t st)
{
    switch (MM) {
    case 24: return launch_fs_c32_t<R, 24>(tr, list, n_list, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    case 50: return launch_fs_c32_t<R, 50>(tr, list, n_list, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    case 100: return launch_fs_c32_t<R, 100>(tr, list, n_list, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    case 127: return launch_fs_c32_t<R, 127>(tr, list, n_list, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

int plan_table_width(int max_component)
{
    if (max_component <= 24) return 24;
    if (max_component <= 50) return 50;
    if (max_component <= 100) return 100;
    if (max_component <= 127) return 127;
    return -1;
}

hipError_t launch_rho(int precision, const TrajDev &tr, const int *list, int n_list, const int *frames,
                      int n_frames, const KPlan &kp, double inv_ell, int n_splits, double2 *out,
                      hipStream_t st)
{
    if (precision == 2) {
        const int per_split = (n_list + n_splits - 1) / n_splits;
        dim3 grid(kp.n_chunks, n_frames, n_splits);
        (void)hipGetLastError();
        switch (kp.R) {
        case 2: hipLaunchKernelGGL(rho_direct_kernel<2>, grid, dim3(kPhaseBlock), 0, st, tr.x, tr.y, tr.z, tr.N, list, n_list, frames, kp.kpack, kp.kdst, kp.K, kp.KB, inv_ell, per_split, n_splits, out); break;
        case 4: hipLaunchKernelGGL(rho_direct_kernel<4>, grid, dim3(kPhaseBlock), 0, st, tr.x, tr.y, tr.z, tr.N, list, n_list, frames, kp.kpack, kp.kdst, kp.K, kp.KB, inv_ell, per_split, n_splits, out); break;
        case 8: hipLaunchKernelGGL(rho_direct_kernel<8>, grid, dim3(kPhaseBlock), 0, st, tr.x, tr.y, tr.z, tr.N, list, n_list, frames, kp.kpack, kp.kdst, kp.K, kp.KB, inv_ell, per_split, n_splits, out); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    if (precision == 1) {
        const int per_split = (n_list + n_splits - 1) / n_splits;
        const double kscale = 2.0 * M_PI * inv_ell;
        dim3 grid(kp.n_chunks, n_frames, n_splits);
        (void)hipGetLastError();
        switch (kp.R) {
        case 4: hipLaunchKernelGGL(rho_f64_kernel<4>, grid, dim3(kPhaseBlock), 0, st, tr.x, tr.y, tr.z, tr.N, list, n_list, frames, kp.kpack, kp.kdst, kp.K, kp.KB, kscale, per_split, n_splits, out); break;
        case 8: hipLaunchKernelGGL(rho_f64_kernel<8>, grid, dim3(kPhaseBlock), 0, st, tr.x, tr.y, tr.z, tr.N, list, n_list, frames, kp.kpack, kp.kdst, kp.K, kp.KB, kscale, per_split, n_splits, out); break;
        case 2: hipLaunchKernelGGL(rho_f64_kernel<2>, grid, dim3(kPhaseBlock), 0, st, tr.x, tr.y, tr.z, tr.N, list, n_list, frames, kp.kpack, kp.kdst, kp.K, kp.KB, kscale, per_split, n_splits, out); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (kp.R) {
    case 4: return dispatch_rho_c32<4>(kp.MM, tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    case 8: return dispatch_rho_c32<8>(kp.MM, tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    case 2: return dispatch_rho_c32<2>(kp.MM, tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fs(int precision, const TrajDev &tr, const int *list, int n_list, const int *t0,
                     const int *t1, int n_pairs, const KPlan &kp, double inv_ell, int n_splits, double *out,
                     hipStream_t st)
{
    if (precision == 1) {
        const int per_split = (n_list + n_splits - 1) / n_splits;
        const double kscale = 2.0 * M_PI * inv_ell;
        dim3 grid(kp.n_chunks, n_pairs, n_splits);
        (void)hipGetLastError();
        switch (kp.R) {
        case 4: hipLaunchKernelGGL(fs_f64_kernel<4>, grid, dim3(kPhaseBlock), 0, st, tr.x, tr.y, tr.z, tr.ix, tr.iy, tr.iz, tr.box, tr.N, list, n_list, t0, t1, kp.kpack, kp.kdst, kp.K, kp.KB, kscale, per_split, n_splits, out); break;
        case 8: hipLaunchKernelGGL(fs_f64_kernel<8>, grid, dim3(kPhaseBlock), 0, st, tr.x, tr.y, tr.z, tr.ix, tr.iy, tr.iz, tr.box, tr.N, list, n_list, t0, t1, kp.kpack, kp.kdst, kp.K, kp.KB, kscale, per_split, n_splits, out); break;
        case 2: hipLaunchKernelGGL(fs_f64_kernel<2>, grid, dim3(kPhaseBlock), 0, st, tr.x, tr.y, tr.z, tr.ix, tr.iy, tr.iz, tr.box, tr.N, list, n_list, t0, t1, kp.kpack, kp.kdst, kp.K, kp.KB, kscale, per_split, n_splits, out); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (kp.R) {
    case 4: return dispatch_fs_c32<4>(kp.MM, tr, list, n_list, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    case 8: return dispatch_fs_c32<8>(kp.MM, tr, list, n_list, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    case 2: return dispatch_fs_c32<2>(kp.MM, tr, list, n_list, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_split_reduce2(const double2 *in, int n_rows, int n_splits, int K, double2 *out, hipStream_t st)
{
    const size_t n = (size_t)n_rows * K;
    int blocks = (int)std::min<size_t>((n + kBlock - 1) / kBlock, 65535);
    if (blocks < 1) blocks = 1;
    (void)hipGetLastError();

    hipLaunchKernelGGL(split_reduce2_kernel, dim3(blocks), dim3(kBlock), 0, st, in, n_rows, n_splits, K, out);
    return hipGetLastError();
}

hipError_t launch_corr_shell(const double2 *rho1, const double2 *rho2, int K, const int *group_off,
                             const int *row0, const int *row1, int n_groups, const int *shell_off,
                             int n_shells, const double *norm, double *out, hipStream_t st)
{
    (void)hipGetLastError();

    hipLaunchKernelGGL(corr_shell_kernel, dim3(n_shells, n_groups), dim3(kBlock), 0, st, rho1, rho2, K,
                       group_off, row0, row1, shell_off, n_shells, norm, out);
    return hipGetLastError();
}

hipError_t launch_fs_shell(const double *V, int n_splits, int K, const int *group_off, const int *order,
                           int n_groups, const int *shell_off, int n_shells, const double *norm,
                           double *out, hipStream_t st)
{
    (void)hipGetLastError();

    hipLaunchKernelGGL(fs_shell_kernel, dim3(n_shells, n_groups), dim3(kBlock), 0, st, V, n_splits, K,
                       group_off, order, shell_off, n_shells, norm, out);
    return hipGetLastError();
}

hipError_t launch_msd(const TrajDev &tr, const int *list, int n_list, const int *t0, const int *t1,
                      int n_pairs, int n_blocks, double *partial, hipStream_t st)
{
    (void)hipGetLastError();
    hipLaunchKernelGGL(msd_kernel, dim3(n_blocks, n_pairs), dim3(kBlock), 0, st, tr.x, tr.y, tr.z, tr.ix,
                       tr.iy, tr.iz, tr.box, tr.N, list, n_list, t0, t1, partial);
    return hipGetLastError();
}

hipError_t launch_msd_reduce(const double *partial, int n_blocks, const int *group_off, const int *order,
                             int n_groups, double *s2, double *s4, hipStream_t st)
{
    (void)hipGetLastError();

    hipLaunchKernelGGL(msd_reduce_kernel, dim3(n_groups), dim3(kBlock), 0, st, partial, n_blocks, group_off,
                       order, s2, s4);
    return hipGetLastError();
}

hipError_t launch_sk_adjoint(const float *x, const float *y, const float *z, int N, double inv_ell, const int *nvec,
                             const float2 *q, int K, double *gx, double *gy, double *gz, hipStream_t st)
{
    const int blocks = (N + kBlock - 1) / kBlock;
    (void)hipGetLastError();
    hipLaunchKernelGGL(sk_adjoint_kernel, dim3(blocks), dim3(kBlock), 0, st, x, y, z, N, inv_ell, nvec, q, K,
                       2.0 * M_PI * inv_ell, gx, gy, gz);
    return hipGetLastError();
}

}  // namespace amdat
