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

// Hot loop over the atom tiles of one block, shared by the single-trajectory
// and batched phase kernels.  FS = false: rho (complex sums into h0/l0 = Re,
// h1/l1 = Im); FS = true: Fs (real part only, into h0/l0).  uf(a, c) gives
// the fixed-point fraction of list position a on axis c.
template <int R, int MM, bool FS, typename UF>
__device__ __forceinline__ void phase_tiles(const uint32_t *__restrict__ kp, int Mc, int a_begin, int a_end,
                                            float2 *tab, float2 *anc, uint32_t *sU, int *sValid, const UF uf,
                                            float (&h0)[R], float (&l0)[R], float (&h1)[R], float (&l1)[R])
{
    using Gm = Geom<MM>;
    const int tid = threadIdx.x;
    const float2 *px[R], *py[R], *pz[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int nx, ny, nz;
        unpack_k(kp[r * kPhaseBlock + tid], nx, ny, nz);
        px[r] = tab + Gm::OX + nx;
        py[r] = tab + Gm::OY + ny;
        pz[r] = tab + Gm::OZ + nz;
        h0[r] = l0[r] = h1[r] = l1[r] = 0.0f;
    }
    for (int a0 = a_begin; a0 < a_end; a0 += Gm::A) {
        __syncthreads();
        if (tid < 3 * Gm::A) {
            const int i = tid / 3, c = tid - 3 * (tid / 3);
            const int a = a0 + i;
            sU[tid] = a < a_end ? uf(a, c) : 0u;
            if (c == 0) sValid[i] = a < a_end;
        }
        __syncthreads();
        build_tables<MM>(tab, anc, sU, sValid, Mc);
        __syncthreads();
        float p0[R], p1[R];
#pragma unroll
        for (int r = 0; r < R; ++r) p0[r] = p1[r] = 0.0f;
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
            p0[r] = fmaf(qr, ez.x, fmaf(-qi, ez.y, p0[r]));
            if constexpr (!FS) p1[r] = fmaf(qr, ez.y, fmaf(qi, ez.x, p1[r]));
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            two_sum_acc(h0[r], l0[r], p0[r]);
            if constexpr (!FS) two_sum_acc(h1[r], l1[r], p1[r]);
        }
    }
}

// Fixed-point fraction of a coordinate (rho).
struct CoordFrac {
    const float *X, *Y, *Z;
    const int *list;
    size_t foff;
    double inv_ell;
    __device__ __forceinline__ uint32_t operator()(int a, int c) const
    {
        const size_t o = foff + (list ? list[a] : a);
        const float v = c == 0 ? X[o] : (c == 1 ? Y[o] : Z[o]);
        return turns_fixed((double)v * inv_ell);
    }
};

// Fixed-point fraction of a displacement dr = (x1 + i1*L1) - (x0 + i0*L0),
// rebuilt exactly in FP64 (Fs).
struct DispFrac {
    const float *X, *Y, *Z;
    const int16_t *IX, *IY, *IZ;
    const float *L0, *L1;
    const int *list;
    size_t o0, o1;
    double inv_ell;
    __device__ __forceinline__ uint32_t operator()(int a, int c) const
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
        return turns_fixed(d * inv_ell);
    }
};

// rho over one block's atoms: compensated sums (hr + lr, hi + li) per slot.
template <int R, int MM>
__device__ __forceinline__ void rho_block(const float *__restrict__ X, const float *__restrict__ Y,
                                          const float *__restrict__ Z, const int *__restrict__ list, size_t foff,
                                          int a_begin, int a_end, double inv_ell, const uint32_t *__restrict__ kp,
                                          int Mc, float2 *tab, float2 *anc, uint32_t *sU, int *sValid, float (&hr)[R],
                                          float (&lr)[R], float (&hi)[R], float (&li)[R])
{
    const CoordFrac uf{X, Y, Z, list, foff, inv_ell};
    phase_tiles<R, MM, false>(kp, Mc, a_begin, a_end, tab, anc, sU, sValid, uf, hr, lr, hi, li);
}

template <int R>
__device__ __forceinline__ void rho_write(const int *__restrict__ kd, const float (&hr)[R], const float (&lr)[R],
                                          const float (&hi)[R], const float (&li)[R], double2 *__restrict__ dst)
{
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int d = kd[r * kPhaseBlock + threadIdx.x];
        if (d >= 0) {
            const double re = (double)hr[r] + (double)lr[r];
            double im = (double)hi[r] + (double)li[r];
            if (d & (1 << 30)) im = -im;
            dst[d & ~(1 << 30)] = make_double2(re, im);
        }
    }
}

// Fs over one block's atoms for the pair (o0, o1): compensated sum of cos(k.dr).
template <int R, int MM>
__device__ __forceinline__ void fs_block(const float *__restrict__ X, const float *__restrict__ Y,
                                         const float *__restrict__ Z, const int16_t *__restrict__ IX,
                                         const int16_t *__restrict__ IY, const int16_t *__restrict__ IZ,
                                         const float *__restrict__ L0, const float *__restrict__ L1,
                                         const int *__restrict__ list, size_t o0, size_t o1, int a_begin, int a_end,
                                         double inv_ell, const uint32_t *__restrict__ kp, int Mc, float2 *tab,
                                         float2 *anc, uint32_t *sU, int *sValid, float (&hs)[R], float (&ls)[R])
{
    const DispFrac uf{X, Y, Z, IX, IY, IZ, L0, L1, list, o0, o1, inv_ell};
    float u1[R], u2[R];
    phase_tiles<R, MM, true>(kp, Mc, a_begin, a_end, tab, anc, sU, sValid, uf, hs, ls, u1, u2);
}

template <int R>
__device__ __forceinline__ void fs_write(const int *__restrict__ kd, const float (&hs)[R], const float (&ls)[R],
                                         double *__restrict__ dst)
{
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int d = kd[r * kPhaseBlock + threadIdx.x];
        if (d >= 0) dst[d & ~(1 << 30)] = (double)hs[r] + (double)ls[r];
    }
}

#define AMDAT_PHASE_SHARED(MM)                         \
    extern __shared__ float2 tab[];                    \
    __shared__ float2 anc[3 * Geom<MM>::A * Geom<MM>::NA]; \
    __shared__ uint32_t sU[3 * Geom<MM>::A];           \
    __shared__ int sValid[Geom<MM>::A]

// rho_k(t), compensated FP32.  grid = (chunks, frames, splits).
template <int R, int MM>
__global__ __launch_bounds__(kPhaseBlock, 4) void rho_c32_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z, int N,
    const int *__restrict__ list, int n_list, const int *__restrict__ frames,
    const uint32_t *__restrict__ kpack, const int *__restrict__ kdst, const int *__restrict__ chunkM,
    int K, int KB, double inv_ell, int per_split, int n_splits, double2 *__restrict__ out)
{
    AMDAT_PHASE_SHARED(MM);
    const int chunk = blockIdx.x, fslot = blockIdx.y, split = blockIdx.z;
    const int a_begin = split * per_split;
    const int a_end = min(n_list, a_begin + per_split);
    float hr[R], lr[R], hi[R], li[R];
    rho_block<R, MM>(X, Y, Z, list, (size_t)frames[fslot] * (size_t)N, a_begin, a_end, inv_ell,
                     kpack + (size_t)chunk * KB, chunkM[chunk], tab, anc, sU, sValid, hr, lr, hi, li);
    rho_write<R>(kdst + (size_t)chunk * KB, hr, lr, hi, li, out + ((size_t)fslot * n_splits + split) * (size_t)K);
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
    AMDAT_PHASE_SHARED(MM);
    const int chunk = blockIdx.x, pslot = blockIdx.y, split = blockIdx.z;
    const int t0 = T0[pslot], t1 = T1[pslot];
    const int a_begin = split * per_split;
    const int a_end = min(n_list, a_begin + per_split);
    float hs[R], ls[R];
    fs_block<R, MM>(X, Y, Z, IX, IY, IZ, box + 3 * t0, box + 3 * t1, list, (size_t)t0 * N, (size_t)t1 * N, a_begin,
                    a_end, inv_ell, kpack + (size_t)chunk * KB, chunkM[chunk], tab, anc, sU, sValid, hs, ls);
    fs_write<R>(kdst + (size_t)chunk * KB, hs, ls, out + ((size_t)pslot * n_splits + split) * (size_t)K);
}

// Next work item of a persistent launch: thread 0 draws from the queue
// counter and the block reads the value from LDS.
__device__ __forceinline__ int next_item(int *counter, int *slot)
{
    __syncthreads();  // every thread has read the previous value, and the tables are free
    if (threadIdx.x == 0) *slot = atomicAdd(counter, 1);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*slot);  // same value in every lane: keep it scalar
}

// Batched S(k) numerators, persistent: item = (trajectory, chunk, frame
// group, -).  For each frame of the group the block forms rho over all of
// the trajectory's atoms and adds |rho_v|^2 into row [group][v] of the
// trajectory's output (FP64; each slot is owned by one thread of one item,
// so the sum order is fixed).  All trajectories share one k-plan: the same
// integer vectors, k = 2 pi n / l_b with each trajectory's own l_b.
template <int R, int MM>
__global__ __launch_bounds__(kPhaseBlock, 4) void sk_batch_kernel(
    const BatchTraj *__restrict__ trajs, const int4 *__restrict__ items, int n_items, int *counter,
    const int *__restrict__ frames, int n_frames, int group_size, const uint32_t *__restrict__ kpack,
    const int *__restrict__ kdst, const int *__restrict__ chunkM, int K, int KB)
{
    AMDAT_PHASE_SHARED(MM);
    __shared__ int s_item;
    for (;;) {
        const int q = next_item(counter, &s_item);
        if (q >= n_items) break;
        const int4 it = items[q];
        const BatchTraj &t = trajs[it.x];
        const int chunk = it.y, g = it.z;
        const int f_end = min(n_frames, (g + 1) * group_size);
        const int *kd = kdst + (size_t)chunk * KB;
        double *row = t.out + (size_t)g * K;
        for (int f = g * group_size; f < f_end; ++f) {
            float hr[R], lr[R], hi[R], li[R];
            rho_block<R, MM>(t.x, t.y, t.z, t.list, (size_t)frames[f] * (size_t)t.N, 0, t.n_list, t.inv_ell,
                             kpack + (size_t)chunk * KB, chunkM[chunk], tab, anc, sU, sValid, hr, lr, hi, li);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int d = kd[r * kPhaseBlock + threadIdx.x];
                if (d >= 0) {
                    const double re = (double)hr[r] + (double)lr[r];
                    const double im = (double)hi[r] + (double)li[r];
                    double *slot = row + (d & ~(1 << 30));
                    const double prev = f == g * group_size ? 0.0 : *slot;
                    *slot = prev + (re * re + im * im);
                }
            }
        }
    }
}

// Batched Fs, persistent: item = (trajectory, global chunk, time gap, -)
// over concatenated per-trajectory plans.  The block runs the gap's pairs
// in reference order and adds each pair's sum over atoms of cos(k.dr) into
// row [gap][v] of the trajectory's output (FP64, fixed order).
template <int R, int MM>
__global__ __launch_bounds__(kPhaseBlock, 4) void fs_batch_kernel(
    const BatchTraj *__restrict__ trajs, const int4 *__restrict__ items, int n_items, int *counter,
    const int *__restrict__ T0, const int *__restrict__ T1, const int *__restrict__ group_off,
    const int *__restrict__ order, const uint32_t *__restrict__ kpack, const int *__restrict__ kdst,
    const int *__restrict__ chunkM, int KB)
{
    AMDAT_PHASE_SHARED(MM);
    __shared__ int s_item;
    for (;;) {
        const int q = next_item(counter, &s_item);
        if (q >= n_items) break;
        const int4 it = items[q];
        const BatchTraj &t = trajs[it.x];
        const int chunk = it.y, g = it.z;
        const int *kd = kdst + (size_t)chunk * KB;
        double *row = t.out + (size_t)g * t.K;
        const int p_begin = group_off[g], p_end = group_off[g + 1];
        for (int pq = p_begin; pq < p_end; ++pq) {
            const int p = order[pq];
            const int t0 = T0[p], t1 = T1[p];
            float hs[R], ls[R];
            fs_block<R, MM>(t.x, t.y, t.z, t.ix, t.iy, t.iz, t.box + 3 * t0, t.box + 3 * t1, t.list,
                            (size_t)t0 * t.N, (size_t)t1 * t.N, 0, t.n_list, t.inv_ell, kpack + (size_t)chunk * KB,
                            chunkM[chunk], tab, anc, sU, sValid, hs, ls);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int d = kd[r * kPhaseBlock + threadIdx.x];
                if (d >= 0) {
                    double *slot = row + (d & ~(1 << 30));
                    const double prev = pq == p_begin ? 0.0 : *slot;
                    *slot = prev + ((double)hs[r] + (double)ls[r]);
                }
            }
        }
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

// Batched MSD/NGP partials: item = (trajectory, pair) -> FP64 sums of
// |dr|^2 and |dr|^4 over the trajectory's atoms, at row trajectory*n_pairs + pair.
__global__ __launch_bounds__(kBlock) void msd_batch_kernel(const BatchTraj *__restrict__ trajs,
                                                            const int2 *__restrict__ items,
                                                            const int *__restrict__ T0, const int *__restrict__ T1,
                                                            int n_pairs, double *__restrict__ partial)
{
    __shared__ double red[kBlock];
    const int2 it = items[blockIdx.x];
    const BatchTraj &t = trajs[it.x];
    const int t0 = T0[it.y], t1 = T1[it.y];
    const size_t o0 = (size_t)t0 * (size_t)t.N, o1 = (size_t)t1 * (size_t)t.N;
    const double L0x = t.box[3 * t0], L0y = t.box[3 * t0 + 1], L0z = t.box[3 * t0 + 2];
    const double L1x = t.box[3 * t1], L1y = t.box[3 * t1 + 1], L1z = t.box[3 * t1 + 2];
    double a2 = 0.0, a4 = 0.0;
    for (int a = threadIdx.x; a < t.n_list; a += kBlock) {
        const int j = t.list ? t.list[a] : a;
        const double dx = ((double)t.x[o1 + j] - (double)t.x[o0 + j]) + ((double)t.ix[o1 + j] * L1x - (double)t.ix[o0 + j] * L0x);
        const double dy = ((double)t.y[o1 + j] - (double)t.y[o0 + j]) + ((double)t.iy[o1 + j] * L1y - (double)t.iy[o0 + j] * L0y);
        const double dz = ((double)t.z[o1 + j] - (double)t.z[o0 + j]) + ((double)t.iz[o1 + j] * L1z - (double)t.iz[o0 + j] * L0z);
        const double r2 = dx * dx + dy * dy + dz * dz;
        a2 += r2;
        a4 += r2 * r2;
    }
    a2 = block_sum(a2, red);
    a4 = block_sum(a4, red);
    if (threadIdx.x == 0) {
        const size_t o = 2 * ((size_t)it.x * n_pairs + it.y);
        partial[o] = a2;
        partial[o + 1] = a4;
    }
}

// grid = (time gaps, trajectories): sums of the partials of the gap's pairs
// in reference order; partial rows are [trajectory][pair].
__global__ __launch_bounds__(64) void msd_batch_reduce_kernel(const double *__restrict__ partial, int n_pairs,
                                                               const int *__restrict__ group_off,
                                                               const int *__restrict__ order, int ntg,
                                                               double *__restrict__ s2, double *__restrict__ s4)
{
    const int g = blockIdx.x, b = blockIdx.y;
    if (threadIdx.x != 0) return;
    const double *pb = partial + 2 * (size_t)b * n_pairs;
    double a2 = 0.0, a4 = 0.0;
    for (int q = group_off[g]; q < group_off[g + 1]; ++q) {
        a2 += pb[2 * (size_t)order[q]];
        a4 += pb[2 * (size_t)order[q] + 1];
    }
    s2[(size_t)b * ntg + g] = a2;
    s4[(size_t)b * ntg + g] = a4;
}

// grid = (shells, trajectories): S[b][s] = sum over frame groups and the
// shell's vectors of the |rho|^2 rows, / H / norm[b].
__global__ __launch_bounds__(kBlock) void sk_batch_reduce_kernel(const BatchTraj *__restrict__ trajs, int n_groups,
                                                                  int K, const int *__restrict__ shell_off,
                                                                  int n_shells, const double *__restrict__ norm,
                                                                  double *__restrict__ S)
{
    __shared__ double red[kBlock];
    const int s = blockIdx.x, b = blockIdx.y;
    const double *rows = trajs[b].out;
    const int v0 = shell_off[s], H = shell_off[s + 1] - v0;
    const long n = (long)n_groups * H;
    double acc = 0.0;
    for (long i = threadIdx.x; i < n; i += kBlock) {
        const int g = (int)(i / H);
        acc += rows[(size_t)g * K + v0 + (int)(i - (long)g * H)];
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) {
        const double nrm = norm[b];
        S[(size_t)b * n_shells + s] = (H > 0 && nrm != 0.0) ? acc / (double)H / nrm : 0.0;
    }
}

// grid = (max shells, time gaps, trajectories) over ragged per-trajectory
// shells: out[out_off[b] + g*nsh[b] + s] = sum_{v in s} V_b[g][v] / H / norm[b*ntg + g].
__global__ __launch_bounds__(kBlock) void fs_batch_reduce_kernel(
    const BatchTraj *__restrict__ trajs, const int *__restrict__ shell_off, const int *__restrict__ shell_base,
    const int *__restrict__ nsh, int ntg, const double *__restrict__ norm, const int *__restrict__ out_off,
    double *__restrict__ out)
{
    __shared__ double red[kBlock];
    const int s = blockIdx.x, g = blockIdx.y, b = blockIdx.z;
    if (s >= nsh[b]) return;  // uniform across the block
    const BatchTraj &t = trajs[b];
    const int *so = shell_off + shell_base[b];
    const int v0 = so[s], H = so[s + 1] - v0;
    const double nrm = norm[(size_t)b * ntg + g];
    double acc = 0.0;
    if (nrm != 0.0)
        for (int v = threadIdx.x; v < H; v += kBlock) acc += t.out[(size_t)g * t.K + v0 + v];
    acc = block_sum(acc, red);
    if (threadIdx.x == 0)
        out[out_off[b] + (size_t)g * nsh[b] + s] = (H > 0 && nrm != 0.0) ? acc / (double)H / nrm : 0.0;
}

__global__ __launch_bounds__(64) void fill_int_kernel(int *p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
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

namespace {

// Blocks of a persistent launch: what the device keeps resident, capped
// by the number of work items.
template <typename F>
hipError_t persistent_grid(F kfn, size_t lds, int n_items, int *grid)
{
    int dev = 0, n_cu = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, kPhaseBlock, lds);
    if (e != hipSuccess) return e;
    *grid = std::max(1, std::min(n_items, std::max(1, per_cu) * n_cu));
    return hipSuccess;
}

hipError_t reset_counter(int *counter, hipStream_t st)
{
    (void)hipGetLastError();
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, st, counter, 1, 0);
    return hipGetLastError();
}

template <int R, int MM>
hipError_t launch_sk_batch_t(const BatchTraj *trajs, const int4 *items, int n_items, int *counter, const int *frames,
                             int n_frames, int group_size, const KPlan &kp, hipStream_t st)
{
    const size_t lds = Geom<MM>::tab_bytes;
    auto kfn = sk_batch_kernel<R, MM>;
    hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int grid = 0;
    if (e == hipSuccess) e = persistent_grid(kfn, lds, n_items, &grid);
    if (e == hipSuccess) e = reset_counter(counter, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kfn, dim3(grid), dim3(kPhaseBlock), lds, st, trajs, items, n_items, counter, frames, n_frames,
                       group_size, kp.kpack, kp.kdst, kp.chunkM, kp.K, kp.KB);
    return hipGetLastError();
}

template <int R, int MM>
hipError_t launch_fs_batch_t(const BatchTraj *trajs, const int4 *items, int n_items, int *counter, const int *t0,
                             const int *t1, const int *group_off, const int *order, const KPlan &kp, hipStream_t st)
{
    const size_t lds = Geom<MM>::tab_bytes;
    auto kfn = fs_batch_kernel<R, MM>;
    hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int grid = 0;
    if (e == hipSuccess) e = persistent_grid(kfn, lds, n_items, &grid);
    if (e == hipSuccess) e = reset_counter(counter, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kfn, dim3(grid), dim3(kPhaseBlock), lds, st, trajs, items, n_items, counter, t0, t1, group_off,
                       order, kp.kpack, kp.kdst, kp.chunkM, kp.KB);
    return hipGetLastError();
}

#define AMDAT_DISPATCH_MM(FN, R, ...)                          \
    switch (kp.MM) {                                           \
    case 24: return FN<R, 24>(__VA_ARGS__);                    \
    case 50: return FN<R, 50>(__VA_ARGS__);                    \
    case 100: return FN<R, 100>(__VA_ARGS__);                  \
    case 127: return FN<R, 127>(__VA_ARGS__);                  \
    default: return hipErrorInvalidValue;                      \
    }

}  // namespace

hipError_t launch_sk_batch(const BatchTraj *trajs, const int4 *items, int n_items, int *counter, const int *frames,
                           int n_frames, int group_size, const KPlan &kp, hipStream_t st)
{
    if (n_items <= 0) return hipSuccess;
    switch (kp.R) {
    case 2: AMDAT_DISPATCH_MM(launch_sk_batch_t, 2, trajs, items, n_items, counter, frames, n_frames, group_size, kp, st)
    case 4: AMDAT_DISPATCH_MM(launch_sk_batch_t, 4, trajs, items, n_items, counter, frames, n_frames, group_size, kp, st)
    case 8: AMDAT_DISPATCH_MM(launch_sk_batch_t, 8, trajs, items, n_items, counter, frames, n_frames, group_size, kp, st)
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fs_batch(const BatchTraj *trajs, const int4 *items, int n_items, int *counter, const int *t0,
                           const int *t1, const int *group_off, const int *order, const KPlan &kp, hipStream_t st)
{
    if (n_items <= 0) return hipSuccess;
    switch (kp.R) {
    case 2: AMDAT_DISPATCH_MM(launch_fs_batch_t, 2, trajs, items, n_items, counter, t0, t1, group_off, order, kp, st)
    case 4: AMDAT_DISPATCH_MM(launch_fs_batch_t, 4, trajs, items, n_items, counter, t0, t1, group_off, order, kp, st)
    case 8: AMDAT_DISPATCH_MM(launch_fs_batch_t, 8, trajs, items, n_items, counter, t0, t1, group_off, order, kp, st)
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_msd_batch(const BatchTraj *trajs, const int2 *items, int n_items, const int *t0, const int *t1,
                            int n_pairs, double *partial, hipStream_t st)
{
    if (n_items <= 0) return hipSuccess;
    (void)hipGetLastError();
    hipLaunchKernelGGL(msd_batch_kernel, dim3(n_items), dim3(kBlock), 0, st, trajs, items, t0, t1, n_pairs, partial);
    return hipGetLastError();
}

hipError_t launch_msd_batch_reduce(const double *partial, int n_pairs, const int *group_off, const int *order,
                                   int ntg, int n_traj, double *s2, double *s4, hipStream_t st)
{
    (void)hipGetLastError();
    hipLaunchKernelGGL(msd_batch_reduce_kernel, dim3(ntg, n_traj), dim3(64), 0, st, partial, n_pairs, group_off,
                       order, ntg, s2, s4);
    return hipGetLastError();
}

hipError_t launch_sk_batch_reduce(const BatchTraj *trajs, int n_traj, int n_groups, int K, const int *shell_off,
                                  int n_shells, const double *norm, double *S, hipStream_t st)
{
    (void)hipGetLastError();
    hipLaunchKernelGGL(sk_batch_reduce_kernel, dim3(n_shells, n_traj), dim3(kBlock), 0, st, trajs, n_groups, K,
                       shell_off, n_shells, norm, S);
    return hipGetLastError();
}

hipError_t launch_fs_batch_reduce(const BatchTraj *trajs, int n_traj, const int *shell_off, const int *shell_base,
                                  const int *nsh, int max_nsh, int ntg, const double *norm, const int *out_off,
                                  double *out, hipStream_t st)
{
    (void)hipGetLastError();
    hipLaunchKernelGGL(fs_batch_reduce_kernel, dim3(max_nsh, ntg, n_traj), dim3(kBlock), 0, st, trajs, shell_off,
                       shell_base, nsh, ntg, norm, out_off, out);
    return hipGetLastError();
}

}  // namespace amdat
