// kernels.hip -- gfx950 kernels of the reciprocal-space and time-correlation
// path: rho_k(t), Fs(k,t), MSD/NGP moments, and the FP64 reductions that turn
// per-vector sums into per-shell S(k), F(k,t) and Fs(k,t).
//
// The shared tile machinery of the phase kernels is in phase_device.h.
#include <algorithm>

#include "abi_internal.h"
#include "kernels.h"
#include "phase_device.h"

namespace amdat {

namespace {

// rho_k(t), FP32 tables and tile partials, FP64 sums.  Work items (chunk, frame slot, split),
// numbered by xcd_item.  GROUPED: the plan's waves carry group codes.
template <int R, int MM, bool GROUPED>
__global__ __launch_bounds__(kPhaseBlock, 4) void rho_c32_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z, int N,
    const int *__restrict__ list, int n_list, const int *__restrict__ frames, int n_frames, int n_chunks,
    const uint32_t *__restrict__ kpack, const int *__restrict__ kdst, const int4 *__restrict__ range,
    int K, int KB, double inv_ell, int per_split, int n_splits, double2 *__restrict__ out)
{
    AMDAT_PHASE_SHARED(MM);
    int chunk, row;
    if (!xcd_item(n_chunks, n_frames * n_splits, chunk, row)) return;
    const int fslot = row % n_frames, split = row / n_frames;
    const int a_begin = split * per_split;
    const int a_end = min(n_list, a_begin + per_split);
    double re[R], im[R];
    rho_block<R, MM, GROUPED>(X, Y, Z, list, (size_t)frames[fslot] * (size_t)N, a_begin, a_end, inv_ell,
                              kpack + (size_t)chunk * KB, range[chunk], tab, anc, sU, re, im);
    rho_write<R>(kdst + (size_t)chunk * KB, re, im, out + ((size_t)fslot * n_splits + split) * (size_t)K);
}

// Fs per (pair, vector), compensated FP32: sum over the split's atoms of
// cos(k.dr).  Work items (chunk, pair slot, split), numbered by xcd_item.
template <int R, int MM, bool GROUPED>
__global__ __launch_bounds__(kPhaseBlock, 4) void fs_c32_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z,
    const int16_t *__restrict__ IX, const int16_t *__restrict__ IY, const int16_t *__restrict__ IZ,
    const float *__restrict__ box, int N, const int *__restrict__ list, int n_list, const int2 *__restrict__ seg,
    const int *__restrict__ T0, const int *__restrict__ T1, int n_pairs, int n_chunks,
    const uint32_t *__restrict__ kpack, const int *__restrict__ kdst, const int4 *__restrict__ range, int K, int KB,
    double inv_ell, int per_split, int n_splits, double *__restrict__ out)
{
    AMDAT_PHASE_SHARED(MM);
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
    double sum[R];
    fs_block<R, MM, GROUPED>(X, Y, Z, IX, IY, IZ, box + 3 * t0, box + 3 * t1, list, (size_t)t0 * N, (size_t)t1 * N,
                             a_begin, a_end, inv_ell, kpack + (size_t)chunk * KB, range[chunk], tab, anc, sU,
                             sum);
    fs_write<R>(kdst + (size_t)chunk * KB, sum, out + ((size_t)pslot * n_splits + split) * (size_t)K);
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
    const float *__restrict__ box, int N, const int *__restrict__ list, int n_list, const int2 *__restrict__ seg,
    const int *__restrict__ T0, const int *__restrict__ T1, const uint32_t *__restrict__ kpack,
    const int *__restrict__ kdst, int K, int KB, double kscale, int per_split, int n_splits,
    double *__restrict__ out)
{
    const int chunk = blockIdx.x, pslot = blockIdx.y, split = blockIdx.z;
    const int tid = threadIdx.x;
    const int t0 = T0[pslot], t1 = T1[pslot];
    if (seg) {
        const int2 sg = seg[pslot];
        list += sg.x;
        n_list = sg.y;
    }
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

// out[f][k] = sum over s of in[f][s][k].  A block takes 32 consecutive k of
// one row and its 8 wave-halves (32 lanes each) sum the splits s = g, g+8,
// g+16, ... (g = 0..7) in order; the eight partials are then added in
// order g = 0..7, so the result does not depend on the launch.  Each
// half-wave reads 32 consecutive entries (512 bytes) per split.
constexpr int kSplitCols = 32, kSplitGroups = kBlock / kSplitCols;
__global__ __launch_bounds__(kBlock) void split_reduce2_kernel(const double2 *__restrict__ in, int n_rows,
                                                                int n_splits, int K, double2 *__restrict__ out)
{
    __shared__ double2 part[kSplitGroups][kSplitCols];
    const int col_blocks = (K + kSplitCols - 1) / kSplitCols;
    const int f = blockIdx.x / col_blocks, k0 = (blockIdx.x - f * col_blocks) * kSplitCols;
    const int c = threadIdx.x % kSplitCols, g = threadIdx.x / kSplitCols;
    const int k = k0 + c;
    double re = 0.0, im = 0.0;
    if (f < n_rows && k < K)
        for (int sp = g; sp < n_splits; sp += kSplitGroups) {
            const double2 v = in[((size_t)f * n_splits + sp) * K + k];
            re += v.x;
            im += v.y;
        }
    part[g][c] = make_double2(re, im);
    __syncthreads();
    if (g == 0 && f < n_rows && k < K) {
        double2 acc = part[0][c];
        for (int q = 1; q < kSplitGroups; ++q) {
            acc.x += part[q][c].x;
            acc.y += part[q][c].y;
        }
        out[(size_t)f * K + k] = acc;
    }
}

// Adjoint weights of the S(k) misfit from the forward sums rho[K] (one
// frame): per shell s (vectors shell_off[s] - shell_off[0] ...), S_s =
// sum |rho_v|^2 / (H_s N), c_s = -4 w_s (S_s - T_s) / (H_s N), q_v = c_s
// rho_v (FP32); the loss sum_s w_s (S_s - T_s)^2 added in shell order by
// thread 0.  One block of kWeightsBlock threads; wave w takes shells w,
// w + 16, ...: its lanes sum vectors l, l + 64, ... and a fixed shuffle
// tree adds the lanes, so the result does not depend on the launch.
// res[0] = loss, res[1 + s] = S_s.  (Each shell is a chain of dependent
// loads and shuffles: 16 waves instead of 4 took the kernel of a 50-shell
// call from 20.6 to 10.4 us.)
constexpr int kWeightsBlock = 1024;
__global__ __launch_bounds__(kWeightsBlock) void adjoint_weights_kernel(const double2 *__restrict__ rho,
                                                                  const int *__restrict__ shell_off, int n_shells,
                                                                  int n_atoms, const double *__restrict__ wt,
                                                                  double *__restrict__ res, float2 *__restrict__ q)
{
    // wt[s] = weight, wt[n_shells + s] = target; res[1 + n_shells + s] holds
    // the shell's loss term until thread 0 adds them
    const int v0 = shell_off[0], lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int sh = wave; sh < n_shells; sh += kWeightsBlock / 64) {
        const int a = shell_off[sh] - v0, b = shell_off[sh + 1] - v0, H = b - a;
        double acc = 0.0;
        for (int v = a + lane; v < b; v += 64) acc += rho[v].x * rho[v].x + rho[v].y * rho[v].y;
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        const double S = H ? acc / ((double)H * n_atoms) : 0.0;
        const double d = S - wt[n_shells + sh];
        const double c = H ? -4.0 * wt[sh] * d / ((double)H * n_atoms) : 0.0;
        for (int v = a + lane; v < b; v += 64) q[v] = make_float2((float)(c * rho[v].x), (float)(c * rho[v].y));
        if (lane == 0) {
            res[1 + sh] = S;
            res[1 + n_shells + sh] = wt[sh] * d * d;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double L = 0.0;
        for (int sh = 0; sh < n_shells; ++sh) L += res[1 + n_shells + sh];
        res[0] = L;
    }
}


// grid = (columns, groups).
__global__ __launch_bounds__(kBlock) void corr_shell_kernel(
    const double2 *__restrict__ rho1, const double2 *__restrict__ rho2, int K,
    const int *__restrict__ group_off, const int *__restrict__ row0, const int *__restrict__ row1,
    const int *__restrict__ vbeg, const int *__restrict__ vend, int n_shells, const double *__restrict__ norm,
    double *__restrict__ out)
{
    __shared__ double red[kBlock];
    const int s = blockIdx.x, g = blockIdx.y;
    const int v0 = vbeg[s], H = vend[s] - v0;
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

__global__ __launch_bounds__(kBlock) void rho_add_kernel(const double2 *__restrict__ a, const double2 *__restrict__ b,
                                                         size_t n, double2 *__restrict__ out)
{
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const double2 x = a[i], y = b[i];
        out[i] = make_double2(x.x + y.x, x.y + y.y);
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

// Work items (atom block, pair), numbered by xcd_atom_block.
__global__ __launch_bounds__(kBlock) void msd_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z,
    const int16_t *__restrict__ IX, const int16_t *__restrict__ IY, const int16_t *__restrict__ IZ,
    const float *__restrict__ box, int N, const int *__restrict__ list, int n_list, const int2 *__restrict__ seg,
    const int *__restrict__ T0, const int *__restrict__ T1, int n_pairs, int n_blocks, double *__restrict__ partial)
{
    __shared__ double red[kBlock];
    int bx, p;
    if (!xcd_atom_block(n_blocks, n_pairs, bx, p)) return;
    const int t0 = T0[p], t1 = T1[p];
    if (seg) {  // the pair's own atoms: list[seg.x .. seg.x + seg.y)
        const int2 sg = seg[p];
        list += sg.x;
        n_list = sg.y;
    }
    const size_t o0 = (size_t)t0 * (size_t)N, o1 = (size_t)t1 * (size_t)N;
    const double L0x = box[3 * t0], L0y = box[3 * t0 + 1], L0z = box[3 * t0 + 2];
    const double L1x = box[3 * t1], L1y = box[3 * t1 + 1], L1z = box[3 * t1 + 2];
    double a2 = 0.0, a4 = 0.0;
    for (int a = bx * kBlock + threadIdx.x; a < n_list; a += n_blocks * kBlock) {
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
        partial[((size_t)p * n_blocks + bx) * 2] = a2;
        partial[((size_t)p * n_blocks + bx) * 2 + 1] = a4;
    }
}

// Displacement statistics per (pair, atom block): FP64 sums of d_c and
// d_c^2 (c = x, y, z) from the FP64 rebuild d = (x1 - x0) + (i1 L1 - i0 L0),
// into partial[(p * blocks + block) * 6 + k] (k = dx, dy, dz, dx^2, dy^2,
// dz^2).  With n_bins > 0 also a histogram of |dr| as the reference forms
// it (unwrapped FP32 u = float(x + i L), FP32 difference,
// sqrtf((dx dx + dy dy) + dz dz)), counted in LDS and added to
// hist[TG[p] * n_bins + bin] (integer atomics: the counts do not depend on
// the order).  hist_mode 0 (self van Hove): bin = int(|dr| / bin_size),
// counted when below n_bins; hist_mode 1 (DebyeWaller_Dist,
// Displacement_Distribution): bin = int(d2 / bin_size) of the FP32 square
// d2 = |dr| |dr|, the last bin taking the overflow.
__global__ __launch_bounds__(kBlock) void disp_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z,
    const int16_t *__restrict__ IX, const int16_t *__restrict__ IY, const int16_t *__restrict__ IZ,
    const float *__restrict__ box, int N, const int *__restrict__ list, int n_list, const int2 *__restrict__ seg,
    const int *__restrict__ T0, const int *__restrict__ T1, const int *__restrict__ TG, int n_pairs, int n_blocks,
    int n_bins, int hist_mode, float bin_size, double *__restrict__ partial, unsigned long long *__restrict__ hist)
{
#pragma clang fp contract(off)
    extern __shared__ unsigned int sh_hist[];
    __shared__ double red[kBlock];
    int bx, p;
    if (!xcd_atom_block(n_blocks, n_pairs, bx, p)) return;
    const int t0 = T0[p], t1 = T1[p];
    if (seg) {  // the pair's own atoms: list[seg.x .. seg.x + seg.y)
        const int2 sg = seg[p];
        list += sg.x;
        n_list = sg.y;
    }
    const size_t o0 = (size_t)t0 * (size_t)N, o1 = (size_t)t1 * (size_t)N;
    const float *B0 = box + 3 * t0, *B1 = box + 3 * t1;
    for (int b = threadIdx.x; b < n_bins; b += kBlock) sh_hist[b] = 0u;
    if (n_bins > 0) __syncthreads();
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const float *src[3] = {X, Y, Z};
    const int16_t *img[3] = {IX, IY, IZ};
    for (int a = bx * kBlock + threadIdx.x; a < n_list; a += n_blocks * kBlock) {
        const int j = list ? list[a] : a;
        float u[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double x0 = src[c][o0 + j], x1 = src[c][o1 + j];
            const double w0 = (double)img[c][o0 + j] * (double)B0[c], w1 = (double)img[c][o1 + j] * (double)B1[c];
            const double d = (x1 - x0) + (w1 - w0);
            s[c] += d;
            s[3 + c] += d * d;
            u[c] = (float)(x1 + w1) - (float)(x0 + w0);
        }
        if (n_bins > 0) {
            const float len = sqrtf((u[0] * u[0] + u[1] * u[1]) + u[2] * u[2]);
            if (hist_mode == 0) {
                const int bin = (int)(len / bin_size);
                if (bin >= 0 && bin < n_bins) atomicAdd(&sh_hist[bin], 1u);
            } else {
                const float d2 = len * len;
                const int bin = min((int)(d2 / bin_size), n_bins - 1);
                if (bin >= 0) atomicAdd(&sh_hist[bin], 1u);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 6; k++) s[k] = block_sum(s[k], red);
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; k++) partial[((size_t)p * n_blocks + bx) * 6 + k] = s[k];
    if (n_bins > 0) {
        __syncthreads();
        unsigned long long *row = hist + (size_t)TG[p] * n_bins;
        for (int b = threadIdx.x; b < n_bins; b += kBlock)
            if (sh_hist[b]) atomicAdd(&row[b], (unsigned long long)sh_hist[b]);
    }
}

// Membership test of Fast_Particles, Slow_Particles and Average_Particles
// (src/fast_particles.cpp:127-139, src/slow_particles.cpp:100-113,
// src/average_particles.cpp:103-116) for atom j between frames t0 and t1
// (element offsets o0, o1; boxes B0, B1).  The squared displacement is
// formed as the reference forms it: pow(distance, 2) stored as float, with
// distance = sqrtf((dx dx + dy dy) + dz dz) of the FP32 difference of
// unwrapped FP32 coordinates u = float(x + i L).  The double square of a
// float is exact, so its rounding to float equals the FP32 product.
// mode 0: d2 > lo; 1: d2 < hi; 2: lo <= d2 <= hi.
__device__ __forceinline__ bool displacement_selected(const float *const (&src)[3], const int16_t *const (&img)[3],
                                                      size_t o0, size_t o1, const float *B0, const float *B1, int j,
                                                      int mode, float lo, float hi)
{
#pragma clang fp contract(off)
    float u[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double x0 = src[c][o0 + j], x1 = src[c][o1 + j];
        const double w0 = (double)img[c][o0 + j] * (double)B0[c], w1 = (double)img[c][o1 + j] * (double)B1[c];
        u[c] = (float)(x1 + w1) - (float)(x0 + w0);
    }
    const float len = sqrtf((u[0] * u[0] + u[1] * u[1]) + u[2] * u[2]);
    const float d2 = len * len;
    if (mode == 0) return d2 > lo;
    if (mode == 1) return d2 < hi;
    return d2 <= hi && d2 >= lo;
}

// Rectangular bins (Trajectory_List_Bins::assign_bins, src/trajectory_list_
// bins.cpp:949-1096): atom j of frame t (blockIdx.y) lies in the bin of its
// stored coordinate when strictly inside (lo, hi) of the frame on every
// axis, with index int((c - lo + 10 eps) / (hi - lo + 20 eps) * n) (FP32
// differences, the rest in double, eps = FLT_EPSILON), as the reference
// forms it; cell[t N + j] = (x ny + y) nz + z, or -1 outside.
__global__ __launch_bounds__(kBlock) void bins_assign_kernel(const float *__restrict__ X, const float *__restrict__ Y,
                                                              const float *__restrict__ Z, int N, int nx, int ny,
                                                              int nz, const float *__restrict__ lo,
                                                              const float *__restrict__ hi, int *__restrict__ cell)
{
#pragma clang fp contract(off)
    const int t = blockIdx.y;
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= N) return;
    const size_t o = (size_t)t * N + j;
    const float c[3] = {X[o], Y[o], Z[o]};
    const int n[3] = {nx, ny, nz};
    const double eps = (double)1.19209290e-07f;
    int b[3];
    bool inside = true;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float l = lo[3 * t + k], h = hi[3 * t + k];
        inside = inside && c[k] < h && c[k] > l;
        b[k] = inside ? (int)(((double)(c[k] - l) + eps * 10.0) / ((double)(h - l) + eps * 20.0) * (double)(float)n[k]) : 0;
    }
    cell[o] = inside ? (b[0] * ny + b[1]) * nz + b[2] : -1;
}

// Distance bins, point and plane forms (Trajectory_List_Bins::assign_bins_
// distance_point, src/trajectory_list_bins.cpp:1321-1488, and assign_bins_
// distance_plane, :1489-1774).  Frame t = blockIdx.y; entry a of the binned
// list at t (seg[t] = {first entry of idx, count}; idx null = all atoms in
// order); L = the frame's box lengths.  FP32 in the reference's order:
//   point p: d = r - p, d -= L * float(int(d / (L * 0.5f))) per axis,
//            distance sqrt((dx dx + dy dy) + dz dz);
//   plane on axis k at p: d = r_k - p; average |L * int(d / double(L * 0.5))
//            - d| (that quotient is a double one); above d < 0 ? L + d : d;
//            below d < 0 ? |d| : L - d.
// bin = int(distance / w), stored in cell[t N + j] when 0 <= bin < n_bins
// (cells not stored keep the caller's -1).
__global__ __launch_bounds__(kBlock) void distance_cells_kernel(const float *__restrict__ X,
                                                                 const float *__restrict__ Y,
                                                                 const float *__restrict__ Z,
                                                                 const float *__restrict__ box, int N,
                                                                 const int2 *__restrict__ seg,
                                                                 const int *__restrict__ idx, int kind, int axis,
                                                                 int dir, float px, float py, float pz, float w,
                                                                 int n_bins, int *__restrict__ cell)
{
#pragma clang fp contract(off)
    const int t = blockIdx.y;
    const int a = blockIdx.x * kBlock + threadIdx.x;
    const int2 sg = seg[t];
    if (a >= sg.y) return;
    const int j = idx ? idx[sg.x + a] : a;
    const size_t o = (size_t)t * N + j;
    const float r[3] = {X[o], Y[o], Z[o]};
    const float L[3] = {box[3 * t], box[3 * t + 1], box[3 * t + 2]};
    float len;
    if (kind == 0) {
        const float p[3] = {px, py, pz};
        float d[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            d[k] = r[k] - p[k];
            d[k] = d[k] - L[k] * (float)(int)(d[k] / (L[k] * 0.5f));
        }
        len = sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    } else {
        const float Lk = axis == 0 ? L[0] : (axis == 1 ? L[1] : L[2]);
        const float d = (axis == 0 ? r[0] : (axis == 1 ? r[1] : r[2])) - px;
        if (dir == 0) len = fabsf(Lk * (float)(int)((double)d / ((double)Lk * 0.5)) - d);
        else if (dir == 1) len = d < 0.0f ? Lk + d : d;
        else len = d < 0.0f ? fabsf(d) : Lk - d;
    }
    const int b = (int)(len / w);
    if (b >= 0 && b < n_bins) cell[o] = b;
}

// Distance bins from a second list (Trajectory_List_Bins::assign_bins_
// distance_clusters, src/trajectory_list_bins.cpp:1097-1320): entry a of the
// binned list at frame t (blockIdx.y) takes the smallest distance to the
// atoms of the cluster list at t (cseg, cidx as seg, idx), starting from the
// largest box length of the frame.  Each difference takes its shortest image
// with the box of frame 0 (Coordinate::vector_unwrapped, src/coordinate.cpp:
// 98-124: d > L/2 -> d - L, d < -L/2 -> d + L); FP32 as in the reference.
// The cluster list is staged through LDS, kBlock atoms at a time.
__global__ __launch_bounds__(kBlock) void cluster_cells_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z,
    const float *__restrict__ box, int N, const int2 *__restrict__ seg, const int *__restrict__ idx,
    const int2 *__restrict__ cseg, const int *__restrict__ cidx, float w, int n_bins, int *__restrict__ cell)
{
#pragma clang fp contract(off)
    __shared__ float sx[kBlock], sy[kBlock], sz[kBlock];
    const int t = blockIdx.y;
    const int a = blockIdx.x * kBlock + threadIdx.x;
    const int2 sg = seg[t], cs = cseg[t];
    const bool active = a < sg.y;
    const size_t ft = (size_t)t * N;
    int j = 0;
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    float len = box[3 * t];
    if (box[3 * t + 1] > len) len = box[3 * t + 1];
    if (box[3 * t + 2] > len) len = box[3 * t + 2];
    if (active) {
        j = idx ? idx[sg.x + a] : a;
        cx = X[ft + j], cy = Y[ft + j], cz = Z[ft + j];
    }
    const float bx = box[0], by = box[1], bz = box[2];
    const float hx = bx / 2.0f, hy = by / 2.0f, hz = bz / 2.0f;
    for (int base = 0; base < cs.y; base += kBlock) {
        const int c = base + threadIdx.x;
        if (c < cs.y) {
            const int k = cidx ? cidx[cs.x + c] : c;
            sx[threadIdx.x] = X[ft + k], sy[threadIdx.x] = Y[ft + k], sz[threadIdx.x] = Z[ft + k];
        }
        __syncthreads();
        const int m = min(kBlock, cs.y - base);
        if (active) {
            for (int i = 0; i < m; i++) {
                float dx = cx - sx[i], dy = cy - sy[i], dz = cz - sz[i];
                dx = dx > hx ? dx - bx : (dx < -hx ? dx + bx : dx);
                dy = dy > hy ? dy - by : (dy < -hy ? dy + by : dy);
                dz = dz > hz ? dz - bz : (dz < -hz ? dz + bz : dz);
                const float l = sqrtf((dx * dx + dy * dy) + dz * dz);
                len = l < len ? l : len;
            }
        }
        __syncthreads();
    }
    if (!active) return;
    const int b = (int)(len / w);
    if (b < n_bins) cell[ft + j] = b;
}

// Selection pass 1.  Pair q (blockIdx.y) has seg[q] = {t0, t1, first entry
// in the input list (-1: all atoms in order), entries}; one thread per
// entry.  Writes the number of selected entries of each block of kBlock
// entries to tile_count[q * n_tiles + block].
__global__ __launch_bounds__(kBlock) void select_count_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z,
    const int16_t *__restrict__ IX, const int16_t *__restrict__ IY, const int16_t *__restrict__ IZ,
    const float *__restrict__ box, int N, const int4 *__restrict__ seg, const int *__restrict__ in_idx, int mode,
    float lo, float hi, int n_tiles, int *__restrict__ tile_count)
{
    const int q = blockIdx.y;
    const int4 sg = seg[q];
    const int a = blockIdx.x * kBlock + threadIdx.x;
    const float *src[3] = {X, Y, Z};
    const int16_t *img[3] = {IX, IY, IZ};
    bool sel = false;
    if (a < sg.w) {
        const int j = sg.z >= 0 ? in_idx[sg.z + a] : a;
        sel = displacement_selected(src, img, (size_t)sg.x * N, (size_t)sg.y * N, box + 3 * sg.x, box + 3 * sg.y, j,
                                    mode, lo, hi);
    }
    const int c = __syncthreads_count(sel);
    if (threadIdx.x == 0) tile_count[(size_t)q * n_tiles + blockIdx.x] = c;
}

// Selection pass 2: the same test; selected atoms are written in input
// order to out[tile_base[q * n_tiles + block] + rank in the block].
__global__ __launch_bounds__(kBlock) void select_write_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z,
    const int16_t *__restrict__ IX, const int16_t *__restrict__ IY, const int16_t *__restrict__ IZ,
    const float *__restrict__ box, int N, const int4 *__restrict__ seg, const int *__restrict__ in_idx, int mode,
    float lo, float hi, int n_tiles, const int *__restrict__ tile_base, int *__restrict__ out)
{
    __shared__ int wsum[kBlock / 64];
    const int q = blockIdx.y;
    const int4 sg = seg[q];
    const int a = blockIdx.x * kBlock + threadIdx.x;
    const float *src[3] = {X, Y, Z};
    const int16_t *img[3] = {IX, IY, IZ};
    bool sel = false;
    int j = 0;
    if (a < sg.w) {
        j = sg.z >= 0 ? in_idx[sg.z + a] : a;
        sel = displacement_selected(src, img, (size_t)sg.x * N, (size_t)sg.y * N, box + 3 * sg.x, box + 3 * sg.y, j,
                                    mode, lo, hi);
    }
    const unsigned long long m = __ballot(sel);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int below = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    int base = tile_base[(size_t)q * n_tiles + blockIdx.x];
    for (int i = 0; i < w; i++) base += wsum[i];
    if (sel) out[base + below] = j;
}

// Per time gap: the six sums of its pairs' partials, pairs in reference order.
__global__ __launch_bounds__(kBlock) void disp_reduce_kernel(const double *__restrict__ partial, int n_blocks,
                                                              const int *__restrict__ group_off,
                                                              const int *__restrict__ order,
                                                              double *__restrict__ out)
{
    __shared__ double red[kBlock];
    const int g = blockIdx.x;
    const int p0 = group_off[g], np = group_off[g + 1] - p0;
    const long items = (long)np * n_blocks;
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (long it = threadIdx.x; it < items; it += kBlock) {
        const int pl = (int)(it / n_blocks);
        const int b = (int)(it - (long)pl * n_blocks);
        const size_t e = ((size_t)order[p0 + pl] * n_blocks + b) * 6;
        for (int k = 0; k < 6; k++) s[k] += partial[e + k];
    }
    for (int k = 0; k < 6; k++) {
        const double v = block_sum(s[k], red);
        if (threadIdx.x == 0) out[(size_t)g * 6 + k] = v;
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
// wave-uniform; sums are FP64.  blockIdx.y = s takes vectors
// [s * per_split, (s + 1) * per_split); with one split the block writes
// g = kscale * sum, else the unscaled partial sums go to
// part[(s * 3 + c) * N + j] for adjoint_reduce_kernel.
__global__ __launch_bounds__(kBlock) void sk_adjoint_kernel(
    const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z, int N, double inv_ell,
    const int *__restrict__ nvec, const float2 *__restrict__ q, int K, int per_split, double kscale,
    double *__restrict__ gx, double *__restrict__ gy, double *__restrict__ gz, double *__restrict__ part)
{
    const int j = blockIdx.x * kBlock + threadIdx.x;
    const int jj = j < N ? j : N - 1;
    const uint32_t ux = turns_fixed((double)X[jj] * inv_ell);
    const uint32_t uy = turns_fixed((double)Y[jj] * inv_ell);
    const uint32_t uz = turns_fixed((double)Z[jj] * inv_ell);
    const int v0 = blockIdx.y * per_split, v1 = min(K, v0 + per_split);
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int v = v0; v < v1; ++v) {
        const int nx = nvec[3 * v], ny = nvec[3 * v + 1], nz = nvec[3 * v + 2];
        const uint32_t P = (uint32_t)nx * ux + (uint32_t)ny * uy + (uint32_t)nz * uz;
        const float2 e = sincos_turns(P);
        const float2 qq = q[v];
        const double t = (double)fmaf(qq.x, e.y, -qq.y * e.x);
        ax = fma((double)nx, t, ax);
        ay = fma((double)ny, t, ay);
        az = fma((double)nz, t, az);
    }
    if (j >= N) return;
    if (gridDim.y == 1) {
        gx[j] = ax * kscale;
        gy[j] = ay * kscale;
        gz[j] = az * kscale;
    } else {
        const size_t o = (size_t)blockIdx.y * 3 * N + j;
        part[o] = ax;
        part[o + N] = ay;
        part[o + 2 * (size_t)N] = az;
    }
}

// g_c[j] = kscale * sum over the splits, in split order, of part[(s * 3 + c)
// * N + j]: the same result on every run.
__global__ __launch_bounds__(kBlock) void adjoint_reduce_kernel(const double *__restrict__ part, int N,
                                                                int n_splits, double kscale, double *__restrict__ gx,
                                                                double *__restrict__ gy, double *__restrict__ gz)
{
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= N) return;
    double a[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < n_splits; ++s)
        for (int c = 0; c < 3; ++c) a[c] += part[((size_t)s * 3 + c) * N + j];
    gx[j] = a[0] * kscale;
    gy[j] = a[1] * kscale;
    gz[j] = a[2] * kscale;
}

template <int R, int MM, bool GROUPED>
hipError_t launch_rho_c32_t(const TrajDev &tr, const int *list, int n_list, const int *frames, int n_frames,
                            const KPlan &kp, double inv_ell, int n_splits, double2 *out, hipStream_t st)
{
    const int per_split = (n_list + n_splits - 1) / n_splits;
    const size_t lds = Geom<MM>::tab_bytes;
    auto kfn = rho_c32_kernel<R, MM, GROUPED>;
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
hipError_t launch_fs_c32_t(const TrajDev &tr, const int *list, int n_list, const int2 *seg, const int *t0,
                           const int *t1, int n_pairs, const KPlan &kp, double inv_ell, int n_splits, double *out,
                           hipStream_t st)
{
    const int per_split = (n_list + n_splits - 1) / n_splits;
    const size_t lds = Geom<MM>::tab_bytes;
    auto kfn = fs_c32_kernel<R, MM, GROUPED>;
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
hipError_t dispatch_rho_c32_p(int MM, const TrajDev &tr, const int *list, int n_list, const int *frames,
                              int n_frames, const KPlan &kp, double inv_ell, int n_splits, double2 *out,
                              hipStream_t st)
{
    switch (MM) {
    case 24: return launch_rho_c32_t<R, 24, GROUPED>(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    case 50: return launch_rho_c32_t<R, 50, GROUPED>(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    case 100: return launch_rho_c32_t<R, 100, GROUPED>(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    case 127: return launch_rho_c32_t<R, 127, GROUPED>(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    default: return hipErrorInvalidValue;
    }
}

// Grouped kernels exist for R = 4 and 8; an ungrouped plan runs on the
// kernel without group codes (every slot carries its full vector).
template <int R>
hipError_t dispatch_rho_c32(int MM, const TrajDev &tr, const int *list, int n_list, const int *frames,
                            int n_frames, const KPlan &kp, double inv_ell, int n_splits, double2 *out,
                            hipStream_t st)
{
    if constexpr (R >= 4) {
        if (kp.pair)
            return dispatch_rho_c32_p<R, true>(MM, tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
    }
    return dispatch_rho_c32_p<R, false>(MM, tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
}

template <int R, bool GROUPED>
hipError_t dispatch_fs_c32_p(int MM, const TrajDev &tr, const int *list, int n_list, const int2 *seg,
                             const int *t0, const int *t1, int n_pairs, const KPlan &kp, double inv_ell,
                             int n_splits, double *out, hipStream_t st)
{
    switch (MM) {
    case 24: return launch_fs_c32_t<R, 24, GROUPED>(tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    case 50: return launch_fs_c32_t<R, 50, GROUPED>(tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    case 100: return launch_fs_c32_t<R, 100, GROUPED>(tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    case 127: return launch_fs_c32_t<R, 127, GROUPED>(tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    default: return hipErrorInvalidValue;
    }
}

template <int R>
hipError_t dispatch_fs_c32(int MM, const TrajDev &tr, const int *list, int n_list, const int2 *seg,
                           const int *t0, const int *t1, int n_pairs, const KPlan &kp, double inv_ell,
                           int n_splits, double *out, hipStream_t st)
{
    if constexpr (R >= 4) {
        if (kp.pair)
            return dispatch_fs_c32_p<R, true>(MM, tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    }
    return dispatch_fs_c32_p<R, false>(MM, tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
}

}  // namespace

int plan_table_width(int max_component) { return table_half_width(max_component); }

static_assert(Geom<24>::OY == table_oy(24) && Geom<24>::OZ == table_oz(24), "table geometry");
static_assert(Geom<50>::OY == table_oy(50) && Geom<50>::OZ == table_oz(50), "table geometry");
static_assert(Geom<100>::OY == table_oy(100) && Geom<100>::OZ == table_oz(100), "table geometry");
static_assert(Geom<127>::OY == table_oy(127) && Geom<127>::OZ == table_oz(127), "table geometry");

hipError_t launch_rho(int precision, const TrajDev &tr, const int *list, int n_list, const int *frames,
                      int n_frames, const KPlan &kp, double inv_ell, int n_splits, double2 *out,
                      hipStream_t st)
{
    if (precision == 3) return launch_rho_t64(tr, list, n_list, frames, n_frames, kp, inv_ell, n_splits, out, st);
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

hipError_t launch_fs(int precision, const TrajDev &tr, const int *list, int n_list, const int2 *seg, const int *t0,
                     const int *t1, int n_pairs, const KPlan &kp, double inv_ell, int n_splits, double *out,
                     hipStream_t st)
{
    if (precision == 3) return launch_fs_t64(tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    if (precision == 1) {
        const int per_split = (n_list + n_splits - 1) / n_splits;
        const double kscale = 2.0 * M_PI * inv_ell;
        dim3 grid(kp.n_chunks, n_pairs, n_splits);
        (void)hipGetLastError();
        switch (kp.R) {
        case 4: hipLaunchKernelGGL(fs_f64_kernel<4>, grid, dim3(kPhaseBlock), 0, st, tr.x, tr.y, tr.z, tr.ix, tr.iy, tr.iz, tr.box, tr.N, list, n_list, seg, t0, t1, kp.kpack, kp.kdst, kp.K, kp.KB, kscale, per_split, n_splits, out); break;
        case 8: hipLaunchKernelGGL(fs_f64_kernel<8>, grid, dim3(kPhaseBlock), 0, st, tr.x, tr.y, tr.z, tr.ix, tr.iy, tr.iz, tr.box, tr.N, list, n_list, seg, t0, t1, kp.kpack, kp.kdst, kp.K, kp.KB, kscale, per_split, n_splits, out); break;
        case 2: hipLaunchKernelGGL(fs_f64_kernel<2>, grid, dim3(kPhaseBlock), 0, st, tr.x, tr.y, tr.z, tr.ix, tr.iy, tr.iz, tr.box, tr.N, list, n_list, seg, t0, t1, kp.kpack, kp.kdst, kp.K, kp.KB, kscale, per_split, n_splits, out); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (kp.R) {
    case 4: return dispatch_fs_c32<4>(kp.MM, tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    case 8: return dispatch_fs_c32<8>(kp.MM, tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    case 2: return dispatch_fs_c32<2>(kp.MM, tr, list, n_list, seg, t0, t1, n_pairs, kp, inv_ell, n_splits, out, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_split_reduce2(const double2 *in, int n_rows, int n_splits, int K, double2 *out, hipStream_t st)
{
    if (n_rows <= 0 || K <= 0) return hipSuccess;
    const long blocks = (long)n_rows * ((K + kSplitCols - 1) / kSplitCols);
    if (blocks > 0x7fffffffL) return hipErrorInvalidValue;
    (void)hipGetLastError();

    hipLaunchKernelGGL(split_reduce2_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, st, in, n_rows, n_splits, K,
                       out);
    return hipGetLastError();
}

hipError_t launch_adjoint_weights(const double2 *rho, const int *shell_off, int n_shells, int n_atoms,
                                  const double *wt, double *res, float2 *q, hipStream_t st)
{
    (void)hipGetLastError();
    hipLaunchKernelGGL(adjoint_weights_kernel, dim3(1), dim3(kWeightsBlock), 0, st, rho, shell_off, n_shells,
                       n_atoms, wt, res, q);
    return hipGetLastError();
}

hipError_t launch_corr_shell(const double2 *rho1, const double2 *rho2, int K, const int *group_off,
                             const int *row0, const int *row1, int n_groups, const int *vbeg, const int *vend,
                             int n_shells, const double *norm, double *out, hipStream_t st)
{
    (void)hipGetLastError();

    hipLaunchKernelGGL(corr_shell_kernel, dim3(n_shells, n_groups), dim3(kBlock), 0, st, rho1, rho2, K,
                       group_off, row0, row1, vbeg, vend, n_shells, norm, out);
    return hipGetLastError();
}

hipError_t launch_rho_add(const double2 *a, const double2 *b, size_t n, double2 *out, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    (void)hipGetLastError();
    const unsigned grid = (unsigned)std::min<size_t>((n + kBlock - 1) / kBlock, 8192);
    hipLaunchKernelGGL(rho_add_kernel, dim3(grid), dim3(kBlock), 0, st, a, b, n, out);
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

hipError_t launch_msd(const TrajDev &tr, const int *list, int n_list, const int2 *seg, const int *t0, const int *t1,
                      int n_pairs, int n_blocks, double *partial, hipStream_t st)
{
    if (n_pairs <= 0 || n_blocks <= 0) return hipSuccess;
    (void)hipGetLastError();
    const long grid = (long)kXcds * ((n_blocks + kXcds - 1) / kXcds) * n_pairs;
    hipLaunchKernelGGL(msd_kernel, dim3((unsigned)grid), dim3(kBlock), 0, st, tr.x, tr.y, tr.z, tr.ix, tr.iy, tr.iz,
                       tr.box, tr.N, list, n_list, seg, t0, t1, n_pairs, n_blocks, partial);
    return hipGetLastError();
}

hipError_t launch_disp(const TrajDev &tr, const int *list, int n_list, const int2 *seg, const int *t0, const int *t1,
                       const int *tg, int n_pairs, int n_blocks, int n_bins, int hist_mode, float bin_size,
                       double *partial, unsigned long long *hist, hipStream_t st)
{
    if (n_pairs <= 0 || n_blocks <= 0) return hipSuccess;
    (void)hipGetLastError();
    const long grid = (long)kXcds * ((n_blocks + kXcds - 1) / kXcds) * n_pairs;
    hipLaunchKernelGGL(disp_kernel, dim3((unsigned)grid), dim3(kBlock), sizeof(unsigned int) * (size_t)n_bins, st,
                       tr.x, tr.y, tr.z, tr.ix, tr.iy, tr.iz, tr.box, tr.N, list, n_list, seg, t0, t1, tg, n_pairs,
                       n_blocks, n_bins, hist_mode, bin_size, partial, hist);
    return hipGetLastError();
}

hipError_t launch_bins_assign(const TrajDev &tr, int nx, int ny, int nz, const float *lo, const float *hi, int *cell,
                              hipStream_t st)
{
    (void)hipGetLastError();
    hipLaunchKernelGGL(bins_assign_kernel, dim3((tr.N + kBlock - 1) / kBlock, tr.T), dim3(kBlock), 0, st, tr.x, tr.y,
                       tr.z, tr.N, nx, ny, nz, lo, hi, cell);
    return hipGetLastError();
}

hipError_t launch_distance_cells(const TrajDev &tr, const int2 *seg, const int *idx, int max_n, int kind, int axis,
                                 int dir, const float *point, float w, int n_bins, int *cell, hipStream_t st)
{
    if (max_n <= 0) return hipSuccess;
    (void)hipGetLastError();
    hipLaunchKernelGGL(distance_cells_kernel, dim3((max_n + kBlock - 1) / kBlock, tr.T), dim3(kBlock), 0, st, tr.x,
                       tr.y, tr.z, tr.box, tr.N, seg, idx, kind, axis, dir, point[0], point[1], point[2], w, n_bins,
                       cell);
    return hipGetLastError();
}

hipError_t launch_cluster_cells(const TrajDev &tr, const int2 *seg, const int *idx, int max_n, const int2 *cseg,
                                const int *cidx, float w, int n_bins, int *cell, hipStream_t st)
{
    if (max_n <= 0) return hipSuccess;
    (void)hipGetLastError();
    hipLaunchKernelGGL(cluster_cells_kernel, dim3((max_n + kBlock - 1) / kBlock, tr.T), dim3(kBlock), 0, st, tr.x,
                       tr.y, tr.z, tr.box, tr.N, seg, idx, cseg, cidx, w, n_bins, cell);
    return hipGetLastError();
}

hipError_t launch_select(const TrajDev &tr, const int4 *seg, int n_seg, const int *in_idx, int mode, float lo,
                         float hi, int n_tiles, int *tile_count, const int *tile_base, int *out, hipStream_t st)
{
    if (n_seg <= 0 || n_tiles <= 0) return hipSuccess;
    (void)hipGetLastError();
    if (!tile_base)
        hipLaunchKernelGGL(select_count_kernel, dim3(n_tiles, n_seg), dim3(kBlock), 0, st, tr.x, tr.y, tr.z, tr.ix,
                           tr.iy, tr.iz, tr.box, tr.N, seg, in_idx, mode, lo, hi, n_tiles, tile_count);
    else
        hipLaunchKernelGGL(select_write_kernel, dim3(n_tiles, n_seg), dim3(kBlock), 0, st, tr.x, tr.y, tr.z, tr.ix,
                           tr.iy, tr.iz, tr.box, tr.N, seg, in_idx, mode, lo, hi, n_tiles, tile_base, out);
    return hipGetLastError();
}

hipError_t launch_disp_reduce(const double *partial, int n_blocks, const int *group_off, const int *order,
                              int n_groups, double *out, hipStream_t st)
{
    (void)hipGetLastError();
    hipLaunchKernelGGL(disp_reduce_kernel, dim3(n_groups), dim3(kBlock), 0, st, partial, n_blocks, group_off, order,
                       out);
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

int sk_adjoint_splits(int N, int K)
{
    // enough blocks for 8 per CU, at least 32 vectors per split
    const int blocks = (N + kBlock - 1) / kBlock;
    const int want = (2048 + blocks - 1) / blocks;
    return std::max(1, std::min(want, K / 32));
}

hipError_t launch_sk_adjoint(const float *x, const float *y, const float *z, int N, double inv_ell, const int *nvec,
                             const float2 *q, int K, int n_splits, double *part, double *gx, double *gy, double *gz,
                             hipStream_t st)
{
    const int blocks = (N + kBlock - 1) / kBlock;
    const int per_split = (K + n_splits - 1) / n_splits;
    const double kscale = 2.0 * M_PI * inv_ell;
    (void)hipGetLastError();
    hipLaunchKernelGGL(sk_adjoint_kernel, dim3(blocks, n_splits), dim3(kBlock), 0, st, x, y, z, N, inv_ell, nvec, q,
                       K, per_split, kscale, gx, gy, gz, part);
    if (n_splits > 1)
        hipLaunchKernelGGL(adjoint_reduce_kernel, dim3(blocks), dim3(kBlock), 0, st, part, N, n_splits, kscale, gx,
                           gy, gz);
    return hipGetLastError();
}

}  // namespace amdat
