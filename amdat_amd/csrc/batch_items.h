// batch_items.h -- the work items of the batched launches (S(k), Fs and
// MSD/NGP over many trajectories), shared by the per-analysis persistent
// kernels and the one-launch job kernel.
#pragma once

#include "kernels.h"
#include "phase_device.h"

namespace amdat {

namespace {

// S(k) item (trajectory b, chunk, frame group g): for each frame of the
// group, rho over all of the trajectory's atoms, and |rho_v|^2 added into
// row [g][v] of t.out (FP64; each slot is owned by one thread of one item,
// so the sum order is fixed).  All trajectories share one k-plan: the same
// integer vectors, k = 2 pi n / l_b with each trajectory's own l_b.  The
// atoms of a frame are never split over items: |rho|^2 is not additive over
// atom parts.
template <int R, int MM, bool GROUPED = false>
__device__ __forceinline__ void sk_item(const BatchTraj &t, int chunk, int g, const int *__restrict__ frames,
                                        int n_frames, int group_size, const KPlan &kp, float2 *tab, float2 *anc,
                                        uint32_t *sU)
{
    const int f_begin = g * group_size, f_end = min(n_frames, f_begin + group_size);
    const int *kd = kp.kdst + (size_t)chunk * kp.KB;
    double *row = t.out + (size_t)g * kp.K;
    for (int f = f_begin; f < f_end; ++f) {
        double re[R], im[R];
        rho_block<R, MM, GROUPED>(t.x, t.y, t.z, t.list, (size_t)frames[f] * (size_t)t.N, 0, t.n_list, t.inv_ell,
                                  kp.kpack + (size_t)chunk * kp.KB, kp.range[chunk], tab, anc, sU, re, im);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int d = kd[r * kPhaseBlock + threadIdx.x];
            if (d >= 0) {
                double *slot = row + (d & ~(1 << 30));
                const double prev = f == f_begin ? 0.0 : *slot;
                *slot = prev + (re[r] * re[r] + im[r] * im[r]);
            }
        }
    }
}

// Fs item (trajectory b, global chunk of the concatenated per-trajectory
// plans, row r): pair chunk pc = r / split holds the pairs
// order[pc_off[pc] .. pc_off[pc+1]) of one time gap (reference order), atom
// part s = r % split the atoms [s n / split, (s+1) n / split) of the list.
// Each pair's sum over the part's atoms of cos(k.dr) is added into row r of
// t.out (FP64, fixed order); the reduction sums a gap's rows in row order.
// pc_off = the gaps' pair offsets and split = 1 give one row per gap.
template <int R, int MM>
__device__ __forceinline__ void fs_item(const BatchTraj &t, int chunk, int r, int split, const int *__restrict__ T0,
                                        const int *__restrict__ T1, const int *__restrict__ pc_off,
                                        const int *__restrict__ order, const KPlan &kp, float2 *tab, float2 *anc,
                                        uint32_t *sU)
{
    const int pc = r / split, s = r - (r / split) * split;
    const int a_begin = (int)(((long long)s * t.n_list) / split);
    const int a_end = (int)(((long long)(s + 1) * t.n_list) / split);
    const int *kd = kp.kdst + (size_t)chunk * kp.KB;
    double *row = t.out + (size_t)r * t.K;
    const int p_begin = pc_off[pc], p_end = pc_off[pc + 1];
    for (int pq = p_begin; pq < p_end; ++pq) {
        // the same pair in every lane: scalar registers
        const int p = __builtin_amdgcn_readfirstlane(order[pq]);
        const int t0 = __builtin_amdgcn_readfirstlane(T0[p]), t1 = __builtin_amdgcn_readfirstlane(T1[p]);
        double sum[R];
        fs_block<R, MM>(t.x, t.y, t.z, t.ix, t.iy, t.iz, t.box + 3 * t0, t.box + 3 * t1, t.list, (size_t)t0 * t.N,
                        (size_t)t1 * t.N, a_begin, a_end, t.inv_ell, kp.kpack + (size_t)chunk * kp.KB,
                        kp.range[chunk], tab, anc, sU, sum);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int d = kd[q * kPhaseBlock + threadIdx.x];
            if (d >= 0) {
                double *slot = row + (d & ~(1 << 30));
                const double prev = pq == p_begin ? 0.0 : *slot;
                *slot = prev + sum[q];
            }
        }
    }
}

// Sum over a wavefront by lane exchanges, then over the block's waves in
// wave order (deterministic); red holds kWaves doubles.  Every thread gets
// the result.
__device__ __forceinline__ double block_sum_waves(double v, double *red)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    __syncthreads();  // red is free
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) s += red[i];
    return s;
}

// MSD item (trajectory b, time gap g): FP64 sums of |dr|^2 and |dr|^4 over
// the gap's pairs (reference order) and the trajectory's atoms, written to
// s2/s4[b*ntg + g].
__device__ __forceinline__ void msd_item(const BatchTraj &t, int b, int g, int ntg, const int *__restrict__ T0,
                                         const int *__restrict__ T1, const int *__restrict__ group_off,
                                         const int *__restrict__ order, double *s2, double *s4, double *red)
{
    double acc2 = 0.0, acc4 = 0.0;
    for (int pq = group_off[g]; pq < group_off[g + 1]; ++pq) {
        const int p = order[pq];
        const int t0 = T0[p], t1 = T1[p];
        const size_t o0 = (size_t)t0 * (size_t)t.N, o1 = (size_t)t1 * (size_t)t.N;
        const double L0x = t.box[3 * t0], L0y = t.box[3 * t0 + 1], L0z = t.box[3 * t0 + 2];
        const double L1x = t.box[3 * t1], L1y = t.box[3 * t1 + 1], L1z = t.box[3 * t1 + 2];
        double a2 = 0.0, a4 = 0.0;
        for (int a = threadIdx.x; a < t.n_list; a += kPhaseBlock) {
            const int j = t.list ? t.list[a] : a;
            const double dx = ((double)t.x[o1 + j] - (double)t.x[o0 + j]) + ((double)t.ix[o1 + j] * L1x - (double)t.ix[o0 + j] * L0x);
            const double dy = ((double)t.y[o1 + j] - (double)t.y[o0 + j]) + ((double)t.iy[o1 + j] * L1y - (double)t.iy[o0 + j] * L0y);
            const double dz = ((double)t.z[o1 + j] - (double)t.z[o0 + j]) + ((double)t.iz[o1 + j] * L1z - (double)t.iz[o0 + j] * L0z);
            const double r2 = dx * dx + dy * dy + dz * dz;
            a2 += r2;
            a4 += r2 * r2;
        }
        acc2 += block_sum_waves(a2, red);
        acc4 += block_sum_waves(a4, red);
    }
    if (threadIdx.x == 0) {
        s2[(size_t)b * ntg + g] = acc2;
        s4[(size_t)b * ntg + g] = acc4;
    }
}

__global__ __launch_bounds__(64) void fill_int_kernel(int *p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Zero the work-queue counter of a persistent launch (a kernel, not a
// memset, so that the sequence can be captured into a graph).
inline hipError_t reset_counter(int *counter, hipStream_t st)
{
    (void)hipGetLastError();
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, st, counter, 1, 0);
    return hipGetLastError();
}

}  // namespace

}  // namespace amdat
