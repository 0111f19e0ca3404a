// batch_kernels.hip -- persistent launches over many trajectories: the
// S(k), Fs and MSD/NGP work items of a batch, drawn from one device queue.
#include <algorithm>

#include "kernels.h"
#include "phase_device.h"

namespace amdat {

namespace {

// S(k) item (trajectory b, chunk, frame group g): for each frame of the
// group, rho over all of the trajectory's atoms, and |rho_v|^2 added into
// row [g][v] of t.out (FP64; each slot is owned by one thread of one item,
// so the sum order is fixed).  All trajectories share one k-plan: the same
// integer vectors, k = 2 pi n / l_b with each trajectory's own l_b.
template <int R, int MM, bool GROUPED = false>
__device__ __forceinline__ void sk_item(const BatchTraj &t, int chunk, int g, const int *__restrict__ frames,
                                        int n_frames, int group_size, const KPlan &kp, float2 *tab, float2 *anc,
                                        uint32_t *sU, int *sValid)
{
    const int f_begin = g * group_size, f_end = min(n_frames, f_begin + group_size);
    const int *kd = kp.kdst + (size_t)chunk * kp.KB;
    double *row = t.out + (size_t)g * kp.K;
    for (int f = f_begin; f < f_end; ++f) {
        float hr[R], lr[R], hi[R], li[R];
        rho_block<R, MM, GROUPED>(t.x, t.y, t.z, t.list, (size_t)frames[f] * (size_t)t.N, 0, t.n_list, t.inv_ell,
                         kp.kpack + (size_t)chunk * kp.KB, kp.range[chunk], tab, anc, sU, sValid, hr, lr, hi, li);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int d = kd[r * kPhaseBlock + threadIdx.x];
            if (d >= 0) {
                const double re = (double)hr[r] + (double)lr[r];
                const double im = (double)hi[r] + (double)li[r];
                double *slot = row + (d & ~(1 << 30));
                const double prev = f == f_begin ? 0.0 : *slot;
                *slot = prev + (re * re + im * im);
            }
        }
    }
}

// Fs item (trajectory b, global chunk of the concatenated per-trajectory
// plans, time gap g): the gap's pairs in reference order, each pair's sum
// over atoms of cos(k.dr) added into row [g][v] of t.out (FP64, fixed order).
template <int R, int MM>
__device__ __forceinline__ void fs_item(const BatchTraj &t, int chunk, int g, const int *__restrict__ T0,
                                        const int *__restrict__ T1, const int *__restrict__ group_off,
                                        const int *__restrict__ order, const KPlan &kp, float2 *tab, float2 *anc,
                                        uint32_t *sU, int *sValid)
{
    const int *kd = kp.kdst + (size_t)chunk * kp.KB;
    double *row = t.out + (size_t)g * t.K;
    const int p_begin = group_off[g], p_end = group_off[g + 1];
    for (int pq = p_begin; pq < p_end; ++pq) {
        const int p = order[pq];
        const int t0 = T0[p], t1 = T1[p];
        float hs[R], ls[R];
        fs_block<R, MM>(t.x, t.y, t.z, t.ix, t.iy, t.iz, t.box + 3 * t0, t.box + 3 * t1, t.list, (size_t)t0 * t.N,
                        (size_t)t1 * t.N, 0, t.n_list, t.inv_ell, kp.kpack + (size_t)chunk * kp.KB,
                        kp.range[chunk], tab, anc, sU, sValid, hs, ls);
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

// Batched S(k), persistent: items (trajectory, chunk, frame group, -).
template <int R, int MM>
__global__ __launch_bounds__(kPhaseBlock, 4) void sk_batch_kernel(const BatchTraj *__restrict__ trajs,
                                                                 const int4 *__restrict__ items, int n_items,
                                                                 int *counter, const int *__restrict__ frames,
                                                                 int n_frames, int group_size, KPlan kp)
{
    AMDAT_PHASE_SHARED(MM);
    __shared__ int s_item;
    for (;;) {
        const int q = next_item(counter, &s_item);
        if (q >= n_items) break;
        const int4 it = items[q];
        sk_item<R, MM>(trajs[it.x], it.y, it.z, frames, n_frames, group_size, kp, tab, anc, sU, sValid);
    }
}

// Batched Fs, persistent: items (trajectory, global chunk, time gap, -).
template <int R, int MM>
__global__ __launch_bounds__(kPhaseBlock, 4) void fs_batch_kernel(const BatchTraj *__restrict__ trajs,
                                                                 const int4 *__restrict__ items, int n_items,
                                                                 int *counter, const int *__restrict__ T0,
                                                                 const int *__restrict__ T1,
                                                                 const int *__restrict__ group_off,
                                                                 const int *__restrict__ order, KPlan kp)
{
    AMDAT_PHASE_SHARED(MM);
    __shared__ int s_item;
    for (;;) {
        const int q = next_item(counter, &s_item);
        if (q >= n_items) break;
        const int4 it = items[q];
        fs_item<R, MM>(trajs[it.x], it.y, it.z, T0, T1, group_off, order, kp, tab, anc, sU, sValid);
    }
}

// One analysis job over a batch in a single persistent launch: S(k), Fs and
// MSD/NGP items (item.w = 0, 1, 2) drawn from one queue, costliest first.
// PS: the S(k) plan carries group codes.
template <int RS, int RF, int MM, bool PS>
__global__ __launch_bounds__(kPhaseBlock, 4) void job_kernel(BatchJob j)
{
    AMDAT_PHASE_SHARED(MM);
    __shared__ int s_item;
    __shared__ double red[kWaves];
    for (;;) {
        const int q = next_item(j.counter, &s_item);
        if (q >= j.n_items) break;
        const int4 it = j.items[q];
        if (it.w == 0)
            sk_item<RS, MM, PS>(j.sk_trajs[it.x], it.y, it.z, j.frames, j.n_frames, j.group_size, j.sk, tab, anc,
                                sU, sValid);
        else if (it.w == 1)
            fs_item<RF, MM>(j.fs_trajs[it.x], it.y, it.z, j.t0, j.t1, j.group_off, j.order, j.fs, tab, anc, sU,
                            sValid);
        else
            msd_item(j.msd_trajs[it.x], it.x, it.y, j.ntg, j.t0, j.t1, j.group_off, j.order, j.s2, j.s4, red);
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

}  // namespace

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
                       group_size, kp);
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
                       order, kp);
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

template <int RS, int RF, bool PS>
const void *job_fn_mm(int MM)
{
    switch (MM) {
    case 24: return (const void *)job_kernel<RS, RF, 24, PS>;
    case 50: return (const void *)job_kernel<RS, RF, 50, PS>;
    case 100: return (const void *)job_kernel<RS, RF, 100, PS>;
    case 127: return (const void *)job_kernel<RS, RF, 127, PS>;
    default: return nullptr;
    }
}

template <int RS, bool PS>
const void *job_fn_rf(int RF, int MM)
{
    switch (RF) {
    case 2: return job_fn_mm<RS, 2, PS>(MM);
    case 4: return job_fn_mm<RS, 4, PS>(MM);
    case 8: return job_fn_mm<RS, 8, PS>(MM);
    default: return nullptr;
    }
}

size_t tab_bytes_of(int MM)
{
    switch (MM) {
    case 24: return Geom<24>::tab_bytes;
    case 50: return Geom<50>::tab_bytes;
    case 100: return Geom<100>::tab_bytes;
    case 127: return Geom<127>::tab_bytes;
    default: return 0;
    }
}

}  // namespace

hipError_t prepare_job(int RS, int RF, int MM, int PS, int n_items, JobLaunch *out)
{
    const void *fn = nullptr;
    switch (RS) {
    case 2: fn = PS ? nullptr : job_fn_rf<2, false>(RF, MM); break;
    case 4: fn = PS ? nullptr : job_fn_rf<4, false>(RF, MM); break;
    case 8: fn = PS ? job_fn_rf<8, true>(RF, MM) : job_fn_rf<8, false>(RF, MM); break;
    default: break;
    }
    if (!fn) return hipErrorInvalidValue;
    const size_t lds = tab_bytes_of(MM);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int dev = 0, n_cu = 0, per_cu = 0;
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kPhaseBlock, lds);
    if (e != hipSuccess) return e;
    out->fn = fn;
    out->lds = lds;
    out->grid = std::max(1, std::min(n_items, std::max(1, per_cu) * n_cu));
    return hipSuccess;
}

hipError_t enqueue_job(const JobLaunch &L, const BatchJob &j, hipStream_t st)
{
    if (j.n_items <= 0) return hipSuccess;
    hipError_t e = reset_counter(j.counter, st);
    if (e != hipSuccess) return e;
    BatchJob arg = j;
    void *args[] = {&arg};
    return hipLaunchKernel(L.fn, dim3(L.grid), dim3(kPhaseBlock), args, L.lds, st);
}

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
