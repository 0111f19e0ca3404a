// batch_kernels.hip -- persistent launches over many trajectories: the
// S(k), Fs and MSD/NGP work items of a batch, drawn from one device queue.
#include <algorithm>

#include "batch_items.h"
#include "kernels.h"
#include "phase_device.h"

namespace amdat {

namespace {

// Batched S(k), persistent: items (trajectory, chunk, frame group, -).
template <int R, int MM, bool GROUPED>
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
        const int4 it = uniform_item(items, q);
        sk_item<R, MM, GROUPED>(trajs[it.x], it.y, it.z, frames, n_frames, group_size, kp, tab, anc, sU);
    }
}

// Batched Fs, persistent: items (trajectory, global chunk, time gap, -);
// one row per gap.
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
        const int4 it = uniform_item(items, q);
        fs_item<R, MM>(trajs[it.x], it.y, it.z, 1, T0, T1, group_off, order, kp, tab, anc, sU);
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
// shells: out[out_off[b] + g*nsh[b] + s] = sum over the gap's rows
// r in [row_off[g], row_off[g+1]) (row_off null: r = g) and v in s of
// V_b[r][v], / H / norm[b*ntg + g].
__global__ __launch_bounds__(kBlock) void fs_batch_reduce_kernel(
    const BatchTraj *__restrict__ trajs, const int *__restrict__ shell_off, const int *__restrict__ shell_base,
    const int *__restrict__ nsh, int ntg, const int *__restrict__ row_off, const double *__restrict__ norm,
    const int *__restrict__ out_off, double *__restrict__ out)
{
    __shared__ double red[kBlock];
    const int s = blockIdx.x, g = blockIdx.y, b = blockIdx.z;
    if (s >= nsh[b]) return;  // uniform across the block
    const BatchTraj &t = trajs[b];
    const int *so = shell_off + shell_base[b];
    const int v0 = so[s], H = so[s + 1] - v0;
    const double nrm = norm[(size_t)b * ntg + g];
    const int r0 = row_off ? row_off[g] : g, r1 = row_off ? row_off[g + 1] : g + 1;
    double acc = 0.0;
    if (nrm != 0.0)
        for (int r = r0; r < r1; ++r)
            for (int v = threadIdx.x; v < H; v += kBlock) acc += t.out[(size_t)r * t.K + v0 + v];
    acc = block_sum(acc, red);
    if (threadIdx.x == 0)
        out[out_off[b] + (size_t)g * nsh[b] + s] = (H > 0 && nrm != 0.0) ? acc / (double)H / nrm : 0.0;
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

template <int R, int MM, bool GROUPED>
hipError_t launch_sk_batch_g(const BatchTraj *trajs, const int4 *items, int n_items, int *counter, const int *frames,
                             int n_frames, int group_size, const KPlan &kp, hipStream_t st)
{
    const size_t lds = Geom<MM>::tab_bytes;
    auto kfn = sk_batch_kernel<R, MM, GROUPED>;
    hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int grid = 0;
    if (e == hipSuccess) e = persistent_grid(kfn, lds, n_items, &grid);
    if (e == hipSuccess) e = reset_counter(counter, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kfn, dim3(grid), dim3(kPhaseBlock), lds, st, trajs, items, n_items, counter, frames, n_frames,
                       group_size, kp);
    return hipGetLastError();
}

// grouped plans (kp.pair) exist for R = 4 and 8
template <int R, int MM>
hipError_t launch_sk_batch_t(const BatchTraj *trajs, const int4 *items, int n_items, int *counter, const int *frames,
                             int n_frames, int group_size, const KPlan &kp, hipStream_t st)
{
    if constexpr (R >= 4) {
        if (kp.pair)
            return launch_sk_batch_g<R, MM, true>(trajs, items, n_items, counter, frames, n_frames, group_size, kp,
                                                  st);
    } else {
        if (kp.pair) return hipErrorInvalidValue;
    }
    return launch_sk_batch_g<R, MM, false>(trajs, items, n_items, counter, frames, n_frames, group_size, kp, st);
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
                                  const int *nsh, int max_nsh, int ntg, const int *row_off, const double *norm,
                                  const int *out_off, double *out, hipStream_t st)
{
    (void)hipGetLastError();
    hipLaunchKernelGGL(fs_batch_reduce_kernel, dim3(max_nsh, ntg, n_traj), dim3(kBlock), 0, st, trajs, shell_off,
                       shell_base, nsh, ntg, row_off, norm, out_off, out);
    return hipGetLastError();
}

}  // namespace amdat
