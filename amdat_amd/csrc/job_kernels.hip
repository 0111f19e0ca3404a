// job_kernels.hip -- one analysis job over a batch of trajectories in a
// single persistent launch: S(k), Fs and MSD/NGP items drawn from one
// device queue, costliest first.
#include <algorithm>

#include "batch_items.h"
#include "kernels.h"
#include "phase_device.h"

namespace amdat {

namespace {

// item.w = 0: S(k) (trajectory, chunk, frame group); 1: Fs (trajectory,
// global chunk, pair chunk * fs_split + atom part); 2: MSD (trajectory, gap).
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
        const int4 it = uniform_item(j.items, q);
        if (it.w == 0)
            sk_item<RS, MM, PS>(j.sk_trajs[it.x], it.y, it.z, j.frames, j.n_frames, j.group_size, j.sk, tab, anc,
                                sU);
        else if (it.w == 1)
            fs_item<RF, MM>(j.fs_trajs[it.x], it.y, it.z, j.fs_split, j.t0, j.t1, j.fs_pc, j.order, j.fs, tab, anc,
                            sU);
        else
            msd_item(j.msd_trajs[it.x], it.x, it.y, j.ntg, j.t0, j.t1, j.group_off, j.order, j.s2, j.s4, red);
    }
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

int table_atoms_per_tile(int MM)
{
    switch (MM) {
    case 24: return Geom<24>::A;
    case 50: return Geom<50>::A;
    case 100: return Geom<100>::A;
    case 127: return Geom<127>::A;
    default: return 0;
    }
}

hipError_t prepare_job(int RS, int RF, int MM, bool PS, int n_items, JobLaunch *out)
{
    const void *fn = nullptr;
    switch (RS) {
    case 2: fn = PS ? nullptr : job_fn_rf<2, false>(RF, MM); break;
    case 4: fn = PS ? job_fn_rf<4, true>(RF, MM) : job_fn_rf<4, false>(RF, MM); break;
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

}  // namespace amdat
