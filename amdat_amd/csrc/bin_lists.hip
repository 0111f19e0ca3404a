// bin_lists.hip -- per-bin atom lists formed on the device from a cell
// table [T][N] (rectangular or distance bins): for every row (a frame, or a
// frame pair of the persistence form) the atoms of a list that lie in each
// bin, in atom order.  Two passes over the table: per-(bin, row) counts
// (LDS histograms), then, after the host has turned the counts into
// offsets, one wavefront per row writes the atoms in ascending index order
// (ballot ranks per bin, a running cursor per bin in LDS).
#include <algorithm>

#include "kernels.h"

namespace amdat {

namespace {

// The atom j of row `row` is a member: the list's mask row (-2: every atom,
// -1: none), then a cell in the window [w0, w0 + nw) at frame row.x and,
// when row.y >= 0, the same cell at frame row.y.  Returns the window bin or -1.
__device__ __forceinline__ int row_bin(const int *__restrict__ cell, size_t N, int4 row,
                                       const uint8_t *__restrict__ mask, int j, int w0, int nw)
{
    if (row.z == -1) return -1;
    if (row.z >= 0 && !mask[(size_t)row.z * N + j]) return -1;
    const int b = cell[(size_t)row.x * N + j];
    if (b < 0) return -1;
    if (row.y >= 0 && cell[(size_t)row.y * N + j] != b) return -1;
    const int w = b - w0;
    return (w >= 0 && w < nw) ? w : -1;
}

// mask[c * N + idx[e]] = 1 for the entries e of list time c
// (off[c] <= e < off[c + 1]; a static list has one list time).
__global__ __launch_bounds__(kBlock) void list_mask_kernel(const int *__restrict__ idx, const int *__restrict__ off,
                                                           int n_times, int n, int N, uint8_t *__restrict__ mask)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= n) return;
    int lo = 0, hi = n_times;  // off[lo] <= e < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= e) lo = mid;
        else hi = mid;
    }
    mask[(size_t)lo * N + idx[e]] = 1;
}

// grid (atom blocks of 1024, rows): cnt[w * n_rows + r] += atoms of row r in
// window bin w.
__global__ __launch_bounds__(kBlock) void bin_count_kernel(const int *__restrict__ cell, int N,
                                                           const int4 *__restrict__ rows,
                                                           const uint8_t *__restrict__ mask, int w0, int nw,
                                                           int n_rows, int *__restrict__ cnt)
{
    extern __shared__ int hist[];
    const int r = blockIdx.y;
    const int4 row = rows[r];
    for (int b = threadIdx.x; b < nw; b += kBlock) hist[b] = 0;
    __syncthreads();
    const int j0 = blockIdx.x * 1024, j1 = min(N, j0 + 1024);
    for (int j = j0 + threadIdx.x; j < j1; j += kBlock) {
        const int w = row_bin(cell, (size_t)N, row, mask, j, w0, nw);
        if (w >= 0) atomicAdd(&hist[w], 1);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nw; b += kBlock)
        if (hist[b]) atomicAdd(&cnt[(size_t)b * n_rows + r], hist[b]);
}

// One wavefront per row (four per block): the row's atoms, ascending, to
// dst[w][pos[w * n_rows + r] + rank].
__global__ __launch_bounds__(kBlock) void bin_write_kernel(const int *__restrict__ cell, int N,
                                                           const int4 *__restrict__ rows,
                                                           const uint8_t *__restrict__ mask, int w0, int nw,
                                                           int n_rows, const int *__restrict__ pos,
                                                           int *const *__restrict__ dst)
{
    extern __shared__ int cursor[];  // [4][nw]
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + wave;
    int *cur = cursor + wave * nw;
    if (r < n_rows)
        for (int b = lane; b < nw; b += 64) cur[b] = pos[(size_t)b * n_rows + r];
    __syncthreads();
    if (r >= n_rows) return;
    const int4 row = rows[r];
    if (row.z == -1) return;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int j0 = 0; j0 < N; j0 += 64) {
        const int j = j0 + lane;
        const int w = j < N ? row_bin(cell, (size_t)N, row, mask, j, w0, nw) : -1;
        unsigned long long todo = __ballot(w >= 0);
        while (todo) {
            const int lead = __ffsll((long long)todo) - 1;
            const int ws = __shfl(w, lead, 64);
            const unsigned long long m = __ballot(w == ws);
            const int base = cur[ws];
            if (w == ws) dst[ws][base + __popcll(m & below)] = j;
            if (lane == lead) cur[ws] = base + __popcll(m);
            todo &= ~m;
        }
    }
}

}  // namespace

hipError_t launch_list_mask(const int *idx, const int *off, int n_times, int n, int N, uint8_t *mask, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    (void)hipGetLastError();
    hipLaunchKernelGGL(list_mask_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, idx, off, n_times, n,
                       N, mask);
    return hipGetLastError();
}

hipError_t launch_bin_count(const int *cell, int N, const int4 *rows, int n_rows, const uint8_t *mask, int w0,
                            int nw, int *cnt, hipStream_t st)
{
    if (n_rows <= 0 || N <= 0) return hipSuccess;
    (void)hipGetLastError();
    hipLaunchKernelGGL(bin_count_kernel, dim3((N + 1023) / 1024, n_rows), dim3(kBlock), sizeof(int) * nw, st, cell,
                       N, rows, mask, w0, nw, n_rows, cnt);
    return hipGetLastError();
}

hipError_t launch_bin_write(const int *cell, int N, const int4 *rows, int n_rows, const uint8_t *mask, int w0,
                            int nw, const int *pos, int *const *dst, hipStream_t st)
{
    if (n_rows <= 0 || N <= 0) return hipSuccess;
    (void)hipGetLastError();
    hipLaunchKernelGGL(bin_write_kernel, dim3((n_rows + 3) / 4), dim3(kBlock), sizeof(int) * 4 * nw, st, cell, N,
                       rows, mask, w0, nw, n_rows, pos, dst);
    return hipGetLastError();
}

}  // namespace amdat
