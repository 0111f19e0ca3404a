// kernels.h -- launch interface of the device kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amdat {

constexpr int kBlock = 256;       // reduction and MSD kernels
constexpr int kPhaseBlock = 512;  // rho / Fs kernels: 8 waves share one atom tile

// Device view of a trajectory: frame-major SoA, x[t*N + j].
struct TrajDev {
    int N = 0, T = 0;
    const float *x = nullptr, *y = nullptr, *z = nullptr;
    const int16_t *ix = nullptr, *iy = nullptr, *iz = nullptr;
    const float *box = nullptr;  // device [T][3]
};

// Device k-vector plan for one shell range.  Vectors are cut into chunks
// of KB = kPhaseBlock*R in their original order, each chunk sorted by
// (nx, ny, nz) so a wave's lanes read nearby table entries, and padded
// with (0,0,0) entries whose destination is -1.  Thread t of a block owns
// positions r*kPhaseBlock + t, r < R, of its chunk.
struct KPlan {
    int K = 0;         // real vectors
    int R = 0;         // vectors per thread
    int KB = 0;        // vectors per chunk
    int n_chunks = 0;
    int MM = 0;        // table half-width the kernels were built for (max |n_c| <= MM)
    uint32_t *kpack = nullptr;  // [n_chunks*KB] (nx+128) | (ny+128)<<8 | nz<<16
    int *kdst = nullptr;        // [n_chunks*KB] destination index, -1 padding, bit 30 = conjugate
    int *chunkM = nullptr;      // [n_chunks] max |component| in the chunk
};

// Table half-widths with compiled kernel variants.
int plan_table_width(int max_component);

// rho_k(t) for frames[n_frames] (device); out rows [frame slot][split][K].
// precision: 0 table products (compensated FP32), 1 plain FP64,
// 2 direct exact-phase FP32 evaluation (compensated sums).
hipError_t launch_rho(int precision, const TrajDev &tr, const int *list, int n_list,
                      const int *frames, int n_frames, const KPlan &kp, double inv_ell,
                      int n_splits, double2 *out, hipStream_t st);
// Sum splits: out[f][k] = sum_s in[f][s][k].
hipError_t launch_split_reduce2(const double2 *in, int n_rows, int n_splits, int K, double2 *out,
                                hipStream_t st);
// Fs per (pair, vector): out[p][split][K] = sum_atoms cos(k.dr).
hipError_t launch_fs(int precision, const TrajDev &tr, const int *list, int n_list,
                     const int *t0, const int *t1, int n_pairs, const KPlan &kp, double inv_ell,
                     int n_splits, double *out, hipStream_t st);
// Per (group, shell): sum over the group's pairs and the shell's vectors of
// Re(rho1[row0] conj rho2[row1]) / H, divided by norm[group].
hipError_t launch_corr_shell(const double2 *rho1, const double2 *rho2, int K, const int *group_off,
                             const int *row0, const int *row1, int n_groups, const int *shell_off,
                             int n_shells, const double *norm, double *out, hipStream_t st);
// Per (group, shell): sum over the group's pairs (order[]) and splits of
// V[p][split][v] over the shell's vectors / H, divided by norm[group].
hipError_t launch_fs_shell(const double *V, int n_splits, int K, const int *group_off,
                           const int *order, int n_groups, const int *shell_off, int n_shells,
                           const double *norm, double *out, hipStream_t st);
// Per (pair, atom block): FP64 sums of |dr|^2 and |dr|^4.
hipError_t launch_msd(const TrajDev &tr, const int *list, int n_list, const int *t0,
                      const int *t1, int n_pairs, int n_blocks, double *partial, hipStream_t st);
// Per group: sums of the partials of its pairs.
hipError_t launch_msd_reduce(const double *partial, int n_blocks, const int *group_off,
                             const int *order, int n_groups, double *s2, double *s4,
                             hipStream_t st);

// Adjoint of the S(k) misfit: g_j = (2 pi/l) sum_v n_v Im(conj(q_v) e^{ik_v.r_j}).
hipError_t launch_sk_adjoint(const float *x, const float *y, const float *z, int N, double inv_ell, const int *nvec,
                             const float2 *q, int K, double *gx, double *gy, double *gz, hipStream_t st);

}  // namespace amdat
