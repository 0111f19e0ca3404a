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

// Device k-vector plan for one shell range (layout built on the host by
// plan_layout): chunks of KB = kPhaseBlock*R slots; thread t of a block
// owns positions r*kPhaseBlock + t, r < R, of its chunk, all with one nx;
// unused slots have destination -1.
struct KPlan {
    int K = 0;         // real vectors
    int R = 0;         // vectors per thread
    int KB = 0;        // vectors per chunk
    int n_chunks = 0;
    int MM = 0;        // table half-width the kernels were built for (max |n_c| <= MM)
    uint32_t *kpack = nullptr;  // [n_chunks*KB] (nx+128) | (ny+128)<<8 | nz<<16
    int *kdst = nullptr;        // [n_chunks*KB] destination index, -1 padding, bit 30 = conjugate
    int4 *range = nullptr;      // [n_chunks] (nx lo, nx hi, max|ny| | max nz << 16,
                                //  units with vectors | group codes << 16, 2 bits per wave)
    int pair = 0;               // 1: grouped plan, wave w's GS = 1 << code(w) slots share e_x e_y
};

// One trajectory of a batch (device-side table entry).
struct BatchTraj {
    const float *x, *y, *z;
    const int16_t *ix, *iy, *iz;
    const float *box;   // device [T][3]
    const int *list;    // nullptr = all atoms in order
    int N, n_list;      // atoms per frame, atoms selected
    int K;              // vectors per output row
    double inv_ell;     // 1 / length scale
    double *out;        // output rows [group][K]
};

// One analysis job over a batch (device view): S(k), Fs and MSD/NGP items in
// one queue.  item = (trajectory, chunk, row, kind): kind 0 = S(k), row =
// frame group; kind 1 = Fs, row = pair chunk * fs_split + atom part (pair
// chunk c holds the pairs order[fs_pc[c] .. fs_pc[c+1]) of one gap); MSD
// items are (trajectory, gap, -, 2).
struct BatchJob {
    const int4 *items = nullptr;
    int n_items = 0;
    int *counter = nullptr;
    const BatchTraj *sk_trajs = nullptr, *fs_trajs = nullptr, *msd_trajs = nullptr;
    const int *frames = nullptr;
    int n_frames = 0, group_size = 1;
    KPlan sk, fs;  // fs: per-trajectory plans concatenated
    const int *t0 = nullptr, *t1 = nullptr, *group_off = nullptr, *order = nullptr;
    const int *fs_pc = nullptr;  // pair-chunk offsets into order
    int fs_split = 1;            // atom parts per Fs pair chunk
    int ntg = 0;
    double *s2 = nullptr, *s4 = nullptr;  // [trajectory][gap]
};

// Launch shape of a job kernel, fixed before any stream capture.
struct JobLaunch {
    const void *fn = nullptr;
    int grid = 0;
    size_t lds = 0;
};

// Table half-widths with compiled kernel variants.
int plan_table_width(int max_component);

// rho_k(t) for frames[n_frames] (device); out rows [frame slot][split][K].
// precision: 0 table products (compensated FP32), 1 plain FP64,
// 2 direct exact-phase FP32 evaluation (compensated sums), 3 table
// products in FP64 (the plan and tables of 0, FP64 entries and sums).
hipError_t launch_rho(int precision, const TrajDev &tr, const int *list, int n_list,
                      const int *frames, int n_frames, const KPlan &kp, double inv_ell,
                      int n_splits, double2 *out, hipStream_t st);
// Per-bin lists from a cell table (bin_lists.hip).  Rows: int4 (frame of
// the cell, second frame whose cell must match or -1, list mask row: -2
// every atom, -1 none).  Bins are processed in windows [w0, w0 + nw) of at
// most kBinWindow bins.
constexpr int kBinWindow = 2048;
// mask[c * N + idx[e]] = 1 for the entries of list time c (off on the device).
hipError_t launch_list_mask(const int *idx, const int *off, int n_times, int n, int N, uint8_t *mask, hipStream_t st);
// cnt[w * n_rows + r] += the row's atoms in window bin w (cnt zeroed by the caller).
hipError_t launch_bin_count(const int *cell, int N, const int4 *rows, int n_rows, const uint8_t *mask, int w0,
                            int nw, int *cnt, hipStream_t st);
// The row's atoms in window bin w, ascending, to dst[w][pos[w * n_rows + r] ...].
hipError_t launch_bin_write(const int *cell, int N, const int4 *rows, int n_rows, const uint8_t *mask, int w0,
                            int nw, const int *pos, int *const *dst, hipStream_t st);
// Tabled-FP64 rho and Fs (precision 3 of launch_rho and launch_fs).
hipError_t launch_rho_t64(const TrajDev &tr, const int *list, int n_list, const int *frames, int n_frames,
                          const KPlan &kp, double inv_ell, int n_splits, double2 *out, hipStream_t st);
hipError_t launch_fs_t64(const TrajDev &tr, const int *list, int n_list, const int2 *seg, const int *t0,
                         const int *t1, int n_pairs, const KPlan &kp, double inv_ell, int n_splits, double *out,
                         hipStream_t st);
// Sum splits: out[f][k] = sum_s in[f][s][k].
hipError_t launch_split_reduce2(const double2 *in, int n_rows, int n_splits, int K, double2 *out,
                                hipStream_t st);
// Fs per (pair, vector): out[p][split][K] = sum_atoms cos(k.dr).  The pair
// kernels (launch_fs, launch_msd, launch_disp) take every pair's atoms from
// list[0 .. n_list) (list null = all atoms in order), or, with seg non-null,
// pair p's own atoms list[seg[p].x .. seg[p].x + seg[p].y) (n_list is then
// the largest count, which sets the atom splits).
hipError_t launch_fs(int precision, const TrajDev &tr, const int *list, int n_list, const int2 *seg,
                     const int *t0, const int *t1, int n_pairs, const KPlan &kp, double inv_ell,
                     int n_splits, double *out, hipStream_t st);
// Per (group, column s): sum over the group's pairs and the column's vectors
// v in [vbeg[s], vend[s]) of Re(rho1[row0] conj rho2[row1]) / H (H =
// vend[s] - vbeg[s]), divided by norm[group].  A shell table serves as
// vbeg = shell_off, vend = shell_off + 1.
hipError_t launch_corr_shell(const double2 *rho1, const double2 *rho2, int K, const int *group_off,
                             const int *row0, const int *row1, int n_groups, const int *vbeg, const int *vend,
                             int n_shells, const double *norm, double *out, hipStream_t st);
// out[i] = a[i] + b[i], i < n.
hipError_t launch_rho_add(const double2 *a, const double2 *b, size_t n, double2 *out, hipStream_t st);
// Per (group, shell): sum over the group's pairs (order[]) and splits of
// V[p][split][v] over the shell's vectors / H, divided by norm[group].
hipError_t launch_fs_shell(const double *V, int n_splits, int K, const int *group_off,
                           const int *order, int n_groups, const int *shell_off, int n_shells,
                           const double *norm, double *out, hipStream_t st);
// Per (pair, atom block): FP64 sums of |dr|^2 and |dr|^4.
hipError_t launch_msd(const TrajDev &tr, const int *list, int n_list, const int2 *seg, const int *t0,
                      const int *t1, int n_pairs, int n_blocks, double *partial, hipStream_t st);
// Per (pair, atom block): FP64 sums of d_c and d_c^2 into partial[(p*n_blocks
// + block)*6 + k]; with n_bins > 0 the reference-form counts are added to
// hist[tg[p]*n_bins + bin]: hist_mode 0, bin = int(|dr| / bin_size) below
// n_bins; hist_mode 1, bin = int(|dr|^2 / bin_size), the last bin taking
// the overflow.
hipError_t launch_disp(const TrajDev &tr, const int *list, int n_list, const int2 *seg, const int *t0, const int *t1,
                       const int *tg, int n_pairs, int n_blocks, int n_bins, int hist_mode, float bin_size,
                       double *partial, unsigned long long *hist, hipStream_t st);
// Per group: the six sums of its pairs' partials (pairs in order[]).
// Rectangular bin of every (frame, atom) into cell[T][N] (-1 outside).
hipError_t launch_bins_assign(const TrajDev &tr, int nx, int ny, int nz, const float *lo, const float *hi, int *cell,
                              hipStream_t st);
// Distance bins: point (kind 0; point = {x, y, z}) or plane (kind 1; axis,
// dir 0 average / 1 above / 2 below, point[0] = position); seg[t] = {first
// entry of idx, count} of the binned list at frame t (idx null = all atoms).
hipError_t launch_distance_cells(const TrajDev &tr, const int2 *seg, const int *idx, int max_n, int kind, int axis,
                                 int dir, const float *point, float w, int n_bins, int *cell, hipStream_t st);
// Distance bins from the atoms of a second list (cseg, cidx as seg, idx).
hipError_t launch_cluster_cells(const TrajDev &tr, const int2 *seg, const int *idx, int max_n, const int2 *cseg,
                                const int *cidx, float w, int n_bins, int *cell, hipStream_t st);
// Displacement selection (Fast/Slow/Average_Particles): with tile_base ==
// nullptr, per-block counts into tile_count[n_seg][n_tiles]; otherwise the
// selected atoms, in input order, into out at tile_base[q][tile] + rank.
hipError_t launch_select(const TrajDev &tr, const int4 *seg, int n_seg, const int *in_idx, int mode, float lo,
                         float hi, int n_tiles, int *tile_count, const int *tile_base, int *out, hipStream_t st);
hipError_t launch_disp_reduce(const double *partial, int n_blocks, const int *group_off, const int *order,
                              int n_groups, double *out, hipStream_t st);
// Per group: sums of the partials of its pairs.
hipError_t launch_msd_reduce(const double *partial, int n_blocks, const int *group_off,
                             const int *order, int n_groups, double *s2, double *s4,
                             hipStream_t st);

// Batched launches over work items of many trajectories.  The phase
// kernels are persistent: as many blocks as stay resident draw items from
// a queue (counter: one device int, reset on the stream before the launch).
// Items are processed in queue order, so callers list the costly ones first.
//
// S(k) items (trajectory, chunk, frame group, -): rows [group][K] of
// sum_{f in group} |rho_v(f)|^2; frames of group g are frames[g*group_size ..].
hipError_t launch_sk_batch(const BatchTraj *trajs, const int4 *items, int n_items, int *counter, const int *frames,
                           int n_frames, int group_size, const KPlan &kp, hipStream_t st);
// Fs items (trajectory, global chunk, time gap, -) over concatenated
// per-trajectory plans (common R and MM): rows [gap][K_b] of
// sum over the gap's pairs of sum_atoms cos(k.dr).
hipError_t launch_fs_batch(const BatchTraj *trajs, const int4 *items, int n_items, int *counter, const int *t0,
                           const int *t1, const int *group_off, const int *order, const KPlan &kp, hipStream_t st);
// Job kernel for R = RS (S(k) items), RF (Fs items), table width MM:
// function attributes and the persistent grid for n_items.  Call outside
// stream capture; enqueue_job (counter reset + launch) may be captured.
// PS: the S(k) plan carries group codes (RS = 4 or 8).
hipError_t prepare_job(int RS, int RF, int MM, bool PS, int n_items, JobLaunch *out);
hipError_t enqueue_job(const JobLaunch &L, const BatchJob &j, hipStream_t st);
// Atoms per tile of the table kernels built for table half-width MM (fixed
// at compile time by the LDS budget of two 512-thread blocks per CU).
int table_atoms_per_tile(int MM);
// MSD items (trajectory b, pair p): partial[2*(b*n_pairs + p)] = sum r^2, [+1] = sum r^4.
hipError_t launch_msd_batch(const BatchTraj *trajs, const int2 *items, int n_items, const int *t0, const int *t1,
                            int n_pairs, double *partial, hipStream_t st);
// Per (gap, trajectory): sums of partial rows [trajectory][pair] over the gap's pairs.
hipError_t launch_msd_batch_reduce(const double *partial, int n_pairs, const int *group_off, const int *order,
                                   int ntg, int n_traj, double *s2, double *s4, hipStream_t st);
// S[b*n_shells + s] = sum_g sum_{v in s} rows_b[g][v] / H / norm[b].
hipError_t launch_sk_batch_reduce(const BatchTraj *trajs, int n_traj, int n_groups, int K, const int *shell_off,
                                  int n_shells, const double *norm, double *S, hipStream_t st);
// Fs per (trajectory, time gap, shell) over ragged per-trajectory shells:
// out[out_off[b] + g*nsh[b] + s] = sum_{r in gap g's rows} sum_{v in s}
// rows_b[r][v] / H / norm[b*ntg + g]; gap g's rows are [row_off[g],
// row_off[g+1]) (row_off null: the single row g).
hipError_t launch_fs_batch_reduce(const BatchTraj *trajs, int n_traj, const int *shell_off, const int *shell_base,
                                  const int *nsh, int max_nsh, int ntg, const int *row_off, const double *norm,
                                  const int *out_off, double *out, hipStream_t st);

// Adjoint of the S(k) misfit: g_j = (2 pi/l) sum_v n_v Im(conj(q_v) e^{ik_v.r_j}),
// the K vectors in n_splits ranges (sk_adjoint_splits); with more than one,
// part holds n_splits * 3 * N doubles of partial sums.
int sk_adjoint_splits(int N, int K);
// Adjoint weights from the forward sums of one frame: res[0] = loss,
// res[1 .. n_shells] = S per shell (res holds 1 + 2 n_shells doubles),
// q[K] = c_s rho_v; wt[2 n_shells] = weights then targets.  shell_off:
// the n_shells + 1 offsets of the shells in the table.
hipError_t launch_adjoint_weights(const double2 *rho, const int *shell_off, int n_shells, int n_atoms,
                                  const double *wt, double *res, float2 *q, hipStream_t st);
hipError_t launch_sk_adjoint(const float *x, const float *y, const float *z, int N, double inv_ell, const int *nvec,
                             const float2 *q, int K, int n_splits, double *part, double *gx, double *gy, double *gz,
                             hipStream_t st);

}  // namespace amdat
