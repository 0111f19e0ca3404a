/*
 * amdat_hip.h -- C ABI of the MI355X (gfx950) implementation of AMDAT's
 * reciprocal-space and time-correlation path.
 *
 * Plain C: no exceptions cross this boundary, every call returns an int
 * status (AMDAT_OK or a negative code) and amdat_last_error() gives the
 * message of the calling thread's last failure.  Host buffers belong to
 * the caller.  Device objects (trajectory, shell table, pair table,
 * density) are opaque handles owned by the context that made them.  All
 * calls are synchronous unless a stream has been set with
 * amdat_ctx_set_stream(); then results written to host memory are valid
 * after amdat_ctx_synchronize().
 *
 * The reference (dssimmons-codes/AMDAT) has no FFI: its operator API is the
 * C++ Analysis class hierarchy.  Each entry point below names the
 * reference function or loop it stands in for; the C++ Analysis-class
 * layer in amdat_amd/host calls these in the places where the reference's
 * classes run their loops.
 *
 * Data layout of a trajectory: frame-major structure of arrays,
 * x[t*n_atoms + j], FP32 wrapped coordinates as the trajectory file gives
 * them, int16 image counts in the same layout, box[3*t + c] the box edge
 * length of frame t along axis c (FP32).
 *
 * k-vectors: integer triples n[3*v + c]; the wave vector is
 * k = 2*pi*n/length_scale.  Shell s holds vectors
 * [shell_off[s], shell_off[s+1]) (AMDAT's wavenumber index wn = s, read
 * from file qvector.(wn+2); src/wave_vectors.cpp:430-484).
 */
#ifndef AMDAT_HIP_H
#define AMDAT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define AMDAT_OK 0
#define AMDAT_EINVAL (-1)  /* bad argument: null pointer, size or index out of range */
#define AMDAT_ENOMEM (-2)  /* host or device allocation failed */
#define AMDAT_EDEVICE (-3) /* HIP runtime or kernel launch error */
#define AMDAT_ERANGE (-4)  /* shell / time-gap / pair index outside the object's range */
#define AMDAT_EIO (-5)     /* file could not be opened or parsed */

/* Precision of the phase kernels. */
#define AMDAT_PREC_FP32C 0 /* FP32 phase tables + compensated FP32 sums, FP64 reductions */
#define AMDAT_PREC_FP64 1  /* plain FP64: FP64 phase, FP64 sin/cos, FP64 sums */
#define AMDAT_PREC_FP32X 2 /* rho only: exact phase and one FP32 sincos per update, compensated sums */
#define AMDAT_PREC_FP64T 3 /* tabled FP64: the plan and phase tables of FP32C with FP64 entries and sums */

typedef struct amdat_ctx amdat_ctx;
typedef struct amdat_traj amdat_traj;
typedef struct amdat_shells amdat_shells;
typedef struct amdat_pairs amdat_pairs;
typedef struct amdat_list amdat_list;
typedef struct amdat_rho amdat_rho;
typedef struct amdat_batch amdat_batch;

/* ---- errors, version ---------------------------------------------- */

/* Message of this thread's last failed call ("" if none).  Stands in for
 * the reference's Error(msg, code) / cout-and-exit convention
 * (src/error.cpp:51-146): the C++ layer turns these codes back into that
 * convention. */
const char *amdat_last_error(void);
/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int amdat_version(void);

/* ---- context ------------------------------------------------------ */

/* Select a device and create a context (device memory pools, streams).
 * No reference counterpart: the reference is a single CPU process. */
int amdat_ctx_create(int device, amdat_ctx **out);
int amdat_ctx_destroy(amdat_ctx *ctx);
/* Run subsequent work on this HIP stream (hipStream_t passed as void*);
 * NULL restores the context's own stream. */
int amdat_ctx_set_stream(amdat_ctx *ctx, void *stream);
int amdat_ctx_synchronize(amdat_ctx *ctx);

/* ---- time scheme (host integer work; bit-exact with the reference) -- */

/* n_timesteps = nes*nexp + 1  (src/system.cpp:120). */
int amdat_ts_n_timesteps(int n_exponentials, int n_exponential_steps);
/* n_timegaps = nes + nexp - 1 + !frt  (src/system.cpp:218). */
int amdat_ts_n_timegaps(int n_exponentials, int n_exponential_steps, int frt);
/* Frame times, System::create_timelist (src/system.cpp:3542-3568).
 * out[n_timesteps]. */
int amdat_ts_timelist(int n_exponentials, int n_exponential_steps, float exp_base,
                      int first_exponent, float time_unit, float *out, int cap);
/* Lag time per time gap, System::displacement_times()
 * (src/system.cpp:3763-3784).  out[n_timegaps]. */
int amdat_ts_displacement_times(int n_exponentials, int n_exponential_steps, int frt,
                                const float *timelist, float *out, int cap);
/* Ordered (time gap, t0, t1) pairs of System::displacement_list
 * (src/system.cpp:4090-4179; limit as set by the `limit` command,
 * src/control.cpp:1860-1865, 0 = none).  Writes at most cap pairs;
 * *n_pairs receives the full count. */
int amdat_ts_pairs(int n_exponentials, int n_exponential_steps, int frt, int fullblock,
                   int limit, int *tg, int *t0, int *t1, int cap, int *n_pairs);
/* Pairs of one time gap, System::displacement_list(analysis, timegap,
 * false) (src/system.cpp:4349-4398): the first frame of each block only;
 * the last gap is frames 0 and n_timesteps-1 (time gap n_timegaps-1, or
 * requested as n_timegaps when frt = 1).  limit as in amdat_ts_pairs.  A
 * time gap out of range is AMDAT_ERANGE ("Requested timegap out of
 * range.").  Writes at most cap pairs; *n_pairs receives the count. */
int amdat_ts_gap_pairs(int n_exponentials, int n_exponential_steps, int frt, int limit, int time_index,
                       int *tg, int *t0, int *t1, int cap, int *n_pairs);
/* System::timegap_weighting(fullblock) (src/system.cpp:4599-4628),
 * including its +1 on cross-block gaps.  out[n_timegaps]. */
int amdat_ts_weighting(int n_exponentials, int n_exponential_steps, int frt, int fullblock,
                       int limit, int *out, int cap);

/* ---- shells -------------------------------------------------------- */

/* Read AMDAT q-vector files stem.NNN (NNN = bin, 3 digits; integer
 * triples in any float text format) for bins first_bin ..
 * first_bin+n_bins-1, as Wave_Vectors::read_vectors does
 * (src/wave_vectors.cpp:430-484; the reference reads bins 2..100).
 * shell_off[n_bins+1] and nvec[3*cap_vectors] are written; *n_vectors gets
 * the total. */
int amdat_shells_read(const char *stem, int first_bin, int n_bins, int *shell_off, int *nvec,
                      int cap_vectors, int *n_vectors);

/* The same for any plane of the reference's Wave_Vectors: "xyz" (3D files),
 * "xy", "xz", "yz" (2D files: columns placed on the axes as
 * src/wave_vectors.cpp:266-336 places them) or "x", "y", "z" (1D files,
 * src/wave_vectors.cpp:343-424, whose shells hold half-integers bin/2).
 * Vectors come back as integers n on (x, y, z) with *denominator set to 1,
 * or 2 for the 1D planes: k = 2 pi n / (denominator * length_scale), so a
 * caller passes denominator * length_scale wherever a length scale is
 * asked for.  Other arguments as amdat_shells_read. */
int amdat_shells_read_plane(const char *stem, const char *plane, int first_bin, int n_bins,
                            int *shell_off, int *nvec, int cap_vectors, int *n_vectors, int *denominator);

/* Build shells by the shell rule of src/qvectors/generate_3d_qvectors.m
 * (half-space integer vectors, bin = floor((|n|+0.25)*2), |n_c| <= 100)
 * and keep at most max_per_shell vectors per shell, chosen by a seeded
 * shuffle.  Same output convention as amdat_shells_read. */
int amdat_shells_generate(int first_bin, int n_bins, int max_per_shell, uint64_t seed,
                          int *shell_off, int *nvec, int cap_vectors, int *n_vectors);
/* Upload a shell table to the device. */
int amdat_shells_upload(amdat_ctx *ctx, int n_shells, const int *shell_off, const int *nvec,
                        amdat_shells **out);
/* Tuning override, no reference counterpart (the reference has no device
 * plans): fix the vectors per thread R (2, 4 or 8) and the pairing of slots
 * that share ny or nz (R = 4 or 8; pair = 1: vectors without a partner are
 * padded, 2: they run in unpaired waves, 3: quads of four vectors sharing
 * ny first, then pairs as 1, 4: groups of 8, 4, 2 or 1 vectors sharing nz
 * or ny, placed against LDS bank clashes, 5 (R = 8 only): as 4, with waves
 * of single vectors holding 4 per thread and waves of pairs 6, so that all
 * waves carry about equal work) for the single-trajectory phase kernels run
 * on this table; R = 0 restores the automatic choice (lowest modelled
 * cost).  Results do not depend on it beyond FP32 rounding. */
int amdat_shells_plan_hint(amdat_shells *sh, int R, int pair);
int amdat_shells_free(amdat_ctx *ctx, amdat_shells *sh);

/* ---- trajectory and atom lists ------------------------------------- */

/* Upload a trajectory (host arrays, layout above).  Stands in for the
 * reference's per-atom Coordinate arrays (src/trajectory.h:37-38) and the
 * image-count unwrap of the custom reader (src/system.cpp:1290-1297):
 * unwrapped positions are rebuilt on the device as x + ix*L. */
int amdat_traj_upload(amdat_ctx *ctx, int n_atoms, int n_frames, const float *x, const float *y,
                      const float *z, const int16_t *ix, const int16_t *iy, const int16_t *iz,
                      const float *box, amdat_traj **out);
/* Same, from device pointers the caller keeps alive (no copy). */
int amdat_traj_wrap_device(amdat_ctx *ctx, int n_atoms, int n_frames, const float *d_x,
                           const float *d_y, const float *d_z, const int16_t *d_ix,
                           const int16_t *d_iy, const int16_t *d_iz, const float *box_host,
                           amdat_traj **out);
int amdat_traj_free(amdat_ctx *ctx, amdat_traj *tr);
/* Atoms, frames and (box != NULL) the per-frame box lengths [n_frames][3]
 * of a trajectory. */
int amdat_traj_shape(const amdat_traj *tr, int *n_atoms, int *n_frames, float *box);

/* ---- trajectory files (src/system.cpp:237-299 read_trajectory) ---------- */

/* Parse a LAMMPS custom dump (src/system.cpp:1026-1367 read_custom) into
 * frame-major SoA host arrays [n_frames][n_atoms], atoms in file order:
 * columns x y z with ix iy iz; or xu yu zu (wrapped coordinates and images
 * then follow from the box: image = floor((u - lo)/L)); or x y z alone
 * (images 0).  box [n_frames][3] = hi - lo; box_lo (optional) = lo; types
 * (optional) = the type column of frame 0.  Frames are parsed in parallel
 * on `threads` threads (0: all hardware threads).  Fails if the file holds
 * fewer than n_frames frames or a frame's atom count differs. */
int amdat_read_custom(const char *path, int n_atoms, int n_frames, int threads, float *x, float *y, float *z,
                      int16_t *ix, int16_t *iy, int16_t *iz, float *box, float *box_lo, int *types);

/* Parse an xyz trajectory (src/system.cpp:610-705 read_xyz; the box is
 * constant, box3 = L_x L_y L_z of the system block, src/system.cpp:385-397):
 * frames of a count line, a comment line and "type x y z" lines.  Image
 * counts follow the reference's unwrapping (src/trajectory.cpp:233-270):
 * between consecutive frames the FP32 jump d moves an atom by
 * -L * int(d / (0.55 L)). */
int amdat_read_xyz(const char *path, int n_atoms, int n_frames, const float *box3, int threads, float *x,
                   float *y, float *z, int16_t *ix, int16_t *iy, int16_t *iz, int *types);

/* Read a trajectory file straight into device memory: format "custom" or
 * "xyz" (box3 required for xyz, ignored for custom).  Frames are parsed in
 * parallel into pinned staging buffers, chunk by chunk, and each chunk's
 * copy to the device runs while the next chunk is parsed; no host copy of
 * the trajectory is kept.  types (optional, host [n_atoms]) receives frame
 * 0's atom types.  Replaces the reader plus the upload of the reference's
 * System construction (src/system.cpp:152). */
int amdat_traj_read(amdat_ctx *ctx, const char *format, const char *path, int n_atoms, int n_frames,
                    const float *box3, int threads, int *types, amdat_traj **out);

/* Box lower bounds [n_frames][3] of a trajectory: the bounds lines of a
 * custom file (amdat_traj_read), zeros for xyz files and uploads.  With the
 * box lengths (amdat_traj_shape) they give System::boundaries
 * (src/system.cpp). */
int amdat_traj_bounds(const amdat_traj *tr, float *lo);

/* Rectangular bins of Trajectory_List_Bins (src/trajectory_list_bins.cpp:
 * 207-279 "all", 280-394 "region", assign_bins 949-1096): for every frame
 * t and atom j, cell[t*n_atoms + j] = (x*n_ybins + y)*n_zbins + z for the
 * atom's stored coordinate c strictly inside (lo, hi) of frame t on every
 * axis (lo, hi: [n_frames][3]), with x = int((c - lo + 10 FLT_EPSILON) /
 * (hi - lo + 20 FLT_EPSILON) * n_xbins) formed as the reference forms it,
 * or -1 outside.  cell is host memory [n_frames][n_atoms]. */
int amdat_bins_assign(amdat_ctx *ctx, const amdat_traj *tr, int n_xbins, int n_ybins, int n_zbins, const float *lo,
                      const float *hi, int *cell);

/* Distance bins of Trajectory_List_Bins (create_bin_list distance ...,
 * src/control.cpp:3915-3985): for every frame t and every atom j of
 * `binned` at t (NULL = all atoms), cell[t*n_atoms + j] = int(distance /
 * thickness) when below n_bins, else -1 (also for atoms not in `binned`).
 * Distances in FP32 as the reference forms them, L the frame's box lengths:
 *  AMDAT_DIST_POINT (src/trajectory_list_bins.cpp:1321-1488): point[3];
 *    d = r - point, d -= L * trunc(d / (L/2)) per axis, |d|;
 *  AMDAT_DIST_PLANE (:1489-1774): axis 0/1/2, point[0] = position,
 *    direction 0 average (|L * trunc(d / (L/2)) - d|), 1 above (d < 0 ?
 *    L + d : d), 2 below (d < 0 ? |d| : L - d), d = r[axis] - position;
 *  AMDAT_DIST_LIST (:1097-1320): the smallest distance to the atoms of
 *    `cluster` at t (NULL = all atoms), shortest images in the box of frame
 *    0, starting from the frame's largest box length.
 * A point or plane position beyond a frame's box length is AMDAT_ERANGE, as
 * the reference stops there.  n_bins is the bin count the caller settled
 * (the constructors cap it at int(reach / thickness) + 1).  cell is host
 * memory [n_frames][n_atoms]. */
#define AMDAT_DIST_POINT 0
#define AMDAT_DIST_PLANE 1
#define AMDAT_DIST_LIST 2
int amdat_bins_distance(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *binned, int kind,
                        const float *point, int axis, int direction, const amdat_list *cluster, float thickness,
                        int n_bins, int *cell);

/* Cell tables kept on the device: the cells of amdat_bins_assign
 * (amdat_cells_assign, n_xbins*n_ybins*n_zbins bins) and of
 * amdat_bins_distance (amdat_cells_distance, n_bins bins), same arguments,
 * for forming per-bin lists without copying the [n_frames][n_atoms] table
 * to the host.  amdat_cells_download copies it ([n_frames][n_atoms]). */
typedef struct amdat_cells amdat_cells;
int amdat_cells_assign(amdat_ctx *ctx, const amdat_traj *tr, int n_xbins, int n_ybins, int n_zbins,
                       const float *lo, const float *hi, amdat_cells **out);
int amdat_cells_distance(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *binned, int kind,
                         const float *point, int axis, int direction, const amdat_list *cluster,
                         float thickness, int n_bins, amdat_cells **out);
int amdat_cells_download(amdat_ctx *ctx, const amdat_cells *cells, int *cell);
int amdat_cells_free(amdat_ctx *ctx, amdat_cells *cells);

/* Per-bin lists formed on the device, lists[b] for bin first_bin + b
 * (b < n_bins):
 *  amdat_cells_bin_lists: Trajectory_List_Bins intersected with list l
 *    (Trajectory_List::operator&&, src/trajectory_list.cpp:518-; l NULL =
 *    all atoms): a time-varying list with one list time per frame holding
 *    the atoms of l at that frame that lie in the bin, in ascending index
 *    order;
 *  amdat_cells_pair_lists: the persistence form of Bin_Dynamics_Analysis
 *    (src/bin_dynamics_analysis.h:120-141): a pair-keyed list (see
 *    amdat_list_create_paired) with key k = (t0[k], t1[k]) holding the
 *    atoms of l at t0 that lie in the bin at t0 and at t1, ascending.
 * Each list is freed with amdat_list_free.  ERANGE: bin range outside the
 * table or frame out of range; EINVAL: pair-keyed l, a pair given twice. */
int amdat_cells_bin_lists(amdat_ctx *ctx, const amdat_cells *cells, const amdat_list *l, int first_bin,
                          int n_bins, amdat_list **lists);
int amdat_cells_pair_lists(amdat_ctx *ctx, const amdat_cells *cells, const amdat_list *l, int n_keys,
                           const int *t0, const int *t1, int first_bin, int n_bins, amdat_list **lists);

/* Static atom selection (Static_Trajectory_List, src/static_trajectory_
 * list.cpp:59-101): atom indices in list order.  indices == NULL with
 * n == n_atoms of the trajectory means "all atoms in order". */
int amdat_list_create(amdat_ctx *ctx, const amdat_traj *tr, int n, const int *indices,
                      amdat_list **out);
int amdat_list_free(amdat_ctx *ctx, amdat_list *l);

/* Time-varying atom selection (Trajectory_List with n_times list times,
 * src/trajectory_list.h:22-30; Exptime_Trajectory_List, src/exptime_
 * trajectory_list.cpp:29-56): list time c holds indices[off[c] ..
 * off[c+1]-1] in list order; time_conv[t] (n_frames entries) is the list
 * time of frame t, and a frame whose list time lies outside [0, n_times)
 * has no atoms.  Every entry point taking an amdat_list takes such a list:
 * rho and S(k) use the list at each frame, Fs, MSD/NGP and the displacement
 * statistics the list at the list time of each pair's t0
 * (src/trajectory_list.cpp:191-211), and the counts follow.  The batched
 * entry points take static lists only. */
int amdat_list_create_timed(amdat_ctx *ctx, const amdat_traj *tr, int n_times, const int *time_conv,
                            const int *off, const int *indices, amdat_list **out);
/* Pair-keyed atom selection: the atoms of key c, indices[off[c] ..
 * off[c+1]-1], apply to the frame pair (t0[c], t1[c]); a frame pair with
 * no key has no atoms, and no pair may be given twice.  It replaces the
 * per-pair list of Bin_Dynamics_Analysis::list_displacementkernel
 * (src/bin_dynamics_analysis.h:120-141: the list's atoms at t0 that lie
 * in one bin at both t0 and t1, Trajectory_List::time_intersection,
 * src/trajectory_list.cpp:252-277), handed to each analysis's bin_hook.
 * Fs, MSD/NGP and the displacement statistics take it, with the counts per
 * time gap summed over the pairs' own lists; rho, S(k), the displacement
 * selection and the batched entry points reject it (AMDAT_EINVAL). */
int amdat_list_create_paired(amdat_ctx *ctx, const amdat_traj *tr, int n_keys, const int *t0, const int *t1,
                             const int *off, const int *indices, amdat_list **out);
/* n_times (0 for a static list, the key count for a pair-keyed one) and
 * total entries. */
int amdat_list_shape(const amdat_list *l, int *n_times, int *n_entries);
/* Copy a list to the host: time_conv[n_frames] (time-varying lists) and
 * off[n_times+1] (time-varying and pair-keyed lists; both ignored for
 * static ones) and indices[n_entries]; any output may be NULL. */
int amdat_list_download(amdat_ctx *ctx, const amdat_list *l, int *time_conv, int *off, int *indices);

/* Atoms selected by squared displacement at one time gap, per exponential
 * block: Fast_Particles (src/fast_particles.cpp:117-139; the
 * find_fast_fixedthreshold command, src/control.cpp:3466-3496),
 * Slow_Particles (src/slow_particles.cpp:93-113) and Average_Particles
 * (src/average_particles.cpp:95-116).  Pairs are those of the time gap as
 * System::displacement_list(analysis, time_index, false) lists them
 * (src/system.cpp:4349-4398, with limit, 0 = none); for each, the atoms of
 * `in` (NULL = all atoms) at the list time of t0 whose squared displacement
 * d2 = pow(distance, 2), formed in FP32 as the reference forms it, passes
 * the test are appended, in list order, to the list of block
 * int((t0 - frt) / n_exponential_steps).  The result is a time-varying
 * list with n_exponentials list times and the exponential time conversion.
 * FAST: d2 > lo; SLOW: d2 < hi; BETWEEN: lo <= d2 <= hi. */
#define AMDAT_SELECT_FAST 0
#define AMDAT_SELECT_SLOW 1
#define AMDAT_SELECT_BETWEEN 2
int amdat_list_select_displacement(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *in, int n_exponentials,
                                   int n_exponential_steps, int frt, int limit, int time_index, int mode, float lo,
                                   float hi, amdat_list **out);

/* ---- pairs --------------------------------------------------------- */

/* Upload (tg, t0, t1) pairs as produced by amdat_ts_pairs. */
int amdat_pairs_upload(amdat_ctx *ctx, int n_pairs, const int *tg, const int *t0, const int *t1,
                       int n_timegaps, amdat_pairs **out);
int amdat_pairs_free(amdat_ctx *ctx, amdat_pairs *p);

/* ---- the path ------------------------------------------------------ */

/* rho_k(t) = sum_j exp(i k.r_j(t)) over wrapped coordinates for every
 * frame in frames[n_frames] (NULL = all frames) and every vector of
 * shells first_shell..last_shell, kept on the device.
 * Wave_Density::analyze + listkernel (src/wave_density.cpp:313-365);
 * k = 2*pi*n/length_scale (the reference's Wave_Vectors spacing,
 * src/wave_vectors.cpp:150). */
int amdat_rho_compute(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list,
                      const amdat_shells *sh, int first_shell, int last_shell,
                      float length_scale, const int *frames, int n_frames, int precision,
                      amdat_rho **out);
/* Copy rho to the host: out[2*(f*K + v)] = Re, +1 = Im, f over the
 * computed frames, v over the vectors of the computed shells in order. */
int amdat_rho_download(amdat_ctx *ctx, const amdat_rho *rho, double *out, size_t n_doubles);
/* Number of frames and vectors held. */
int amdat_rho_shape(const amdat_rho *rho, int *n_frames, int *n_vectors);
int amdat_rho_free(amdat_ctx *ctx, amdat_rho *rho);

/* S(k) numerator per shell: sum over the rho's frames of
 * sum_{v in shell} Re(rho1 conj rho2)/H, FP64, and the atom count
 * divisor is applied: S[s] = numerator / (n_frames * list size).
 * Structure_Factor::timekernel/timekernel2 + postprocess_list
 * (src/structure_factor.cpp:201-285).  rho2 may equal rho1
 * (symmetric form).  S[n_shells of rho]. */
int amdat_sk(amdat_ctx *ctx, const amdat_rho *rho1, const amdat_rho *rho2, double *S);

/* F(k,t) from two densities over all frames:
 * F[tg][s] = sum_pairs sum_{v in s} Re(rho1(t0) conj rho2(t1))/H
 *            / sum_pairs N(t0).
 * Intermediate_Scattering_Function(sys,wd1,wd2,wn1,wn2,fullblock)
 * (src/intermediate_scattering_function.cpp:237-322, 551-566) with the
 * pair loop of System::displacement_loop (src/system.cpp:3918-3964).
 * Shells wn1..wn2 are given relative to the densities' first shell.
 * F[n_timegaps*(wn2-wn1+1)], count[n_timegaps] (sum_pairs N(t0)). */
int amdat_fkt(amdat_ctx *ctx, const amdat_rho *rho1, const amdat_rho *rho2,
              const amdat_pairs *pairs, int wn1, int wn2, double *F, double *count);

/* F(k,t) over arbitrary vector ranges of two densities: column c sums
 * the vectors v_first[c] .. v_first[c]+v_count[c]-1 (positions relative to
 * the densities' first vector) and divides by v_count[c]; otherwise as
 * amdat_fkt.  Expresses the reference's reading of shells wn1..wn2 by
 * Intermediate_Scattering_Function(sys, wd1, wd2, wn1, wn2, ...): it takes
 * the density's slot wn - wn1, counted from the density's first shell,
 * with the vector count of shell wn (src/intermediate_scattering_function.
 * cpp:556-562), which only names shell wn when wn1 is the density's first
 * shell.  F[n_timegaps*n_cols], count[n_timegaps]. */
int amdat_fkt_vectors(amdat_ctx *ctx, const amdat_rho *rho1, const amdat_rho *rho2, const amdat_pairs *pairs,
                      int n_cols, const int *v_first, const int *v_count, double *F, double *count);

/* A density that is the sum of two over the same frames and vectors: the
 * reference's Wave_Density::analyze adds each call's rho_k(t) and atom
 * counts to what earlier calls left (src/wave_density.cpp:313-336). */
int amdat_rho_sum(amdat_ctx *ctx, const amdat_rho *a, const amdat_rho *b, amdat_rho **out);

/* Fs(k,t) = sum_pairs sum_atoms mean_{v in s} cos(k.(r(t1)-r(t0)))
 *           / (atom-pair count), with r(t1)-r(t0) rebuilt in FP64 from
 * wrapped coordinates and image counts.
 * Incoherent_Scattering_Function (src/incoherent_scattering_function.cpp:
 * 176-299) + Correlation_2D::postprocess_list (src/correlation_2d.cpp:
 * 150-164).  Fs[n_timegaps*(last-first+1)], count[n_timegaps]. */
int amdat_fs(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list,
             const amdat_shells *sh, int first_shell, int last_shell, float length_scale,
             const amdat_pairs *pairs, int precision, double *Fs, double *count);

/* MSD and the non-Gaussian parameter's fourth moment: per time gap the
 * FP64 sums of |dr|^2 and |dr|^4 over pairs and atoms and the atom-pair
 * count.  Mean_Square_Displacement (src/mean_square_displacement.cpp:
 * 140-180) and Non_Gaussian_Parameter (src/non_gaussian_parameter.cpp:
 * 96-143).  s2, s4, count [n_timegaps]. */
int amdat_msd_ngp(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list,
                  const amdat_pairs *pairs, double *s2, double *s4, double *count);

/* amdat_msd_ngp with no allocation: the same results, bit for bit, for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  It calls no hipMalloc, hipFree or
 * allocating library routine: its partial sums live in a device array the
 * library holds per GPU (32 MiB), which bounds n_pairs * atom blocks * 16 B
 * + n_timegaps * 16 B (atom blocks = min(64, ceil(list size / 1024)); a
 * larger call fails with AMDAT_ERANGE).  Static lists only (AMDAT_EINVAL
 * otherwise).  Every call on one GPU shares that array: calls that overlap
 * on one GPU (two host threads, or two streams) race and must be
 * serialised by the caller. */
int amdat_msd_ngp_noalloc(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list,
                          const amdat_pairs *pairs, double *s2, double *s4, double *count);

/* Displacement statistics over the pairs (src/mean_displacement.cpp:163-176,
 * src/mean_square_displacement_2d.cpp:144-156, src/van_hove_self.cpp:98-102):
 * per time gap g, FP64 sums over its pairs and the list's atoms of the
 * displacement components d_c (sum_d[3g + c]) and their squares
 * (sum_d2[3g + c]; a plane's MSD is the sum of two of them), with
 * d = (x1 - x0) + (i1 L1 - i0 L0) rebuilt in FP64; count[g] = atom pairs.
 * With n_bins > 0, hist[g*n_bins + b] counts the atom pairs whose |dr|,
 * formed as the reference forms it (FP32 unwrapped coordinates
 * float(x + i L), FP32 difference and length: src/system.cpp:1292-1294,
 * src/coordinate.h:55), falls in bin b = int(|dr| / (max_range / n_bins))
 * (src/space_time_correlation_function.cpp:216-230), exactly; n_bins <=
 * 16384.  The reference keeps these counts in FP32, which stops counting
 * exactly past 2^24 per bin; here they are 64-bit integers. */
int amdat_displacements(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list,
                        const amdat_pairs *pairs, int n_bins, float max_range, double *sum_d,
                        double *sum_d2, unsigned long long *hist, double *count);

/* Distribution of squared displacements at one time gap: DebyeWaller_Dist
 * (the u2dist command, src/debyewaller_dist.cpp:47-95, src/control.cpp:
 * 3247-3273) and Displacement_Distribution (displacement_dist, src/
 * displacement_distribution.cpp:50-98, src/control.cpp:3312-3339; its
 * power argument is not used there either).  Pass the pairs of the time
 * gap (amdat_ts_gap_pairs).  counts[n_bins + 1] receives, over all pairs
 * and the list's atoms, the number whose squared displacement d2 =
 * pow(distance, 2) stored as float (the FP32 length of the FP32 unwrapped
 * difference) falls in bin int(d2 / (max_value / n_bins)), the last bin
 * taking the overflow, exactly; sum_d2 and sum_d4 receive FP64 sums of
 * |dr|^2 and |dr|^4 (FP64 rebuild) and count the atom pairs: the mean is
 * sum_d2 / count, the variance sum_d4 / count - mean^2, and bin b's value
 * counts[b] / count (the reference's 1/weighting factors cancel).
 * n_bins <= 16383. */
int amdat_displacement_dist(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list, const amdat_pairs *pairs,
                            int n_bins, float max_value, unsigned long long *counts, double *sum_d2,
                            double *sum_d4, double *count);

/* ---- differentiation (no reference counterpart) ----------------------- */

/* Structure-matching loss on one frame given by device FP32 coordinate
 * arrays d_x, d_y, d_z [n_atoms] (wrapped or not: k = 2*pi*n/length_scale):
 *   L = sum_s w_s (S_s - T_s)^2,  S_s = sum_{v in s} |rho_v|^2 / (H_s N),
 * over shells first_shell..last_shell; S[n] receives S_s (may be NULL).
 * When d_gx, d_gy, d_gz (device FP64 [n_atoms]) are given, also writes
 * dL/dr_j from an adjoint kernel that recomputes every phase
 * exp(i k.r_j) instead of storing the N x K of them.  Used by the
 * structure-matching torch.autograd.Function in amdat_amd/torch_sk.py. */
int amdat_sk_loss_grad(amdat_ctx *ctx, int n_atoms, const float *d_x, const float *d_y, const float *d_z,
                       const amdat_shells *sh, int first_shell, int last_shell, float length_scale,
                       const double *weights, const double *target, double *loss, double *S,
                       double *d_gx, double *d_gy, double *d_gz);

/* ---- batches of trajectories (no reference counterpart) ---------------- */
/* The reference analyses one system per process run.  A batch analyses
 * many trajectories of uneven size in one persistent launch per analysis:
 * the launch keeps as many blocks resident as fit and they draw work items
 * (trajectory x vector chunk x frame group or time gap) from a queue,
 * costliest first.  Each member follows the single-trajectory semantics
 * cited below.  Members must hold the same number of frames (one time
 * scheme, one pair table).  The handle keeps pointers to the members:
 * free the batch before them.  Concurrent batch calls on one batch race on
 * its queue counter. */

/* Build a batch of n_traj trajectories; lists[i] (or lists == NULL, or a
 * NULL entry) selects the atoms of member i (NULL = all atoms in order). */
int amdat_batch_create(amdat_ctx *ctx, int n_traj, const amdat_traj *const *trajs,
                       const amdat_list *const *lists, amdat_batch **out);
int amdat_batch_free(amdat_ctx *ctx, amdat_batch *batch);

/* S(k) of every member over one shell range and one frame list (NULL = all
 * frames), k = 2*pi*n/length_scales[i] (each member's own length scale):
 * S[i*n_shells + s] = sum_frames sum_{v in s} |rho_v|^2 / (H_s * count[i]),
 * count[i] = n_frames * atoms of member i.  Per member this is
 * amdat_rho_compute + amdat_sk (Structure_Factor::analyze,
 * src/structure_factor.cpp:171-320; Wave_Density::listkernel,
 * src/wave_density.cpp:339-365). */
int amdat_batch_sk(amdat_ctx *ctx, amdat_batch *batch, const amdat_shells *sh, int first_shell,
                   int last_shell, const float *length_scales, const int *frames, int n_frames,
                   double *S, double *count);

/* Fs(k,t) of every member, member i on its own shell range
 * first_shell[i]..last_shell[i] of one table and its own length scale, all
 * over one pair table.  Member i's rows [n_timegaps][n_shells_i] start at
 * Fs + sum_{j<i} n_timegaps * n_shells_j; count[i*n_timegaps + g] = atom
 * pairs of gap g.  Per member this is amdat_fs
 * (Incoherent_Scattering_Function, src/incoherent_scattering_function.cpp:176-299). */
int amdat_batch_fs(amdat_ctx *ctx, amdat_batch *batch, const amdat_shells *sh, const int *first_shell,
                   const int *last_shell, const float *length_scales, const amdat_pairs *pairs,
                   double *Fs, double *count);

/* MSD / NGP sums of every member: s2, s4, count are [n_traj][n_timegaps].
 * Per member this is amdat_msd_ngp (Mean_Square_Displacement,
 * src/mean_square_displacement.cpp:140-180; Non_Gaussian_Parameter,
 * src/non_gaussian_parameter.cpp:96-143). */
int amdat_batch_msd_ngp(amdat_ctx *ctx, amdat_batch *batch, const amdat_pairs *pairs, double *s2,
                        double *s4, double *count);

/* One analysis job over a batch in a single persistent launch: S(k) items
 * (trajectory x vector chunk x frame group), Fs items (trajectory x chunk
 * x time gap) and MSD/NGP items (trajectory x time gap) share one queue,
 * costliest first, followed by the per-shell reductions.  Any part may be
 * left out: sk_shells == NULL (no S(k)), fs_shells == NULL (no Fs),
 * want_msd == 0.  Arguments and outputs are those of amdat_batch_sk,
 * amdat_batch_fs and amdat_batch_msd_ngp (pairs serves Fs and MSD).
 * flags: AMDAT_RUN_GRAPH captures the launch sequence in a hipGraph on the
 * first call and replays it while the arguments stay the same.  Without
 * AMDAT_RUN_TUNE the launch settings come from a cost model: the S(k) plan
 * (vectors per thread, slot grouping) the single-trajectory path would
 * choose, frames per S(k) item for about 16,384 items, Fs vectors per thread
 * by the same model, whole time gaps per Fs item, no atom parts.
 * AMDAT_RUN_TUNE times, once per argument set and separately for the S(k)
 * and the Fs items: vectors per thread and plan mode (k-vectors per LDS
 * tile: 1024, 2048 or 4096; S(k) over every 8th frame), then the S(k)
 * frames per item (x1/2, x1, x2 the model's), the Fs pairs per item (whole
 * gaps, 2, 1) and Fs atom parts (1, 2, 4); it keeps
 * the tuned settings only when the whole job runs no slower with them than
 * with the model's (amdat_batch_job_settings reports both timings).  S(k)
 * items never split a frame's atoms (|rho|^2 is not additive over atom
 * parts); the atoms per LDS tile are fixed per table width at compile time
 * (the largest count whose tables keep two blocks per CU).  The batch keeps the
 * planned job (device tables, queue, graph) until another job replaces it
 * or the batch is freed: free the batch before the shells and pairs its
 * last job used. */
#define AMDAT_RUN_GRAPH 1
#define AMDAT_RUN_TUNE 2
int amdat_batch_run(amdat_ctx *ctx, amdat_batch *batch, const float *length_scales,
                    const amdat_shells *sk_shells, int sk_first, int sk_last, const int *frames,
                    int n_frames, const amdat_shells *fs_shells, const int *fs_first, const int *fs_last,
                    const amdat_pairs *pairs, int want_msd, int flags, double *S, double *S_count,
                    double *Fs, double *Fs_count, double *s2, double *s4, double *msd_count);

/* The last job's vectors per thread for S(k) (RS) and Fs (RF) items,
 * whether its S(k) plan pairs slots sharing ny or nz, its number of work
 * items, and whether it runs as a captured graph. */
int amdat_batch_job_info(const amdat_batch *batch, int *RS, int *RF, int *paired, int *n_items, int *graph);

/* Every launch setting of the last job (no counterpart in the reference,
 * which has no batched launch): S(k) vectors per thread, plan mode (0
 * unpaired, 1 paired, 4 grouped in 8, 4, 2 or 1 slots, 5 grouped with waves
 * of singles and pairs holding 4 and 6 slots), frames per item and
 * frame groups; Fs vectors per thread, pairs per item (0: whole time
 * gaps), atom parts per pair chunk and pair chunks; atoms per LDS tile and
 * table half-width; work items and resident blocks of the persistent
 * launch; graph replay; whether the settings came from the autotuner, with
 * the tuner's whole-job timings (ms, all trajectories) of the cost-model
 * settings and of the kept ones (0 when not tuned).  Zero for parts the
 * job does not compute. */
typedef struct amdat_job_settings {
    int sk_R, sk_plan_mode, sk_frames_per_item, sk_frame_groups;
    int fs_R, fs_pairs_per_item, fs_atom_parts, fs_pair_chunks;
    int atoms_per_tile, table_half_width, n_items, grid;
    int graph, tuned;
    float model_ms, tuned_ms;
} amdat_job_settings;
int amdat_batch_job_settings(const amdat_batch *batch, amdat_job_settings *out);

/* Fix the launch settings of the batch's next jobs (no counterpart in the
 * reference): sk_R with sk_plan_mode, sk_frames_per_item, fs_R,
 * fs_pairs_per_item and fs_atom_parts of *s; a zero field keeps the cost
 * model's choice (fs_pairs_per_item 0: whole time gaps), the other fields
 * are ignored, and AMDAT_RUN_TUNE is then ignored.  s == NULL returns to
 * the cost model and tuner.  For settings read back from
 * amdat_batch_job_settings of a tuned run, or for tests of each item form.
 * EINVAL: R not 2, 4 or 8; plan mode not 0, 1, 4 or 5 (1 and 4 need R = 4
 * or 8, 5 needs R = 8); negative counts; more than 64 atom parts. */
int amdat_batch_use_settings(amdat_batch *batch, const amdat_job_settings *s);

#ifdef __cplusplus
}
#endif

#endif /* AMDAT_HIP_H */
