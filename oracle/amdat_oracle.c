/*
 * amdat_oracle.c -- CPU oracle for AMDAT's reciprocal-space and
 * time-correlation path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this file: only tests/, __graft_entry__.smoke() (as the checker) and the
 * cpu_baseline leg of bench.py use it.
 *
 * Three families of functions:
 *
 *   orc_ts_*      integer time-scheme work (frame times, time-gap count,
 *                 time-origin pairs, pair weights).  Follows the reference
 *                 bit for bit, quirks included.
 *   orc_*_f64     FP64 restatement of the physics: exact k = 2*pi*n/l in
 *                 double, double trig, on the same FP32 coordinates and
 *                 integer image counts the product receives.  This is the
 *                 "plain FP64" truth every GPU output is gated against.
 *   orc_ref32_*   restatement of the reference's own FP32 arithmetic (same
 *                 types, operation order and loop order; build with
 *                 -ffp-contract=off and no -march, as the reference is
 *                 built).  Used to time the reference-equivalent CPU path
 *                 and to show the reference's own error level.
 *
 * How each function is pinned (tests/test_oracle*.py):
 *   orc_ts_*            closed-form pair counts, hand-enumerated small
 *                       schemes, weighting quirk (+1) values.
 *   orc_shell_enumerate set equality with the reference's q-vector data
 *                       (bins 2..19, complete shells) and with the shell
 *                       rule; per-shell counts against the data file sizes.
 *   orc_rho_f64         one atom (|rho|=1), two atoms (|rho|^2 = 2+2cos),
 *                       perfect fcc Bragg peak (|rho| = N), lattice
 *                       translation invariance, numpy brute force with
 *                       complex exp on tiny inputs.
 *   orc_sk_f64          ideal-gas mean -> 1, one atom -> 1, fcc Bragg -> N.
 *   orc_fkt_f64         F(k,0) = S(k) on the same frames, Gaussian-walk
 *                       F(k,t)/S(k) decay, normalisation by atom count.
 *   orc_fs_f64          Fs(k,0) = 1, Gaussian random walk
 *                       Fs = exp(-k^2 <r^2>/6), sign/axis permutations.
 *   orc_msd_ngp_f64     MSD(0) = 0, single-atom closed forms, image-count
 *                       rebuild equal to an explicit unwrap, Gaussian
 *                       walk MSD = 3 s^2 n.
 *   orc_bins_assign     interior points map to floor((c - lo)/w), bin
 *                       boundaries go to the upper bin, box faces and
 *                       points outside the region are excluded.
 *   orc_bins_distance   FP64 minimum-image distances (point, plane in each
 *                       direction, nearest atom of a second list by brute
 *                       force) away from bin edges, the above/below
 *                       complement L, the reach cap, list membership.
 *   orc_displacement_dist  constructed trajectories with exactly
 *                       representable squared displacements (bin counts
 *                       in closed form, bin edges, the overflow bin).
 *   orc_select_displacement  constructed trajectories with exactly
 *                       representable displacements (membership known in
 *                       closed form per block, image crossings, the squared
 *                       threshold), list order, limit, cross-block and last
 *                       gaps, the list-time conversion.
 *   orc_ref32_*         against the FP64 functions within the FP32 error
 *                       the reference itself carries.  Bit-level parity
 *                       with the reference's binary: parity unpinned (the
 *                       reference cannot be built here: its system.cpp
 *                       includes a generated version header and its
 *                       control.cpp needs fftw3.h, absent in this image).
 *
 * Data layout (shared with the product): coordinates are frame-major SoA,
 * x[t*N + j]; image counts int16 in the same layout; box[t*3 + c] is the
 * box edge length of frame t along axis c.  k-vectors are integer triples
 * n[3*v + c]; shell s holds vectors [shell_off[s], shell_off[s+1]).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* The reference's wave_vectors.cpp:16 value of pi. */
#define REF_PI 3.14159265358979323846

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }

/* ------------------------------------------------------------------ */
/* Time scheme (integer / FP32 work, bit-exact with the reference)      */
/* ------------------------------------------------------------------ */

/* src/system.cpp:120 -- n_timesteps = n_exponential_steps*n_exponentials+1 */
int orc_ts_n_timesteps(int nexp, int nes) { return nes * nexp + 1; }

/* src/system.cpp:218 -- n_timegaps = nes + nexp - 1 + !frt */
int orc_ts_n_timegaps(int nexp, int nes, int frt) { return nes + nexp - 1 + (frt ? 0 : 1); }

/* src/system.cpp:3542-3568 (create_timelist).  exp_base is a float member
 * (src/system.h:53); pow(float,int) promotes to double. */
void orc_ts_timelist(int nexp, int nes, float base, int first_exp, float dt, float *out)
{
    int timeii = 0;
    float block_start = 0.0f;
    out[0] = 0.0f;
    for (int b = 0; b < nexp; b++) {
        for (int e = 1; e <= nes; e++) {
            timeii++;
            double p = pow((double)base, (double)(e - 1 + first_exp));
            if (p <= (double)e)
                out[timeii] = block_start + (float)e * dt;
            else
                out[timeii] = block_start + (float)floor(p) * dt;
        }
        block_start = out[timeii];
    }
}

/* src/system.cpp:3763-3784 (displacement_times). */
void orc_ts_displacement_times(int nexp, int nes, int frt, const float *timelist, float *out)
{
    int ntg = orc_ts_n_timegaps(nexp, nes, frt);
    int T = orc_ts_n_timesteps(nexp, nes);
    for (int i = 0; i < nes; i++)
        out[i] = timelist[i + frt] - timelist[frt];
    for (int i = 1; i < nexp; i++)
        out[i + nes - 1] = timelist[nes * i + frt] - timelist[frt];
    if (!frt)
        out[ntg - 1] = timelist[T - 1] - timelist[0];
}

/* src/system.cpp:4090-4179 (displacement_list): ordered (tg, t0, t1)
 * triples.  limit == 0 means no cap.  Returns the number of pairs; writes
 * at most cap of them. */
int orc_ts_pairs(int nexp, int nes, int frt, int fullblock, int limit,
                 int *tg_out, int *t0_out, int *t1_out, int cap)
{
    int n = 0;
    int ntg = orc_ts_n_timegaps(nexp, nes, frt);
    int T = orc_ts_n_timesteps(nexp, nes);
#define EMIT(a, b, c) do { if (n < cap) { tg_out[n] = (a); t0_out[n] = (b); t1_out[n] = (c); } n++; } while (0)
    for (int tg = 0; tg < nes; tg++) {
        int count = 0, abort_ = 0;
        for (int b = 0; b < nexp; b++) {
            if (!abort_) {
                int t0 = nes * b + frt;
                EMIT(tg, t0, t0 + tg);
                count++;
                abort_ = (count == limit && limit != 0);
            }
        }
    }
    for (int tg = nes; tg < ntg - 1 + frt; tg++) {
        int count = 0, abort_ = 0;
        int btg = tg - nes + 1;
        for (int e = 0; e < (fullblock ? 1 : 0) * (nes - 1) + 1; e++) {
            if (abort_) continue;
            for (int b = 0; b < nexp - btg; b++) {
                if (abort_) continue;
                int t0 = nes * b + e + frt;
                EMIT(tg, t0, t0 + nes * btg);
                count++;
                abort_ = (count == limit && limit != 0);
            }
        }
    }
    if (!frt)
        EMIT(ntg - 1, 0, T - 1);
#undef EMIT
    return n;
}

/* src/system.cpp:4599-4628 (timegap_weighting), including the +1 on the
 * cross-block entries. */
void orc_ts_weighting(int nexp, int nes, int frt, int fullblock, int limit, int *w)
{
    int ntg = orc_ts_n_timegaps(nexp, nes, frt);
    for (int tg = 0; tg < nes; tg++) {
        w[tg] = nexp;
        if (w[tg] > limit && limit != 0) w[tg] = limit;
    }
    for (int tg = nes; tg < ntg - 1 + frt; tg++) {
        int btg = tg - nes + 1;
        w[tg] = (nexp - btg) + (fullblock ? 1 : 0) * (nexp - btg) * (nes - 1) + 1;
        if (w[tg] > limit && limit != 0) w[tg] = limit;
    }
    if (!frt) w[ntg - 1] = 1;
}

/* ------------------------------------------------------------------ */
/* Shells                                                              */
/* ------------------------------------------------------------------ */

/* The shell rule of src/qvectors/generate_3d_qvectors.m:6-45: half-space
 * integer vectors (z>0, or z==0 and x>0, or x==0,z==0,y>0) with
 * bin = floor((|n| + 0.25) * 2), components bounded by maxc (100 in the
 * generator).  Vectors are written in the generator's loop order.
 * Returns the count; writes at most cap triples. */
int orc_shell_enumerate(int bin, int maxc, int *out, int cap)
{
    int n = 0;
#define ADD(i, j, k) do { double l = sqrt((double)((i)*(i) + (j)*(j) + (k)*(k))); \
        int bb = (int)floor((l + 0.25) * 2.0); \
        if (bb == bin) { if (n < cap) { out[3*n] = (i); out[3*n+1] = (j); out[3*n+2] = (k); } n++; } } while (0)
    for (int j = 0; j <= maxc; j++) ADD(0, j, 0);
    for (int i = 1; i <= maxc; i++)
        for (int j = -maxc; j <= maxc; j++) ADD(i, j, 0);
    for (int k = 1; k <= maxc; k++)
        for (int j = -maxc; j <= maxc; j++)
            for (int i = -maxc; i <= maxc; i++) ADD(i, j, k);
#undef ADD
    return n;
}

/* src/wave_vectors.cpp:150-152: wavegrid_spacing = float(2*PI/min_size),
 * delta = spacing/2; approx_wavenumber[wn] = spacing + delta*wn (:167). */
float orc_wave_spacing(float length_scale) { return (float)(2.0 * REF_PI / length_scale); }
void orc_approx_wavenumbers(float length_scale, int n, float *out)
{
    float spacing = orc_wave_spacing(length_scale);
    float delta = spacing / 2.0f;
    for (int wn = 0; wn < n; wn++) out[wn] = spacing + delta * (float)wn;
}

/* ------------------------------------------------------------------ */
/* FP64 oracle                                                         */
/* ------------------------------------------------------------------ */

static inline double unwrap64(float x, int16_t img, float L) { return (double)x + (double)img * (double)L; }

/* rho_k(t) = sum_j exp(i k.r_j(t)) over wrapped coordinates
 * (src/wave_density.cpp:339-365, show_coordinate at :348), exact
 * k = 2*pi*n/ell in double.  One frame: x,y,z point at that frame's N
 * coordinates.  atoms == NULL means all N atoms in index order.
 * rho[2*v] = Re, rho[2*v+1] = Im for each of the K vectors. */
void orc_rho_f64(int N, const float *x, const float *y, const float *z,
                 const int *atoms, int n_list, int K, const int *nvec,
                 double ell, double *rho)
{
    int cnt = atoms ? n_list : N;
    double c = 2.0 * M_PI / ell;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int v = 0; v < K; v++) {
        double kx = c * nvec[3*v], ky = c * nvec[3*v+1], kz = c * nvec[3*v+2];
        double re = 0.0, im = 0.0;
        for (int a = 0; a < cnt; a++) {
            int j = atoms ? atoms[a] : a;
            double ph = kx * (double)x[j] + ky * (double)y[j] + kz * (double)z[j];
            re += cos(ph);
            im += sin(ph);
        }
        rho[2*v] = re;
        rho[2*v+1] = im;
    }
}

/* S(k) per shell from per-frame rho arrays (src/structure_factor.cpp:
 * 201-285): S[s] = sum_frames sum_{v in s} Re(rho1 conj rho2)/H_s
 *                  / sum_frames N(frame).
 * rho1/rho2: [n_frames][K][2].  natoms: per chosen frame. */
void orc_sk_f64(int n_frames, int K, const double *rho1, const double *rho2,
                const long long *natoms, int n_shells, const int *shell_off, double *S)
{
    long long ntot = 0;
    for (int f = 0; f < n_frames; f++) ntot += natoms[f];
    for (int s = 0; s < n_shells; s++) {
        double acc = 0.0;
        int H = shell_off[s+1] - shell_off[s];
        for (int f = 0; f < n_frames; f++) {
            double part = 0.0;
            for (int v = shell_off[s]; v < shell_off[s+1]; v++) {
                const double *a = rho1 + 2*((size_t)f*K + v);
                const double *b = rho2 + 2*((size_t)f*K + v);
                part += a[0]*b[0] + a[1]*b[1];
            }
            acc += H ? part / (double)H : 0.0;  /* empty shell: the per-vector division never runs */
        }
        S[s] = ntot ? acc / (double)ntot : 0.0;
    }
}

/* F(k,t) (src/intermediate_scattering_function.cpp:237-322, 551-566):
 * F[tg][s] = sum_pairs sum_{v in s} Re(rho1(t0) conj rho2(t1)) / H_s
 *            / sum_pairs N(t0).
 * The reference divides by n_atoms[tg]*w[tg] with n_atoms[tg] =
 * sum_pairs N(t0)/w[tg]; in exact arithmetic w cancels.  rho1/rho2 are
 * [T][K][2] over all frames; natoms[T].  Writes F[ntg][n_shells] and the
 * atom-origin count per time gap. */
void orc_fkt_f64(int T, int K, const double *rho1, const double *rho2,
                 const long long *natoms, int n_pairs, const int *tg, const int *t0,
                 const int *t1, int ntg, int n_shells, const int *shell_off,
                 double *F, double *count)
{
    (void)T;
    for (int i = 0; i < ntg * n_shells; i++) F[i] = 0.0;
    for (int i = 0; i < ntg; i++) count[i] = 0.0;
    for (int p = 0; p < n_pairs; p++) {
        for (int s = 0; s < n_shells; s++) {
            int H = shell_off[s+1] - shell_off[s];
            double part = 0.0;
            for (int v = shell_off[s]; v < shell_off[s+1]; v++) {
                const double *a = rho1 + 2*((size_t)t0[p]*K + v);
                const double *b = rho2 + 2*((size_t)t1[p]*K + v);
                part += a[0]*b[0] + a[1]*b[1];
            }
            F[tg[p]*n_shells + s] += H ? part / (double)H : 0.0;
        }
        count[tg[p]] += (double)natoms[t0[p]];
    }
    for (int g = 0; g < ntg; g++)
        for (int s = 0; s < n_shells; s++)
            F[g*n_shells + s] = count[g] != 0.0 ? F[g*n_shells + s] / count[g] : 0.0;
}

/* Fs(k,t) (src/incoherent_scattering_function.cpp:220-299,
 * src/correlation_2d.cpp:150-164): per pair and atom, mean over the
 * shell's H vectors of cos(k.dr), dr = r(t1)-r(t0) rebuilt exactly in
 * FP64 from wrapped coordinates and image counts; summed over atoms and
 * pairs, divided by the atom-pair count of the time gap.
 * Coordinates: [T][N] SoA; box[T][3].  Writes Fs[ntg][n_shells] and the
 * per-gap atom-pair counts. */
void orc_fs_f64(int N, const float *x, const float *y, const float *z,
                const int16_t *ix, const int16_t *iy, const int16_t *iz, const float *box,
                const int *atoms, int n_list, int n_pairs, const int *tg, const int *t0,
                const int *t1, int ntg, int n_shells, const int *shell_off, const int *nvec,
                double ell, double *Fs, double *count)
{
    int cnt = atoms ? n_list : N;
    double c = 2.0 * M_PI / ell;
    for (int i = 0; i < ntg * n_shells; i++) Fs[i] = 0.0;
    for (int i = 0; i < ntg; i++) count[i] = 0.0;
    for (int p = 0; p < n_pairs; p++) {
        size_t o0 = (size_t)t0[p] * N, o1 = (size_t)t1[p] * N;
        const float *b0 = box + 3*t0[p], *b1 = box + 3*t1[p];
        double *row = Fs + (size_t)tg[p] * n_shells;
#pragma omp parallel for schedule(static) num_threads(g_threads)
        for (int s = 0; s < n_shells; s++) {
            int H = shell_off[s+1] - shell_off[s];
            double acc = 0.0;
            for (int a = 0; a < cnt; a++) {
                int j = atoms ? atoms[a] : a;
                double dx = unwrap64(x[o1+j], ix[o1+j], b1[0]) - unwrap64(x[o0+j], ix[o0+j], b0[0]);
                double dy = unwrap64(y[o1+j], iy[o1+j], b1[1]) - unwrap64(y[o0+j], iy[o0+j], b0[1]);
                double dz = unwrap64(z[o1+j], iz[o1+j], b1[2]) - unwrap64(z[o0+j], iz[o0+j], b0[2]);
                double part = 0.0;
                for (int v = shell_off[s]; v < shell_off[s+1]; v++)
                    part += cos(c * (nvec[3*v]*dx + nvec[3*v+1]*dy + nvec[3*v+2]*dz));
                acc += H ? part / (double)H : 0.0;
            }
            row[s] += acc;
        }
        count[tg[p]] += (double)cnt;
    }
    for (int g = 0; g < ntg; g++)
        for (int s = 0; s < n_shells; s++)
            Fs[g*n_shells + s] = count[g] != 0.0 ? Fs[g*n_shells + s] / count[g] : 0.0;
}

/* MSD and NGP (src/mean_square_displacement.cpp:140-180,
 * src/non_gaussian_parameter.cpp:96-143) in FP64: per time gap the sums
 * of |dr|^2 and |dr|^4 over pairs and atoms, and the atom-pair count.
 * msd = s2/count; alpha2 = 3*s4/(count*5*msd^2) - 1 (NaN at msd = 0,
 * as in the reference). */
void orc_msd_ngp_f64(int N, const float *x, const float *y, const float *z,
                     const int16_t *ix, const int16_t *iy, const int16_t *iz, const float *box,
                     const int *atoms, int n_list, int n_pairs, const int *tg, const int *t0,
                     const int *t1, int ntg, double *s2, double *s4, double *count,
                     double *msd, double *ngp)
{
    int cnt = atoms ? n_list : N;
    for (int g = 0; g < ntg; g++) { s2[g] = 0.0; s4[g] = 0.0; count[g] = 0.0; }
    for (int p = 0; p < n_pairs; p++) {
        size_t o0 = (size_t)t0[p] * N, o1 = (size_t)t1[p] * N;
        const float *b0 = box + 3*t0[p], *b1 = box + 3*t1[p];
        double a2 = 0.0, a4 = 0.0;
        for (int a = 0; a < cnt; a++) {
            int j = atoms ? atoms[a] : a;
            double dx = unwrap64(x[o1+j], ix[o1+j], b1[0]) - unwrap64(x[o0+j], ix[o0+j], b0[0]);
            double dy = unwrap64(y[o1+j], iy[o1+j], b1[1]) - unwrap64(y[o0+j], iy[o0+j], b0[1]);
            double dz = unwrap64(z[o1+j], iz[o1+j], b1[2]) - unwrap64(z[o0+j], iz[o0+j], b0[2]);
            double r2 = dx*dx + dy*dy + dz*dz;
            a2 += r2;
            a4 += r2 * r2;
        }
        s2[tg[p]] += a2;
        s4[tg[p]] += a4;
        count[tg[p]] += (double)cnt;
    }
    for (int g = 0; g < ntg; g++) {
        msd[g] = s2[g] / count[g];
        ngp[g] = 3.0 * s4[g] / (count[g] * 5.0 * msd[g] * msd[g]) - 1.0;
    }
}

/* Displacement sums per time gap in FP64 (Mean_Displacement::listkernel,
 * src/mean_displacement.cpp:163-176; Mean_Square_Displacement_2D::
 * listkernel, src/mean_square_displacement_2d.cpp:144-156, whose plane MSD
 * is the sum of two of the squared components): sum_d[3g+c] += d_c,
 * sum_d2[3g+c] += d_c^2 over the gap's pairs and the atoms. */
void orc_disp_f64(int N, const float *x, const float *y, const float *z,
                  const int16_t *ix, const int16_t *iy, const int16_t *iz, const float *box,
                  const int *atoms, int n_list, int n_pairs, const int *tg, const int *t0,
                  const int *t1, int ntg, double *sum_d, double *sum_d2, double *count)
{
    int cnt = atoms ? n_list : N;
    for (int i = 0; i < 3 * ntg; i++) { sum_d[i] = 0.0; sum_d2[i] = 0.0; }
    for (int g = 0; g < ntg; g++) count[g] = 0.0;
    for (int p = 0; p < n_pairs; p++) {
        size_t o0 = (size_t)t0[p] * N, o1 = (size_t)t1[p] * N;
        const float *b0 = box + 3*t0[p], *b1 = box + 3*t1[p];
        for (int a = 0; a < cnt; a++) {
            int j = atoms ? atoms[a] : a;
            double d[3];
            d[0] = unwrap64(x[o1+j], ix[o1+j], b1[0]) - unwrap64(x[o0+j], ix[o0+j], b0[0]);
            d[1] = unwrap64(y[o1+j], iy[o1+j], b1[1]) - unwrap64(y[o0+j], iy[o0+j], b0[1]);
            d[2] = unwrap64(z[o1+j], iz[o1+j], b1[2]) - unwrap64(z[o0+j], iz[o0+j], b0[2]);
            for (int c = 0; c < 3; c++) {
                sum_d[3*tg[p] + c] += d[c];
                sum_d2[3*tg[p] + c] += d[c] * d[c];
            }
        }
        count[tg[p]] += (double)cnt;
    }
}

/* ------------------------------------------------------------------ */
/* Reference-arithmetic FP32 restatement                               */
/* ------------------------------------------------------------------ */

/* k = float(n) * float spacing, component-wise (src/wave_vectors.cpp:471,
 * Coordinate::operator*= float). */
static inline void kvec32(const int *n, float spacing, float *k)
{
    k[0] = (float)n[0] * spacing;
    k[1] = (float)n[1] * spacing;
    k[2] = (float)n[2] * spacing;
}

/* Wave_Density::listkernel (src/wave_density.cpp:339-365) over one frame:
 * FP32 dot product ((kx*x + ky*y) + kz*z), cosf/sinf, complex<double>
 * accumulation; atoms outer, vectors inner, list order. */
void orc_ref32_rho(int N, const float *x, const float *y, const float *z,
                   const int *atoms, int n_list, int K, const int *nvec,
                   float length_scale, double *rho)
{
    int cnt = atoms ? n_list : N;
    float spacing = orc_wave_spacing(length_scale);
    float *kk = (float *)malloc(sizeof(float) * 3 * (size_t)(K > 0 ? K : 1));
    for (int v = 0; v < K; v++) kvec32(nvec + 3*v, spacing, kk + 3*v);
    /* Each vector's sum runs over the atoms in list order, as in the
     * reference's atom-outer loop; threads (orc_set_threads) take vectors,
     * so the results do not depend on the thread count. */
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int v = 0; v < K; v++) {
        double re = 0.0, im = 0.0;
        for (int a = 0; a < cnt; a++) {
            int j = atoms ? atoms[a] : a;
            float kr = kk[3*v] * x[j] + kk[3*v+1] * y[j] + kk[3*v+2] * z[j];
            re += (double)cosf(kr);
            im += (double)sinf(kr);
        }
        rho[2*v] = re;
        rho[2*v+1] = im;
    }
    free(kk);
}

/* Structure_Factor::timekernel + listkernel (src/structure_factor.cpp:
 * 201-230, 293-320): complex<float> rho per frame, S[wn] +=
 * (re*re+im*im)/float(H) in FP32; postprocess divides by float(n_atoms)
 * (:272-285).  coords [T][N]; frames: the frame indices the time scheme
 * picks (src/analysis_onetime.cpp:22-52). */
void orc_ref32_sk(int N, const float *x, const float *y, const float *z,
                  const int *atoms, int n_list, int n_frames, const int *frames,
                  int n_shells, const int *shell_off, const int *nvec, float length_scale,
                  float *S)
{
    int cnt = atoms ? n_list : N;
    int K = shell_off[n_shells];
    float spacing = orc_wave_spacing(length_scale);
    float *kk = (float *)malloc(sizeof(float) * 3 * (size_t)(K > 0 ? K : 1));
    float *rr = (float *)malloc(sizeof(float) * 2 * (size_t)(K > 0 ? K : 1));
    for (int v = 0; v < K; v++) kvec32(nvec + 3*v, spacing, kk + 3*v);
    for (int s = 0; s < n_shells; s++) S[s] = 0.0f;
    int n_atoms = 0;
    for (int f = 0; f < n_frames; f++) {
        size_t off = (size_t)frames[f] * N;
        for (int v = 0; v < 2*K; v++) rr[v] = 0.0f;
        for (int a = 0; a < cnt; a++) {
            int j = atoms ? atoms[a] : a;
            float cx = x[off+j], cy = y[off+j], cz = z[off+j];
            for (int v = 0; v < K; v++) {
                float kr = kk[3*v] * cx + kk[3*v+1] * cy + kk[3*v+2] * cz;
                rr[2*v] = (float)((double)rr[2*v] + (double)cosf(kr));
                rr[2*v+1] = (float)((double)rr[2*v+1] + (double)sinf(kr));
            }
        }
        for (int s = 0; s < n_shells; s++) {
            int H = shell_off[s+1] - shell_off[s];
            for (int v = shell_off[s]; v < shell_off[s+1]; v++) {
                S[s] += (rr[2*v] * rr[2*v] + rr[2*v+1] * rr[2*v+1]) / (float)H;
            }
        }
        n_atoms += cnt;
    }
    for (int s = 0; s < n_shells; s++) S[s] = (float)n_atoms != 0.0f ? S[s] / (float)n_atoms : 0.0f;
    free(kk);
    free(rr);
}

/* Unwrapped FP32 coordinate as the custom reader forms it
 * (src/system.cpp:1290-1297): float(x + ix*L) computed in double. */
static inline float unwrap32(float x, int16_t img, float L) { return (float)((double)x + (double)img * (double)L); }

/* Self van Hove counts as the reference bins them: Van_Hove_Self::
 * listkernel (src/van_hove_self.cpp:98-102) takes the FP32 length of the
 * difference of the FP32 unwrapped coordinates (src/coordinate.h:55:
 * sqrt(x*x + y*y + z*z)), and Space_Time_Correlation_Function::bin
 * (src/space_time_correlation_function.cpp:216-230) counts it in bin
 * int(distance / bin_size) when 0 <= bin < n_bins, bin_size = max/float(n)
 * (src/van_hove_self.cpp:56).  hist[g*n_bins + b] += 1 (64-bit counts);
 * weighting[g] counts every atom pair. */
void orc_vhs_counts(int N, const float *x, const float *y, const float *z,
                    const int16_t *ix, const int16_t *iy, const int16_t *iz, const float *box,
                    const int *atoms, int n_list, int n_pairs, const int *tg, const int *t0,
                    const int *t1, int ntg, int n_bins, float max_range,
                    unsigned long long *hist, double *weighting)
{
    int cnt = atoms ? n_list : N;
    float bin_size = max_range / (float)n_bins;
    for (size_t i = 0; i < (size_t)ntg * n_bins; i++) hist[i] = 0;
    for (int g = 0; g < ntg; g++) weighting[g] = 0.0;
    for (int p = 0; p < n_pairs; p++) {
        size_t o0 = (size_t)t0[p] * N, o1 = (size_t)t1[p] * N;
        const float *b0 = box + 3*t0[p], *b1 = box + 3*t1[p];
        for (int a = 0; a < cnt; a++) {
            int j = atoms ? atoms[a] : a;
            float dx = unwrap32(x[o1+j], ix[o1+j], b1[0]) - unwrap32(x[o0+j], ix[o0+j], b0[0]);
            float dy = unwrap32(y[o1+j], iy[o1+j], b1[1]) - unwrap32(y[o0+j], iy[o0+j], b0[1]);
            float dz = unwrap32(z[o1+j], iz[o1+j], b1[2]) - unwrap32(z[o0+j], iz[o0+j], b0[2]);
            float distance = sqrtf(dx*dx + dy*dy + dz*dz);
            int bin = (int)(distance / bin_size);
            if (bin >= 0 && bin < n_bins) hist[(size_t)tg[p] * n_bins + bin]++;
            weighting[tg[p]] += 1.0;
        }
    }
}

/* Trajectory_List_Bins::assign_bins (src/trajectory_list_bins.cpp:
 * 949-1096): cell[t*N + j] = (xi*ny + yi)*nz + zi when the stored
 * coordinate is strictly inside (lo, hi) of frame t on every axis, with
 * xi = int((x - lo + FLT_EPSILON*10.0) / (hi - lo + FLT_EPSILON*20.0) *
 * float(nx)) in the reference's types (FP32 differences, double
 * arithmetic after), else -1. */
void orc_bins_assign(int N, int T, const float *x, const float *y, const float *z, int nx, int ny, int nz,
                     const float *lo, const float *hi, int *cell)
{
    const float feps = 1.19209290e-07f;
    int n[3] = {nx, ny, nz};
    for (int t = 0; t < T; t++)
        for (int j = 0; j < N; j++) {
            size_t o = (size_t)t * N + j;
            float c[3] = {x[o], y[o], z[o]};
            int inside = 1, b[3] = {0, 0, 0};
            for (int k = 0; k < 3; k++)
                if (!(c[k] < hi[3*t + k] && c[k] > lo[3*t + k])) inside = 0;
            if (inside)
                for (int k = 0; k < 3; k++) {
                    float l = lo[3*t + k], h = hi[3*t + k];
                    b[k] = (int)(((double)(c[k] - l) + feps * 10.0) / ((double)(h - l) + feps * 20.0) * (double)(float)n[k]);
                }
            cell[o] = inside ? (b[0] * ny + b[1]) * nz + b[2] : -1;
        }
}

/* Distance bins of Trajectory_List_Bins: point (kind 0,
 * assign_bins_distance_point, src/trajectory_list_bins.cpp:1321-1488),
 * plane (kind 1, assign_bins_distance_plane, :1489-1774) and distance to a
 * second list (kind 2, assign_bins_distance_clusters, :1097-1320).  The
 * binned atoms of frame t are idx[off[t] .. off[t+1]-1], the second list's
 * cidx[coff[t] .. coff[t+1]-1].  FP32 in the reference's types and order:
 * Coordinate arithmetic is per component in float, box_size*.5 is a float
 * product for the point form (Coordinate * float) and a double one for the
 * plane form (float * double literal), Coordinate::length is
 * sqrt((x*x + y*y) + z*z) in float, vector_unwrapped (src/coordinate.cpp:
 * 98-124) uses the box of frame 0.  cell[t*N + j] = int(length / w) when
 * that is in [0, n_bins), else -1. */
void orc_bins_distance(int N, int T, const float *x, const float *y, const float *z, const float *box, int kind,
                       const float *point, int axis, int dir, const int *off, const int *idx, const int *coff,
                       const int *cidx, float w, int n_bins, int *cell)
{
    for (size_t o = 0; o < (size_t)N * T; o++) cell[o] = -1;
    for (int t = 0; t < T; t++) {
        const float *L = box + 3 * t;
        const float *r[3] = {x + (size_t)t * N, y + (size_t)t * N, z + (size_t)t * N};
        for (int a = off[t]; a < off[t + 1]; a++) {
            const int j = idx[a];
            float len = 0.0f;
            if (kind == 0) {
                float d[3];
                for (int k = 0; k < 3; k++) {
                    d[k] = r[k][j] - point[k];
                    const float h = L[k] * 0.5f;
                    const float q = d[k] / h;
                    const float im = (float)(int)q;
                    d[k] -= L[k] * im;
                }
                len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            } else if (kind == 1) {
                const float Lk = L[axis];
                const float d = r[axis][j] - point[0];
                if (dir == 0) len = fabsf(Lk * (float)(int)((double)d / ((double)Lk * .5)) - d);
                else if (dir == 1) len = d < 0 ? Lk + d : d;
                else len = d < 0 ? fabsf(d) : Lk - d;
            } else {
                len = L[0];
                if (L[1] > len) len = L[1];
                if (L[2] > len) len = L[2];
                for (int c = coff[t]; c < coff[t + 1]; c++) {
                    const int k = cidx[c];
                    float d[3];
                    for (int m = 0; m < 3; m++) {
                        d[m] = r[m][j] - r[m][k];
                        if (d[m] > box[m] / 2) d[m] = d[m] - box[m];
                        else if (d[m] < -box[m] / 2) d[m] = d[m] + box[m];
                    }
                    const float l = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                    if (l < len) len = l;
                }
            }
            const int b = (int)(len / w);
            if (b >= 0 && b < n_bins) cell[(size_t)t * N + j] = b;
        }
    }
}

/* DebyeWaller_Dist / Displacement_Distribution (src/debyewaller_dist.cpp:
 * 47-95, src/displacement_distribution.cpp:50-98) over given pairs: per
 * pair, every atom of the list (all atoms when atoms is NULL) adds
 * msd = pow(distance, 2) stored as float, with distance the FP32 length of
 * the FP32 unwrapped difference, to bin int(msd / binsize), binsize =
 * max_value / n_bins in FP32, bins >= n_bins + 1 clamped to the last
 * (overflow) bin.  counts[n_bins + 1]; sum_d2, sum_d4 in FP64 of the FP32
 * msd; *count = atom pairs.  The reference accumulates distribution, mean
 * and square_term in FP32 with 1/weighting factors that cancel. */
void orc_displacement_dist(int N, const float *x, const float *y, const float *z,
                           const int16_t *ix, const int16_t *iy, const int16_t *iz, const float *box,
                           const int *atoms, int n_list, int n_pairs, const int *t0, const int *t1,
                           int n_bins, float max_value, unsigned long long *counts, double *sum_d2,
                           double *sum_d4, double *count)
{
    int cnt = atoms ? n_list : N;
    int nb = n_bins + 1;
    float binsize = max_value / (float)n_bins;
    for (int b = 0; b < nb; b++) counts[b] = 0;
    *sum_d2 = *sum_d4 = *count = 0.0;
    for (int p = 0; p < n_pairs; p++) {
        size_t o0 = (size_t)t0[p] * N, o1 = (size_t)t1[p] * N;
        const float *b0 = box + 3*t0[p], *b1 = box + 3*t1[p];
        for (int a = 0; a < cnt; a++) {
            int j = atoms ? atoms[a] : a;
            float dx = unwrap32(x[o1+j], ix[o1+j], b1[0]) - unwrap32(x[o0+j], ix[o0+j], b0[0]);
            float dy = unwrap32(y[o1+j], iy[o1+j], b1[1]) - unwrap32(y[o0+j], iy[o0+j], b0[1]);
            float dz = unwrap32(z[o1+j], iz[o1+j], b1[2]) - unwrap32(z[o0+j], iz[o0+j], b0[2]);
            float msd = (float)pow((double)sqrtf(dx*dx + dy*dy + dz*dz), 2.0);
            int index = (int)(msd / binsize);
            if (index >= nb) index = nb - 1;
            counts[index]++;
            *sum_d2 += (double)msd;
            *sum_d4 += (double)msd * (double)msd;
            *count += 1.0;
        }
    }
}

/* Fast_Particles / Slow_Particles / Average_Particles (src/fast_particles.
 * cpp:117-139, src/slow_particles.cpp:93-113, src/average_particles.cpp:
 * 95-116).  Pairs of System::displacement_list(analysis, time_index, false)
 * (src/system.cpp:4349-4398); for each, the atoms of the input list at the
 * list time of t0 (in_n_times == 0: static list in_idx[in_n], or all atoms
 * when in_idx is NULL; otherwise in_conv[t] and in_off/in_idx) are tested
 * with square_displacement = pow(distance, 2) stored as float, distance =
 * sqrtf of the FP32 squared length of the difference of FP32 unwrapped
 * coordinates (src/trajectory.cpp:326-330, src/coordinate.h:55), and
 * appended to block int((float(t0) - float(frt)) / float(nes)) when
 * mode 0: d2 > lo (Fast), 1: d2 < hi (Slow), 2: d2 <= hi && d2 >= lo
 * (Average).  Output: off[nexp+1], idx (cap entries), conv[T] =
 * int(float(t - frt) / float(nes)) (src/exptime_trajectory_list.cpp:87-91).
 * Returns the entry count, or -1 for a time gap out of range. */
long orc_select_displacement(int N, int T, const float *x, const float *y, const float *z,
                             const int16_t *ix, const int16_t *iy, const int16_t *iz, const float *box,
                             int in_n_times, const int *in_conv, const int *in_off, const int *in_idx, int in_n,
                             int nexp, int nes, int frt, int limit, int time_index, int mode, float lo, float hi,
                             int *off, int *idx, long cap, int *conv)
{
    int f = frt ? 1 : 0;
    int ntg = nes + nexp - 1 + (1 - f);
    int g = time_index;
    int np = 0, *p0 = malloc(sizeof(int) * (nexp + 1)), *p1 = malloc(sizeof(int) * (nexp + 1));
    if (g >= 0 && g < nes) {
        int count = 0;
        for (int b = 0; b < nexp; b++) {
            p0[np] = nes * b + f; p1[np] = nes * b + f + g; np++;
            count++;
            if (count == limit) break;
        }
    } else if (g >= nes && g < ntg - 1 + f) {
        int btg = g - nes + 1, count = 0;
        for (int e = 0; e < 1; e++) {
            for (int b = 0; b < nexp - btg; b++) {
                p0[np] = nes * b + e + f; p1[np] = p0[np] + nes * btg; np++;
                count++;
                if (count == limit) break;
            }
            if (count == limit) break;
        }
    } else if (g == ntg - 1 + f) {
        p0[np] = 0; p1[np] = nes * nexp; np++;
    } else {
        free(p0); free(p1);
        return -1;
    }
    /* per-block lists, appended in pair order */
    long *bn = calloc(nexp, sizeof(long));
    int **bl = calloc(nexp, sizeof(int *));
    for (int q = 0; q < np; q++) {
        int t0 = p0[q], t1 = p1[q];
        int expindex = (int)(((float)t0 - (float)f) / (float)nes);
        const int *atoms = in_idx;
        int cnt;
        if (in_n_times == 0) {
            cnt = in_idx ? in_n : N;
        } else {
            int c = in_conv[t0];
            if (c < 0 || c >= in_n_times) { cnt = 0; atoms = in_idx; }
            else { atoms = in_idx + in_off[c]; cnt = in_off[c + 1] - in_off[c]; }
        }
        size_t o0 = (size_t)t0 * N, o1 = (size_t)t1 * N;
        const float *b0 = box + 3*t0, *b1 = box + 3*t1;
        for (int a = 0; a < cnt; a++) {
            int j = atoms ? atoms[a] : a;
            float dx = unwrap32(x[o1+j], ix[o1+j], b1[0]) - unwrap32(x[o0+j], ix[o0+j], b0[0]);
            float dy = unwrap32(y[o1+j], iy[o1+j], b1[1]) - unwrap32(y[o0+j], iy[o0+j], b0[1]);
            float dz = unwrap32(z[o1+j], iz[o1+j], b1[2]) - unwrap32(z[o0+j], iz[o0+j], b0[2]);
            float distance = sqrtf(dx*dx + dy*dy + dz*dz);
            float square_displacement = (float)pow((double)distance, 2.0);
            int keep;
            if (mode == 0) keep = square_displacement > lo;
            else if (mode == 1) keep = square_displacement < hi;
            else keep = square_displacement <= hi && square_displacement >= lo;
            if (keep) {
                bl[expindex] = realloc(bl[expindex], sizeof(int) * (bn[expindex] + 1));
                bl[expindex][bn[expindex]++] = j;
            }
        }
    }
    long total = 0;
    off[0] = 0;
    for (int b = 0; b < nexp; b++) {
        for (long i = 0; i < bn[b]; i++) {
            if (total < cap) idx[total] = bl[b][i];
            total++;
        }
        off[b + 1] = (int)total;
        free(bl[b]);
    }
    for (int t = 0; t < T; t++) conv[t] = (int)((float)(t - f) / (float)nes);
    free(bl); free(bn); free(p0); free(p1);
    return total;
}

/* Incoherent_Scattering_Function::listkernel (src/incoherent_scattering_
 * function.cpp:272-299) for one pair: per atom and shell,
 * sum_v double(cosf(k.(u1-u0)))/double(H) in double, then
 * correlation[s] += float(...) in FP32.  corr[n_shells] accumulates across
 * calls; returns the atom count added to n_atoms[tg]. */
int orc_ref32_fs_pair(int N, const float *x, const float *y, const float *z,
                      const int16_t *ix, const int16_t *iy, const int16_t *iz, const float *box,
                      const int *atoms, int n_list, int t0, int t1, int n_shells,
                      const int *shell_off, const int *nvec, float length_scale, float *corr)
{
    int cnt = atoms ? n_list : N;
    int K = shell_off[n_shells];
    float spacing = orc_wave_spacing(length_scale);
    float *kk = (float *)malloc(sizeof(float) * 3 * (size_t)(K > 0 ? K : 1));
    for (int v = 0; v < K; v++) kvec32(nvec + 3*v, spacing, kk + 3*v);
    size_t o0 = (size_t)t0 * N, o1 = (size_t)t1 * N;
    const float *b0 = box + 3*t0, *b1 = box + 3*t1;
    /* Each shell's sum runs over the atoms in list order, as in the
     * reference's atom-outer loop; threads take shells. */
#pragma omp parallel for schedule(dynamic) num_threads(g_threads)
    for (int s = 0; s < n_shells; s++) {
        int H = shell_off[s+1] - shell_off[s];
        for (int a = 0; a < cnt; a++) {
            int j = atoms ? atoms[a] : a;
            float dx = unwrap32(x[o1+j], ix[o1+j], b1[0]) - unwrap32(x[o0+j], ix[o0+j], b0[0]);
            float dy = unwrap32(y[o1+j], iy[o1+j], b1[1]) - unwrap32(y[o0+j], iy[o0+j], b0[1]);
            float dz = unwrap32(z[o1+j], iz[o1+j], b1[2]) - unwrap32(z[o0+j], iz[o0+j], b0[2]);
            double t = 0.0;
            for (int v = shell_off[s]; v < shell_off[s+1]; v++) {
                float kr = kk[3*v] * dx + kk[3*v+1] * dy + kk[3*v+2] * dz;
                t += (double)cosf(kr) / (double)H;
            }
            corr[s] = (float)((double)corr[s] + (double)(float)t);
        }
    }
    free(kk);
    return cnt;
}

/* Mean_Square_Displacement / Non_Gaussian_Parameter listkernels for one
 * pair (src/mean_square_displacement.cpp:164-169,
 * src/non_gaussian_parameter.cpp:125-129): d = sqrtf(|u1-u0|^2) in FP32,
 * msd += d*d (FP32), ngp += pow(d,4) (double, stored to FP32). */
int orc_ref32_msd_pair(int N, const float *x, const float *y, const float *z,
                       const int16_t *ix, const int16_t *iy, const int16_t *iz, const float *box,
                       const int *atoms, int n_list, int t0, int t1, float *msd, float *ngp)
{
    int cnt = atoms ? n_list : N;
    size_t o0 = (size_t)t0 * N, o1 = (size_t)t1 * N;
    const float *b0 = box + 3*t0, *b1 = box + 3*t1;
    for (int a = 0; a < cnt; a++) {
        int j = atoms ? atoms[a] : a;
        float dx = unwrap32(x[o1+j], ix[o1+j], b1[0]) - unwrap32(x[o0+j], ix[o0+j], b0[0]);
        float dy = unwrap32(y[o1+j], iy[o1+j], b1[1]) - unwrap32(y[o0+j], iy[o0+j], b0[1]);
        float dz = unwrap32(z[o1+j], iz[o1+j], b1[2]) - unwrap32(z[o0+j], iz[o0+j], b0[2]);
        float d = sqrtf(dx * dx + dy * dy + dz * dz);
        *msd += d * d;
        *ngp = (float)((double)*ngp + pow((double)d, 4.0));
    }
    return cnt;
}

/* ------------------------------------------------------------------ */
/* S(k) misfit loss and its gradient (no reference counterpart)        */
/* ------------------------------------------------------------------ */

/* Structure-matching loss on one frame,
 *   L = sum_s w_s (S_s - T_s)^2,   S_s = sum_{v in s} |rho_v|^2 / (H_s N),
 * with rho_v = sum_j exp(i k_v.r_j), k_v = 2 pi n_v / ell, and its gradient
 *   dL/dr_j = sum_v c_v k_v Im(conj(rho_v) e^{i k_v.r_j}),
 *   c_v = -4 w_s (S_s - T_s) / (H_s N)   for v in shell s.
 * (d|rho|^2/dr_j = 2 Re(conj(rho) i k e^{ik.r_j}) = -2 k Im(conj(rho) e^{ik.r_j}).)
 * Pinned by central finite differences of this same loss
 * (tests/test_oracle_adjoint.py).  S[n_shells] receives S_s. */
double orc_sk_loss_grad_f64(int N, const double *x, const double *y, const double *z, int n_shells,
                            const int *shell_off, const int *nvec, double ell, const double *w,
                            const double *target, double *S, double *gx, double *gy, double *gz)
{
    int K = shell_off[n_shells];
    double c = 2.0 * M_PI / ell;
    double *rho = (double *)malloc(sizeof(double) * 2 * (size_t)(K > 0 ? K : 1));
    double *cv = (double *)malloc(sizeof(double) * (size_t)(K > 0 ? K : 1));
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int v = 0; v < K; v++) {
        double kx = c * nvec[3*v], ky = c * nvec[3*v+1], kz = c * nvec[3*v+2];
        double re = 0.0, im = 0.0;
        for (int j = 0; j < N; j++) {
            double ph = kx * x[j] + ky * y[j] + kz * z[j];
            re += cos(ph);
            im += sin(ph);
        }
        rho[2*v] = re;
        rho[2*v+1] = im;
    }
    double loss = 0.0;
    for (int s = 0; s < n_shells; s++) {
        int H = shell_off[s+1] - shell_off[s];
        double acc = 0.0;
        for (int v = shell_off[s]; v < shell_off[s+1]; v++) acc += rho[2*v]*rho[2*v] + rho[2*v+1]*rho[2*v+1];
        S[s] = H ? acc / ((double)H * N) : 0.0;
        double d = S[s] - target[s];
        loss += w[s] * d * d;
        for (int v = shell_off[s]; v < shell_off[s+1]; v++) cv[v] = -4.0 * w[s] * d / ((double)H * N);
    }
    if (gx) {
#pragma omp parallel for schedule(static) num_threads(g_threads)
        for (int j = 0; j < N; j++) {
            double ax = 0.0, ay = 0.0, az = 0.0;
            for (int v = 0; v < K; v++) {
                double kx = c * nvec[3*v], ky = c * nvec[3*v+1], kz = c * nvec[3*v+2];
                double ph = kx * x[j] + ky * y[j] + kz * z[j];
                double s = sin(ph), co = cos(ph);
                /* Im(conj(rho) e^{i ph}) = re*sin - im*cos */
                double t = cv[v] * (rho[2*v] * s - rho[2*v+1] * co);
                ax += kx * t;
                ay += ky * t;
                az += kz * t;
            }
            gx[j] = ax;
            gy[j] = ay;
            gz[j] = az;
        }
    }
    free(rho);
    free(cv);
    return loss;
}
