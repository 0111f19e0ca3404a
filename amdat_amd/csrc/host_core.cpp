// host_core.cpp -- host-side integer work of the C ABI: the blocked-
// exponential time scheme, time-origin pairs and weights, and shell
// tables (file reader and rule-based generator).  No device code.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <array>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "../../include/amdat_hip.h"
#include "abi_internal.h"

using namespace amdat;

extern "C" {

int amdat_version(void) { return 1 * 10000 + 0 * 100 + 0; }

static int amdat_ts_n_timesteps_body(int nexp, int nes)
{
    if (nexp < 1 || nes < 0) return set_error(AMDAT_EINVAL, "time scheme: n_exponentials must be >= 1 and n_exponential_steps >= 0");
    return nes * nexp + 1;
}

int amdat_ts_n_timesteps(int nexp, int nes)
{
    AMDAT_GUARD("amdat_ts_n_timesteps", amdat_ts_n_timesteps_body(nexp, nes))
}

static int amdat_ts_n_timegaps_body(int nexp, int nes, int frt)
{
    if (nexp < 1 || nes < 0) return set_error(AMDAT_EINVAL, "time scheme: n_exponentials must be >= 1 and n_exponential_steps >= 0");
    return nes + nexp - 1 + (frt ? 0 : 1);
}

int amdat_ts_n_timegaps(int nexp, int nes, int frt)
{
    AMDAT_GUARD("amdat_ts_n_timegaps", amdat_ts_n_timegaps_body(nexp, nes, frt))
}

// System::create_timelist, src/system.cpp:3542-3568.  exp_base is a float
// member; pow(float, int) is evaluated in double.
static int amdat_ts_timelist_body(int nexp, int nes, float base, int first_exp, float dt, float *out, int cap)
{
    if (!out || nexp < 1 || nes < 0) return set_error(AMDAT_EINVAL, "amdat_ts_timelist: bad arguments");
    int T = nes * nexp + 1;
    if (cap < T) return set_error(AMDAT_EINVAL, "amdat_ts_timelist: output capacity smaller than n_timesteps");
    int ti = 0;
    float block_start = 0.0f;
    out[0] = 0.0f;
    for (int b = 0; b < nexp; b++) {
        for (int e = 1; e <= nes; e++) {
            ++ti;
            double p = std::pow((double)base, (double)(e - 1 + first_exp));
            out[ti] = (p <= (double)e) ? block_start + (float)e * dt
                                       : block_start + (float)std::floor(p) * dt;
        }
        block_start = out[ti];
    }
    return AMDAT_OK;
}

int amdat_ts_timelist(int nexp, int nes, float base, int first_exp, float dt, float *out, int cap)
{
    AMDAT_GUARD("amdat_ts_timelist", amdat_ts_timelist_body(nexp, nes, base, first_exp, dt, out, cap))
}

// System::displacement_times, src/system.cpp:3763-3784.
static int amdat_ts_displacement_times_body(int nexp, int nes, int frt, const float *tl, float *out, int cap)
{
    if (!tl || !out || nexp < 1 || nes < 0) return set_error(AMDAT_EINVAL, "amdat_ts_displacement_times: bad arguments");
    int ntg = nes + nexp - 1 + (frt ? 0 : 1);
    if (cap < ntg) return set_error(AMDAT_EINVAL, "amdat_ts_displacement_times: output capacity smaller than n_timegaps");
    int T = nes * nexp + 1;
    int f = frt ? 1 : 0;
    for (int i = 0; i < nes; i++) out[i] = tl[i + f] - tl[f];
    for (int i = 1; i < nexp; i++) out[i + nes - 1] = tl[nes * i + f] - tl[f];
    if (!frt) out[ntg - 1] = tl[T - 1] - tl[0];
    return AMDAT_OK;
}

int amdat_ts_displacement_times(int nexp, int nes, int frt, const float *tl, float *out, int cap)
{
    AMDAT_GUARD("amdat_ts_displacement_times", amdat_ts_displacement_times_body(nexp, nes, frt, tl, out, cap))
}

// System::displacement_list, src/system.cpp:4090-4179.  The pair order is
// the reference's: in-block gaps by gap then block, cross-block gaps by gap,
// block offset, block; then the (0, last) pair when frt == 0.
static int amdat_ts_pairs_body(int nexp, int nes, int frt, int fullblock, int limit, int *tg, int *t0, int *t1,
                   int cap, int *n_pairs)
{
    if (nexp < 1 || nes < 0 || limit < 0) return set_error(AMDAT_EINVAL, "amdat_ts_pairs: bad time scheme");
    if (cap > 0 && (!tg || !t0 || !t1)) return set_error(AMDAT_EINVAL, "amdat_ts_pairs: null output");
    int f = frt ? 1 : 0;
    int ntg = nes + nexp - 1 + (1 - f);
    int T = nes * nexp + 1;
    long n = 0;
    auto emit = [&](int g, int a, int b) {
        if (n < cap) { tg[n] = g; t0[n] = a; t1[n] = b; }
        ++n;
    };
    for (int g = 0; g < nes; g++) {
        int count = 0;
        for (int b = 0; b < nexp; b++) {
            int a = nes * b + f;
            emit(g, a, a + g);
            if (++count == limit && limit != 0) break;
        }
    }
    for (int g = nes; g < ntg - 1 + f; g++) {
        int count = 0;
        bool stop = false;
        int btg = g - nes + 1;
        int n_off = (fullblock ? 1 : 0) * (nes - 1) + 1;
        for (int e = 0; e < n_off && !stop; e++) {
            for (int b = 0; b < nexp - btg; b++) {
                int a = nes * b + e + f;
                emit(g, a, a + nes * btg);
                if (++count == limit && limit != 0) { stop = true; break; }
            }
        }
    }
    if (!frt) emit(ntg - 1, 0, T - 1);
    if (n > 0x7fffffffL) return set_error(AMDAT_EINVAL, "amdat_ts_pairs: too many pairs");
    if (n_pairs) *n_pairs = (int)n;
    return AMDAT_OK;
}

int amdat_ts_pairs(int nexp, int nes, int frt, int fullblock, int limit, int *tg, int *t0, int *t1,
                   int cap, int *n_pairs)
{
    AMDAT_GUARD("amdat_ts_pairs", amdat_ts_pairs_body(nexp, nes, frt, fullblock, limit, tg, t0, t1, cap, n_pairs))
}

// System::displacement_list(analysis, timegap, false), src/system.cpp:
// 4349-4398: the pairs of one time gap with the first time of each block
// only (in-block gaps: every block; cross-block gaps: blocks that reach;
// the last gap: frames 0 and n_timesteps - 1, which with frt = 1 is time
// gap n_timegaps, one past the full enumeration).  limit > 0 keeps the
// first limit pairs.
static int amdat_ts_gap_pairs_body(int nexp, int nes, int frt, int limit, int time_index, int *tg, int *t0, int *t1, int cap,
                       int *n_pairs)
{
    if (nexp < 1 || nes < 1 || limit < 0) return set_error(AMDAT_EINVAL, "amdat_ts_gap_pairs: bad time scheme");
    if (cap > 0 && (!tg || !t0 || !t1)) return set_error(AMDAT_EINVAL, "amdat_ts_gap_pairs: null output");
    const int f = frt ? 1 : 0;
    const int ntg = nes + nexp - 1 + (1 - f);
    const int g = time_index;
    int n = 0;
    auto emit = [&](int gg, int a, int b) {
        if (n < cap) { tg[n] = gg; t0[n] = a; t1[n] = b; }
        ++n;
    };
    if (g >= 0 && g < nes) {
        int count = 0;
        for (int b = 0; b < nexp; b++) {
            emit(g, nes * b + f, nes * b + f + g);
            if (++count == limit) break;
        }
    } else if (g >= nes && g < ntg - 1 + f) {
        const int btg = g - nes + 1;
        int count = 0;
        for (int b = 0; b < nexp - btg; b++) {
            emit(g, nes * b + f, nes * b + f + nes * btg);
            if (++count == limit) break;
        }
    } else if (g == ntg - 1 + f) {
        emit(ntg - 1, 0, nes * nexp);
    } else {
        return set_error(AMDAT_ERANGE, "Requested timegap out of range.");
    }
    if (n_pairs) *n_pairs = n;
    return AMDAT_OK;
}

int amdat_ts_gap_pairs(int nexp, int nes, int frt, int limit, int time_index, int *tg, int *t0, int *t1, int cap,
                       int *n_pairs)
{
    AMDAT_GUARD("amdat_ts_gap_pairs", amdat_ts_gap_pairs_body(nexp, nes, frt, limit, time_index, tg, t0, t1, cap, n_pairs))
}

// System::timegap_weighting, src/system.cpp:4599-4628 (cross-block entries
// carry the reference's +1).
static int amdat_ts_weighting_body(int nexp, int nes, int frt, int fullblock, int limit, int *w, int cap)
{
    if (!w || nexp < 1 || nes < 0) return set_error(AMDAT_EINVAL, "amdat_ts_weighting: bad arguments");
    int f = frt ? 1 : 0;
    int ntg = nes + nexp - 1 + (1 - f);
    if (cap < ntg) return set_error(AMDAT_EINVAL, "amdat_ts_weighting: output capacity smaller than n_timegaps");
    for (int g = 0; g < nes; g++) {
        w[g] = nexp;
        if (w[g] > limit && limit != 0) w[g] = limit;
    }
    for (int g = nes; g < ntg - 1 + f; g++) {
        int btg = g - nes + 1;
        w[g] = (nexp - btg) + (fullblock ? 1 : 0) * (nexp - btg) * (nes - 1) + 1;
        if (w[g] > limit && limit != 0) w[g] = limit;
    }
    if (!frt) w[ntg - 1] = 1;
    return AMDAT_OK;
}

int amdat_ts_weighting(int nexp, int nes, int frt, int fullblock, int limit, int *w, int cap)
{
    AMDAT_GUARD("amdat_ts_weighting", amdat_ts_weighting_body(nexp, nes, frt, fullblock, limit, w, cap))
}

// Wave_Vectors::read_vectors, src/wave_vectors.cpp:430-484: one file per
// bin, whitespace-separated numbers, three per vector.  Numbers must be
// integral (the reference scales them by the grid spacing).
// Shell files stem.%03d for bins first_bin .. first_bin+n_bins-1: whitespace
// separated number triples (t0, t1, t2).  Each triple is placed on the axes
// by the plane as the reference's readers place it, times `scale`, and must
// then be an integer within +-127.
static int read_shell_files(const char *stem, const char *plane, int first_bin, int n_bins, int *shell_off, int *nvec,
                            int cap_vectors, int *n_vectors, int *denominator)
{
    if (!stem || !plane || !shell_off || n_bins < 1 || first_bin < 0 || first_bin + n_bins - 1 > 999)
        return set_error(AMDAT_EINVAL, "amdat_shells_read: bad arguments");
    // axis of each file column (-1: ignored), and the scale that makes the
    // entries integers: src/wave_vectors.cpp:314-331 (2D), 394-413 (1D)
    int ax[3];
    double scale = 1.0;
    const std::string pl(plane);
    if (pl == "xyz" || pl == "xy") { ax[0] = 0; ax[1] = 1; ax[2] = 2; }
    else if (pl == "xz") { ax[0] = 0; ax[1] = 2; ax[2] = 1; }
    else if (pl == "yz") { ax[0] = 1; ax[1] = 2; ax[2] = 0; }
    else if (pl == "x" || pl == "y" || pl == "z") {
        ax[0] = pl == "x" ? 0 : (pl == "y" ? 1 : 2);
        ax[1] = ax[2] = -1;
        scale = 2.0;  // 1D shells hold half-integers (bin/2)
    } else {
        return set_error(AMDAT_EINVAL, "plane command not recognized. valid inputs are: x, y, z, xy, xz, yz, xyz");
    }
    if (denominator) *denominator = (int)scale;
    int n = 0;
    shell_off[0] = 0;
    for (int s = 0; s < n_bins; s++) {
        char suffix[16];
        std::snprintf(suffix, sizeof(suffix), ".%03d", first_bin + s);
        std::string path = std::string(stem) + suffix;
        FILE *fp = std::fopen(path.c_str(), "r");
        if (!fp) return set_error(AMDAT_EIO, "Error opening vector file " + path);
        double v[3];
        int c = 0;
        while (std::fscanf(fp, "%lf", &v[c]) == 1) {
            if (++c == 3) {
                int out[3] = {0, 0, 0};
                for (int d = 0; d < 3; d++) {
                    if (ax[d] < 0) continue;
                    const double w = v[d] * scale;
                    if (w != std::floor(w) || std::fabs(w) > 127.0) {
                        std::fclose(fp);
                        return set_error(AMDAT_EIO, "vector component not representable on the integer grid in " + path);
                    }
                    out[ax[d]] = (int)w;
                }
                if (n < cap_vectors && nvec)
                    for (int d = 0; d < 3; d++) nvec[3 * n + d] = out[d];
                ++n;
                c = 0;
            }
        }
        std::fclose(fp);
        if (c != 0) return set_error(AMDAT_EIO, "vector file " + path + " has a trailing partial vector");
        shell_off[s + 1] = n;
    }
    if (n_vectors) *n_vectors = n;
    if (n > cap_vectors) return set_error(AMDAT_EINVAL, "amdat_shells_read: vector capacity too small");
    return AMDAT_OK;
}

static int amdat_shells_read_body(const char *stem, int first_bin, int n_bins, int *shell_off, int *nvec,
                      int cap_vectors, int *n_vectors)
{
    return read_shell_files(stem, "xyz", first_bin, n_bins, shell_off, nvec, cap_vectors, n_vectors, nullptr);
}

int amdat_shells_read(const char *stem, int first_bin, int n_bins, int *shell_off, int *nvec,
                      int cap_vectors, int *n_vectors)
{
    AMDAT_GUARD("amdat_shells_read", amdat_shells_read_body(stem, first_bin, n_bins, shell_off, nvec, cap_vectors, n_vectors))
}

static int amdat_shells_read_plane_body(const char *stem, const char *plane, int first_bin, int n_bins, int *shell_off,
                            int *nvec, int cap_vectors, int *n_vectors, int *denominator)
{
    return read_shell_files(stem, plane, first_bin, n_bins, shell_off, nvec, cap_vectors, n_vectors, denominator);
}

int amdat_shells_read_plane(const char *stem, const char *plane, int first_bin, int n_bins, int *shell_off,
                            int *nvec, int cap_vectors, int *n_vectors, int *denominator)
{
    AMDAT_GUARD("amdat_shells_read_plane", amdat_shells_read_plane_body(stem, plane, first_bin, n_bins, shell_off, nvec, cap_vectors, n_vectors, denominator))
}

namespace {
inline uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
}  // namespace

// The shell rule of src/qvectors/generate_3d_qvectors.m:6-64: half-space
// integer vectors with bin = floor((|n|+0.25)*2), components within +-100,
// at most max_per_shell per shell by a seeded shuffle.  The reference
// ships one fixed random choice; this generator makes a reproducible one
// of the same sizes.
static int amdat_shells_generate_body(int first_bin, int n_bins, int max_per_shell, uint64_t seed,
                          int *shell_off, int *nvec, int cap_vectors, int *n_vectors)
{
    if (!shell_off || n_bins < 1 || first_bin < 1 || first_bin + n_bins - 1 > 300 || max_per_shell < 1)
        return set_error(AMDAT_EINVAL, "amdat_shells_generate: bad arguments");
    const int maxc = 100;
    int last_bin = first_bin + n_bins - 1;
    std::vector<std::vector<int>> bins(n_bins);
    auto add = [&](int i, int j, int k) {
        double l = std::sqrt((double)(i * i + j * j + k * k));
        int b = (int)std::floor((l + 0.25) * 2.0);
        if (b >= first_bin && b <= last_bin) {
            auto &v = bins[b - first_bin];
            v.push_back(i); v.push_back(j); v.push_back(k);
        }
    };
    double rmax = last_bin / 2.0 + 0.25;
    int lim = std::min(maxc, (int)std::ceil(rmax));
    for (int j = 0; j <= lim; j++) add(0, j, 0);
    for (int i = 1; i <= lim; i++)
        for (int j = -lim; j <= lim; j++) add(i, j, 0);
    for (int k = 1; k <= lim; k++)
        for (int j = -lim; j <= lim; j++)
            for (int i = -lim; i <= lim; i++) add(i, j, k);
    uint64_t st = seed;
    int n = 0;
    shell_off[0] = 0;
    for (int s = 0; s < n_bins; s++) {
        auto &v = bins[s];
        int cnt = (int)v.size() / 3;
        // Fisher-Yates on the vector order, then keep the first max_per_shell.
        for (int a = cnt - 1; a > 0; a--) {
            int b = (int)(splitmix64(st) % (uint64_t)(a + 1));
            for (int d = 0; d < 3; d++) std::swap(v[3 * a + d], v[3 * b + d]);
        }
        int keep = std::min(cnt, max_per_shell);
        for (int a = 0; a < keep; a++) {
            if (n < cap_vectors && nvec)
                for (int d = 0; d < 3; d++) nvec[3 * n + d] = v[3 * a + d];
            ++n;
        }
        shell_off[s + 1] = n;
    }
    if (n_vectors) *n_vectors = n;
    if (n > cap_vectors) return set_error(AMDAT_EINVAL, "amdat_shells_generate: vector capacity too small");
    return AMDAT_OK;
}

int amdat_shells_generate(int first_bin, int n_bins, int max_per_shell, uint64_t seed,
                          int *shell_off, int *nvec, int cap_vectors, int *n_vectors)
{
    AMDAT_GUARD("amdat_shells_generate", amdat_shells_generate_body(first_bin, n_bins, max_per_shell, seed, shell_off, nvec, cap_vectors, n_vectors))
}

}  // extern "C"

// ---- block layout of k-vector plans -------------------------------------

namespace amdat {

namespace {

constexpr int kLanesPerPass = 32;  // a wave's 8-byte LDS reads are served 32 lanes at a time
constexpr int kBankPairs = 32;     // 64 banks of 4 bytes = 32 eight-byte entries

struct Vec {
    int nx, ny, nz, d;
    bool sw = false;  // the group shares nz: the shared read is z, the own read y
};

// Table entries a slot reads besides x: the one its group shares (y, or z
// when swapped) and its own (z, or y).  Their bank pairs follow the entry
// index: every lane of a read addresses the same atom row.
struct Entries {
    int oy, oz;
    int shared(const Vec &v) const { return v.sw ? oz + v.nz : oy + v.ny; }
    int own(const Vec &v) const { return v.sw ? oy + v.ny : oz + v.nz; }
};

// Two-bit group code of a wave whose groups hold GS slots.
int gs_code(int gs) { return gs >= 8 ? 3 : (gs == 4 ? 2 : (gs == 2 ? 1 : 0)); }

// Extra LDS passes of one read instruction: the largest number of distinct
// entries falling into one bank pair, minus one, over the lanes of a pass.
long passes_minus_one(const std::vector<int> &vals)
{
    std::array<std::vector<int>, kBankPairs> seen;
    int worst = 0;
    for (int v : vals) {
        auto &s = seen[((v % kBankPairs) + kBankPairs) % kBankPairs];
        if (std::find(s.begin(), s.end(), v) == s.end()) s.push_back(v);
        worst = std::max(worst, (int)s.size());
    }
    return worst > 0 ? worst - 1 : 0;
}

// Slot order of one half-wave of units (kLanesPerPass consecutive units):
// greedy placement that keeps reads of one slot by the half-wave's lanes in
// distinct bank pairs where it can.  Groups of S slots (S = 1: single
// slots, each reading y and z; S = 2: pairs (2p, 2p + 1) sharing ny, of
// which only slot 2p reads y, in either orientation; S = 4: quads sharing
// ny, read by their first slot) are placed whole.  (Modes 1 to 3.)
void order_half_wave(std::vector<std::vector<Vec>> &units, size_t h0, int R, int S, const Entries &E)
{
    std::vector<std::array<std::vector<int>, kBankPairs>> by(R), bz(R);
    auto add = [](std::array<std::vector<int>, kBankPairs> &b, int v) {
        auto &s = b[((v % kBankPairs) + kBankPairs) % kBankPairs];
        if (std::find(s.begin(), s.end(), v) == s.end()) s.push_back(v);
    };
    auto cost1 = [](const std::array<std::vector<int>, kBankPairs> &b, int v) {
        const auto &s = b[((v % kBankPairs) + kBankPairs) % kBankPairs];
        return std::find(s.begin(), s.end(), v) == s.end() ? (int)s.size() : 0;
    };
    const int G = R / S, NO = S == 2 ? 2 : 1;  // orientations tried
    for (size_t u = h0; u < h0 + kLanesPerPass && u < units.size(); u++) {
        std::vector<Vec> &un = units[u];
        std::vector<Vec> placed(R);
        std::vector<char> used_v(G, 0), used_p(G, 0);
        for (int k = 0; k < G; k++) {
            int bv = -1, bp = -1, bo = 0, bc = 1 << 30;
            for (int i = 0; i < G; i++) {
                if (used_v[i]) continue;
                for (int p = 0; p < G; p++) {
                    if (used_p[p]) continue;
                    for (int o = 0; o < NO; o++) {
                        const Vec &a = un[S * i + o];
                        int c = cost1(by[S * p], E.shared(a)) + cost1(bz[S * p], E.own(a));
                        if (S == 2) c += cost1(bz[S * p + 1], E.own(un[S * i + 1 - o]));
                        for (int q = S == 2 ? 2 : 1; q < S; q++) c += cost1(bz[S * p + q], E.own(un[S * i + q]));
                        if (c < bc) { bc = c; bv = i; bp = p; bo = o; }
                    }
                }
            }
            used_v[bv] = used_p[bp] = 1;
            for (int q = 0; q < S; q++) {
                const Vec &v = un[S * bv + (q >= 2 ? q : (q == 0 ? bo : (S == 1 ? 0 : 1 - bo)))];
                placed[S * bp + q] = v;
                if (q == 0) add(by[S * bp], E.shared(v));
                add(bz[S * bp + q], E.own(v));
            }
        }
        un.swap(placed);
    }
}

// ---- grouped layout (mode 4) --------------------------------------------

// Greedy cover of one nx class by sub-classes that share nz (the shared
// read is then z, the own read y) or ny (shared y, own z): repeatedly the
// largest sub-class left.
void cover_groups(const std::vector<Vec> &vs, std::vector<std::vector<Vec>> &out)
{
    const int n = (int)vs.size();
    std::vector<std::vector<int>> byz(256), byy(256);
    std::vector<int> cz(256, 0), cy(256, 0);
    for (int i = 0; i < n; i++) {
        byz[vs[i].nz].push_back(i);
        byy[vs[i].ny + 128].push_back(i);
        cz[vs[i].nz]++;
        cy[vs[i].ny + 128]++;
    }
    std::vector<char> alive(n, 1);
    for (int left = n; left > 0;) {
        const int kz = (int)(std::max_element(cz.begin(), cz.end()) - cz.begin());
        const int ky = (int)(std::max_element(cy.begin(), cy.end()) - cy.begin());
        const bool use_z = cz[kz] >= cy[ky];
        const std::vector<int> &mem = use_z ? byz[kz] : byy[ky];
        std::vector<Vec> grp;
        for (int i : mem) {
            if (!alive[i]) continue;
            alive[i] = 0;
            Vec v = vs[i];
            v.sw = use_z;
            grp.push_back(v);
            cz[vs[i].nz]--;
            cy[vs[i].ny + 128]--;
            left--;
        }
        out.push_back(std::move(grp));
    }
}

// Sizes of the pieces a class of m vectors is cut into (powers of two up
// to gmax; the last piece may carry padding slots), cheapest by
// cost(piece of s slots) = s * c_mac + c_q: each slot's accumulate plus one
// product q per piece.
std::vector<int> piece_sizes(int m, int gmax, double c_mac, double c_q)
{
    std::vector<int> out;
    while (m > 2 * gmax) {
        out.push_back(gmax);
        m -= gmax;
    }
    std::vector<double> best(m + 1, 1e300);
    std::vector<int> pick(m + 1, 1);
    best[0] = 0.0;
    for (int k = 1; k <= m; k++)
        for (int s = 1; s <= gmax; s *= 2) {
            const double c = s * c_mac + c_q + best[k - std::min(s, k)];
            if (c < best[k]) { best[k] = c; pick[k] = s; }
        }
    for (int k = m; k > 0; k -= std::min(pick[k], k)) out.push_back(pick[k]);
    return out;
}

// Bank-pair clashes of the reads of half-waves, for the local search below.
// Read q of a half-wave is the 32 lanes' entry indices of one instruction;
// its cost is the sum over bank pairs of (distinct entries in the pair)^2
// (32 when every lane hits its own pair or a broadcast); wildcards (-1:
// padding slots, set afterwards to an entry another lane reads) cost
// nothing.
struct ReadCounters {
    int n_reads = 0, W = 0;
    std::vector<int> cnt;  // [n_reads][W]: lanes reading each entry
    std::vector<int> d;    // [n_reads][32]: distinct entries per bank pair
    void init(int nr, int w)
    {
        n_reads = nr;
        W = w;
        cnt.assign((size_t)nr * w, 0);
        d.assign((size_t)nr * kBankPairs, 0);
    }
    // cost change of adding (+1) or removing (-1) one lane's entry a
    long change(int q, int a, int sgn)
    {
        if (a < 0) return 0;
        int &c = cnt[(size_t)q * W + a];
        int &b = d[(size_t)q * kBankPairs + a % kBankPairs];
        long delta = 0;
        if (sgn > 0) {
            if (c++ == 0) { delta = 2L * b + 1; b++; }
        } else {
            if (--c == 0) { delta = -(2L * b - 1); b--; }
        }
        return delta;
    }
    // whether entry a of read q shares its bank pair with another entry
    bool clashes(int q, int a) const { return a >= 0 && d[(size_t)q * kBankPairs + a % kBankPairs] > 1; }
    int worst(int q) const
    {
        int m = 0;
        for (int b = 0; b < kBankPairs; b++) m = std::max(m, d[(size_t)q * kBankPairs + b]);
        return std::max(m, 1);
    }
};

// Reads of a unit of R slots in groups of GS: own read of slot r is read r,
// the shared read of group g (by its slot g * GS) is read R + g.
struct UnitReads {
    int R, GS;
    const Entries &E;
    int own(const Vec &v) const { return v.d < 0 ? -1 : E.own(v); }
    int shared(const std::vector<Vec> &un, int g) const
    {
        for (int i = 0; i < GS; i++)
            if (un[g * GS + i].d >= 0) return E.shared(un[g * GS + i]);
        return -1;  // padding group
    }
    bool clashes(const ReadCounters &rc, const std::vector<Vec> &un) const
    {
        for (int r = 0; r < R; r++)
            if (rc.clashes(r, own(un[r]))) return true;
        for (int g = 0; g < R / GS; g++)
            if (rc.clashes(R + g, shared(un, g))) return true;
        return false;
    }
    long apply(ReadCounters &rc, const std::vector<Vec> &un, int sgn) const
    {
        long delta = 0;
        for (int r = 0; r < R; r++) delta += rc.change(r, own(un[r]), sgn);
        for (int g = 0; g < R / GS; g++) delta += rc.change(R + g, shared(un, g), sgn);
        return delta;
    }
};

// Local search over the units of one chunk that share a group size: for a
// unit with a clashing read, swap two slots of one group, two groups of the
// unit, or the unit with one of another half-wave, keeping each move that
// does not raise the summed clash cost.  hw[h] lists the unit indices of
// half-wave h.
void optimise_reads(std::vector<std::vector<Vec>> &units, const std::vector<std::vector<int>> &hw, int R, int GS,
                    const Entries &E, int W, uint64_t seed, int NS)
{
    // NS: slots that hold vectors (the rest are dead and must stay past
    // them); moves only permute the first NS / GS groups
    const int H = (int)hw.size(), G = R / GS, GA = NS / GS;
    if (H == 0) return;
    const UnitReads ur{R, GS, E};
    std::vector<ReadCounters> rc(H);
    for (int h = 0; h < H; h++) {
        rc[h].init(R + G, W);
        for (int u : hw[h]) ur.apply(rc[h], units[u], +1);
    }
    std::vector<std::pair<int, int>> pos;  // (half-wave, lane)
    for (int h = 0; h < H; h++)
        for (int l = 0; l < (int)hw[h].size(); l++) pos.push_back({h, l});
    // 1,500 tries per unit: the 28,181-vector shell set of the 64,000-atom
    // workload falls from 2,071 extra LDS passes per atom unplaced to 782
    // at 150 tries, 310 at 1,200 and 235 at 2,400; with 782 its rho kernel
    // ran 404 ms, with 235 378 ms.  The search stops early once draws stop
    // finding clashing lanes.
    const long moves = 1500L * (long)pos.size(), draws = 40L * moves;
    uint64_t st = seed;
    long tried = 0;
    // annealing: a move that raises the cost by delta is kept with
    // probability exp(-delta / temp), temp falling linearly to 0
    auto keep = [&](long delta) {
        if (delta <= 0) return true;
        const double temp = 2.0 * (1.0 - (double)tried / (double)moves);
        if (temp <= 0.0) return false;
        const double u = (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
        return u < std::exp(-(double)delta / temp);
    };
    for (long it = 0, miss = 0; it < draws && tried < moves && miss < 64L * (long)pos.size(); it++) {
        const uint64_t rnd = splitmix64(st);
        const int kind = (int)(rnd % 3);
        const auto &p1 = pos[(rnd >> 8) % pos.size()];
        std::vector<Vec> &a = units[hw[p1.first][p1.second]];
        ReadCounters &ra = rc[p1.first];
        if (!ur.clashes(ra, a)) {
            miss++;
            continue;
        }
        miss = 0;
        tried++;
        if (kind == 0 && GS > 1) {
            // two slots of one group: the group's shared read is unchanged
            const int g = (int)((rnd >> 32) % GA), i = (int)((rnd >> 40) % GS), j = (int)((rnd >> 48) % GS);
            if (i == j) continue;
            const int ri = g * GS + i, rj = g * GS + j;
            long d = ra.change(ri, ur.own(a[ri]), -1) + ra.change(rj, ur.own(a[rj]), -1) +
                     ra.change(ri, ur.own(a[rj]), +1) + ra.change(rj, ur.own(a[ri]), +1);
            if (keep(d)) {
                std::swap(a[ri], a[rj]);
            } else {
                ra.change(ri, ur.own(a[rj]), -1);
                ra.change(rj, ur.own(a[ri]), -1);
                ra.change(ri, ur.own(a[ri]), +1);
                ra.change(rj, ur.own(a[rj]), +1);
            }
        } else if (kind == 1 && GA > 1) {
            // two groups of one unit
            const int g1 = (int)((rnd >> 32) % GA), g2 = (int)((rnd >> 40) % GA);
            if (g1 == g2) continue;
            long d = ur.apply(ra, a, -1);
            std::vector<Vec> b = a;
            for (int i = 0; i < GS; i++) std::swap(b[g1 * GS + i], b[g2 * GS + i]);
            d += ur.apply(ra, b, +1);
            if (keep(d)) {
                a.swap(b);
            } else {
                ur.apply(ra, b, -1);
                ur.apply(ra, a, +1);
            }
        } else if (H > 1) {
            // two units of different half-waves
            const auto &p2 = pos[(rnd >> 32) % pos.size()];
            if (p2.first == p1.first) continue;
            std::vector<Vec> &b = units[hw[p2.first][p2.second]];
            ReadCounters &rb = rc[p2.first];
            long d = ur.apply(ra, a, -1) + ur.apply(rb, b, -1) + ur.apply(ra, b, +1) + ur.apply(rb, a, +1);
            if (keep(d)) {
                a.swap(b);
            } else {
                ur.apply(ra, b, -1);
                ur.apply(rb, a, -1);
                ur.apply(ra, a, +1);
                ur.apply(rb, b, +1);
            }
        }
    }
}

// Padding slots read what another lane of their half-wave reads in the
// same instruction (a broadcast): own reads copy the first real lane's own
// entry, a padding group's lead also its shared entry where the two lie in
// different sections.  A padding unit takes that lane's nx as well.
void fill_padding(std::vector<std::vector<Vec>> &units, size_t h0, int R, int GS, const Entries &E)
{
    const size_t h1 = std::min(units.size(), h0 + kLanesPerPass);
    for (int r = 0; r < R; r++) {
        int ref = -1;
        for (size_t u = h0; u < h1 && ref < 0; u++)
            if (units[u][r].d >= 0) ref = (int)u;
        if (ref < 0) continue;
        const Vec &src = units[ref][r];
        for (size_t u = h0; u < h1; u++) {
            Vec &v = units[u][r];
            if (v.d >= 0) continue;
            const bool lead = r % GS == 0;
            bool lead_pad = lead;
            for (int i = r - r % GS; i < r - r % GS + GS && lead_pad; i++)
                if (units[u][i].d >= 0) lead_pad = false;
            if (!lead || lead_pad) {
                // a slot that reads no shared entry, or a padding group's
                // lead: read both of src's entries
                v.ny = src.ny;
                v.nz = src.nz;
                v.sw = src.sw;
            } else if (v.sw == src.sw) {
                // a padding lead of a real group keeps the group's shared
                // entry; its own entry follows src's where on the same axis
                if (v.sw) v.ny = src.ny;
                else v.nz = src.nz;
            }
        }
    }
    // padding units: every slot padding; take the reference lane's nx
    for (size_t u = h0; u < h1; u++) {
        bool all_pad = true;
        for (int r = 0; r < R && all_pad; r++)
            if (units[u][r].d >= 0) all_pad = false;
        if (!all_pad) continue;
        for (size_t w = h0; w < h1; w++) {
            bool real = false;
            for (int r = 0; r < R && !real; r++) real = units[w][r].d >= 0;
            if (real) {
                for (int r = 0; r < R; r++) units[u][r].nx = units[w][0].nx;
                break;
            }
        }
    }
}

// Waves w and w + 4 of a 512-thread block run on one SIMD (measured on
// gfx950 from HW_ID: in every block the two share a SIMD, while which SIMD
// wave 0 lands on varies).  Waves of different group sizes differ in work
// per atom, R ((c - 4) + 4 / GS) VALU ops per lane, and the block meets a
// barrier every tile, so its tile time is that of its busiest SIMD.  Per
// chunk, the active waves (a prefix: the kernels skip waves past the
// chunk's units) are permuted so that the largest SIMD load is smallest;
// ties keep the first order found, so the layout stays deterministic.
template <typename SlotsOf>
void balance_simds(std::vector<std::vector<Vec>> &units, std::vector<char> &unit_g, int n_units, int R, double c,
                   const SlotsOf &slots_of)
{
    (void)R;
    const int n_chunks = (int)(units.size() / 512);
    for (int ch = 0; ch < n_chunks; ch++) {
        const int n_act = std::min(512, n_units - ch * 512);
        const int a = (n_act + 63) / 64;
        if (a <= 4) continue;  // one wave per SIMD at most
        double load[8];
        for (int w = 0; w < a; w++) {
            const int gs = unit_g[(size_t)ch * 512 + w * 64];
            load[w] = slots_of(gs) * ((c - 4.0) + 4.0 / gs);
        }
        std::array<int, 8> perm{}, best{};
        for (int w = 0; w < 8; w++) perm[w] = best[w] = w;
        auto worst = [&](const std::array<int, 8> &p) {
            double m = 0.0;
            for (int s = 0; s < 4; s++) m = std::max(m, load[p[s]] + (s + 4 < a ? load[p[s + 4]] : 0.0));
            return m;
        };
        double best_cost = worst(perm);
        while (std::next_permutation(perm.begin(), perm.begin() + a)) {
            const double cst = worst(perm);
            if (cst < best_cost - 1e-9) { best_cost = cst; best = perm; }
        }
        std::vector<std::vector<Vec>> moved((size_t)a * 64);
        std::vector<char> moved_g((size_t)a * 64);
        for (int w = 0; w < a; w++)
            for (int l = 0; l < 64; l++) {
                const size_t src = (size_t)ch * 512 + (size_t)best[w] * 64 + l;
                moved[(size_t)w * 64 + l] = std::move(units[src]);
                moved_g[(size_t)w * 64 + l] = unit_g[src];
            }
        for (size_t i = 0; i < moved.size(); i++) {
            units[(size_t)ch * 512 + i] = std::move(moved[i]);
            unit_g[(size_t)ch * 512 + i] = moved_g[i];
        }
    }
}

}  // namespace

PlanLayout plan_layout(const int *nvec, int K, int R, int mode, int MM, double c, bool place)
{
    const Entries E{table_oy(MM), table_oz(MM)};
    const bool pair = mode != 0, mixed = mode == 2, quad = mode == 3, grouped = mode == 4 || mode == 5;
    // mode 5: waves of single slots hold R/2 vectors per thread and waves of
    // pairs 3R/4 (R = 8), so that every wave carries about the same work per
    // atom: slots past a unit's count are dead (padding the kernels skip)
    const bool short_waves = mode == 5 && R == 8;
    auto slots_of = [&](int gs) { return short_waves ? (gs == 1 ? R / 2 : (gs == 2 ? 3 * R / 4 : R)) : R; };
    const int W = std::max(5 * table_h(MM) + 3, 257);
    PlanLayout L;
    L.R = R;
    L.KB = 512 * R;
    L.pair = pair;
    L.quad = quad;
    std::vector<Vec> vs(K);
    for (int v = 0; v < K; v++) {
        int nx = nvec[3 * v], ny = nvec[3 * v + 1], nz = nvec[3 * v + 2], d = v;
        if (nz < 0) { nx = -nx; ny = -ny; nz = -nz; d |= (1 << 30); }
        vs[v] = {nx, ny, nz, d};
    }
    std::stable_sort(vs.begin(), vs.end(), [](const Vec &a, const Vec &b) {
        if (a.nx != b.nx) return a.nx < b.nx;
        if (a.ny != b.ny) return a.ny < b.ny;
        return a.nz < b.nz;
    });
    // Thread units.  Unpaired (mode 0): per nx group, R vectors per unit,
    // dealt with stride over ceil(count / R) units.  Paired: per nx group,
    // pairs of two vectors with one ny, then (of the vectors left) two with
    // one nz (swapped roles), dealt R / 2 pairs per unit; each vector left
    // over either forms a pair with a padding slot (modes 1, 3) or goes to
    // unpaired units (mode 2).  Mode 3 first takes quads, four vectors with
    // one ny, dealt R / 4 per unit.  Grouped (mode 4): per nx group, the
    // classes of cover_groups cut into pieces of 8, 4, 2 and 1 slots
    // (piece_sizes), R / GS pieces of one size per unit.  In modes 2 to 4
    // every wave holds one kind of unit, emitted in nx order as 64 units of
    // the kind fill, so a chunk's waves of every kind cover a narrow nx
    // range.
    const int kinds[4] = {8, 4, 2, 1};
    std::vector<std::vector<Vec>> pend[4];  // per group size 8, 4, 2, 1
    size_t sent[4] = {0, 0, 0, 0};
    std::vector<std::vector<Vec>> units_q, units_p, units_s;
    std::vector<std::vector<Vec>> waved;
    std::vector<char> wave_g;
    size_t sent_q = 0, sent_p = 0, sent_s = 0;
    auto emit = [&](std::vector<std::vector<Vec>> &src, size_t &sent_n, char g, bool flush) {
        while (src.size() - sent_n >= 64 || (flush && src.size() > sent_n)) {
            const size_t take = std::min<size_t>(64, src.size() - sent_n);
            for (size_t i = 0; i < take; i++) waved.push_back(src[sent_n + i]);
            for (size_t i = take; i < 64; i++)
                waved.push_back(std::vector<Vec>(R, Vec{src[sent_n + take - 1][0].nx, 0, 0, -1}));
            wave_g.push_back(g);
            sent_n += take;
        }
    };
    // deal groups of S vectors, per_unit groups per unit, with stride
    auto deal = [&](const std::vector<std::array<Vec, 4>> &grp, int per_unit, int S, int nx,
                    std::vector<std::vector<Vec>> &dst) {
        const int cnt = (int)grp.size(), U = (cnt + per_unit - 1) / per_unit;
        for (int u = 0; u < U; u++) {
            std::vector<Vec> un;
            for (int q = u; q < cnt; q += U)
                for (int k = 0; k < S; k++) un.push_back(grp[q][k]);
            while ((int)un.size() < R) un.push_back({nx, 0, 0, -1});
            dst.push_back(un);
        }
    };
    const Vec none{0, 0, 0, -1};
    const int gmax = std::min(R, 8);
    for (int g0 = 0; g0 < K;) {
        int g1 = g0;
        while (g1 < K && vs[g1].nx == vs[g0].nx) g1++;
        const int nx = vs[g0].nx;
        if (grouped) {
            std::vector<std::vector<Vec>> classes;
            cover_groups(std::vector<Vec>(vs.begin() + g0, vs.begin() + g1), classes);
            std::vector<std::vector<Vec>> pieces[4];
            for (const auto &cl : classes) {
                size_t at = 0;
                for (int s : piece_sizes((int)cl.size(), gmax, c - 4.0, 4.0)) {
                    std::vector<Vec> pc;
                    for (int i = 0; i < s; i++) {
                        if (at < cl.size()) pc.push_back(cl[at++]);
                        else pc.push_back(Vec{nx, cl.back().ny, cl.back().nz, -1, cl.back().sw});
                    }
                    pieces[s == 8 ? 0 : (s == 4 ? 1 : (s == 2 ? 2 : 3))].push_back(std::move(pc));
                }
            }
            for (int k = 0; k < 4; k++) {
                const int s = kinds[k], per = slots_of(s) / s;
                if (s > gmax) continue;
                for (size_t i = 0; i < pieces[k].size(); i += per) {
                    std::vector<Vec> un;
                    for (int j = 0; j < per; j++) {
                        if (i + j < pieces[k].size()) un.insert(un.end(), pieces[k][i + j].begin(), pieces[k][i + j].end());
                        else for (int q = 0; q < s; q++) un.push_back(Vec{nx, 0, 0, -1});
                    }
                    while ((int)un.size() < R) un.push_back(Vec{nx, 0, 0, -1});  // dead slots
                    pend[k].push_back(std::move(un));
                }
                emit(pend[k], sent[k], (char)s, false);
            }
            g0 = g1;
            continue;
        }
        std::vector<std::array<Vec, 4>> quads, grp, single;
        std::vector<Vec> left;
        for (int i = g0; i < g1;) {
            int j = i;
            while (j < g1 && vs[j].ny == vs[i].ny) j++;
            // run [i, j) of one ny: quads (mode 3), then pairs, then one left over
            int k = i;
            if (quad)
                for (; k + 4 <= j; k += 4) quads.push_back({vs[k], vs[k + 1], vs[k + 2], vs[k + 3]});
            if (pair)
                for (; k + 2 <= j; k += 2) grp.push_back({vs[k], vs[k + 1], none, none});
            for (; k < j; k++) left.push_back(vs[k]);
            i = j;
        }
        if (pair) std::stable_sort(left.begin(), left.end(), [](const Vec &a, const Vec &b) { return a.nz < b.nz; });
        for (size_t i = 0; i < left.size();) {
            if (pair && i + 1 < left.size() && left[i + 1].nz == left[i].nz) {
                Vec a = left[i], b = left[i + 1];
                a.sw = b.sw = true;
                grp.push_back({a, b, none, none});
                i += 2;
            } else if (pair && !mixed) {
                grp.push_back({left[i], Vec{nx, left[i].ny, 0, -1}, none, none});
                i += 1;
            } else {
                single.push_back({left[i], none, none, none});
                i += 1;
            }
        }
        if (quad) deal(quads, R / 4, 4, nx, units_q);
        if (pair) deal(grp, R / 2, 2, nx, units_p);
        deal(single, R, 1, nx, units_s);
        if (mixed || quad) {
            emit(units_q, sent_q, 4, false);
            emit(units_p, sent_p, 2, false);
            emit(units_s, sent_s, 1, false);
        }
        g0 = g1;
    }
    // unit_g[u]: group size of unit u's wave
    std::vector<std::vector<Vec>> units;
    std::vector<char> unit_g;
    if (grouped) {
        for (int k = 0; k < 4; k++) emit(pend[k], sent[k], (char)kinds[k], true);
    }
    if (mixed || quad) {
        emit(units_q, sent_q, 4, true);
        emit(units_p, sent_p, 2, true);
        emit(units_s, sent_s, 1, true);
    }
    if (grouped || mixed || quad) {
        units = std::move(waved);
        for (char g : wave_g) unit_g.insert(unit_g.end(), 64, g);
    } else {
        units = std::move(units_p);
        unit_g.assign(units.size(), pair ? 2 : 1);
        units.insert(units.end(), units_s.begin(), units_s.end());
        unit_g.resize(units.size(), 1);
    }
    const int n_units = (int)units.size();
    L.n_chunks = std::max(1, (n_units + 511) / 512);
    while ((int)units.size() < L.n_chunks * 512) {
        const int nx = units.empty() ? 0 : units.back()[0].nx;
        units.push_back(std::vector<Vec>(R, Vec{nx, 0, 0, -1}));
        unit_g.push_back(1);
    }
    if (grouped && place) {
        // per chunk and group size: local search over the half-waves of the
        // chunk's real units, then padding slots pointed at broadcasts.
        // The searches touch disjoint units and carry their own seeds, so
        // they run on a thread pool and give the same layout on any number
        // of threads.
        struct Task {
            int c, gs;
            std::vector<std::vector<int>> hw;
        };
        std::vector<Task> tasks;
        for (int c = 0; c < L.n_chunks; c++) {
            const int n_act = std::min(512, n_units - c * 512);
            for (int gs = 1; gs <= 8; gs *= 2) {
                Task t{c, gs, {}};
                for (int h = 0; h < n_act; h += kLanesPerPass) {
                    const size_t u0 = (size_t)c * 512 + h;
                    if (unit_g[u0] != gs) continue;
                    std::vector<int> lanes;
                    for (int l = 0; l < kLanesPerPass && h + l < n_act; l++) lanes.push_back((int)u0 + l);
                    t.hw.push_back(std::move(lanes));
                }
                if (!t.hw.empty()) tasks.push_back(std::move(t));
            }
        }
        // largest searches first
        std::stable_sort(tasks.begin(), tasks.end(), [](const Task &a, const Task &b) { return a.hw.size() > b.hw.size(); });
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t i; (i = next.fetch_add(1)) < tasks.size();)
                optimise_reads(units, tasks[i].hw, R, tasks[i].gs, E, W,
                               0x9E3779B97F4A7C15ull * (uint64_t)(tasks[i].c * 16 + tasks[i].gs + 1),
                               slots_of(tasks[i].gs));
        };
        const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
        const size_t n_threads = std::min<size_t>(std::min<size_t>(hc, 32), tasks.size());
        std::vector<std::thread> pool;
        for (size_t t = 1; t < n_threads; t++) {
            // a thread that cannot be started leaves its share to the others
            // (the tasks are drawn from one counter)
            try {
                pool.emplace_back(work);
            } catch (...) {
                break;
            }
        }
        work();
        for (auto &th : pool) th.join();
        for (size_t h0 = 0; h0 < units.size(); h0 += kLanesPerPass) fill_padding(units, h0, R, unit_g[h0], E);
    } else if (!grouped) {
        for (size_t h0 = 0; h0 < units.size(); h0 += kLanesPerPass) order_half_wave(units, h0, R, unit_g[h0], E);
    }
    if (grouped || mixed || quad) balance_simds(units, unit_g, n_units, R, c, slots_of);
    // emit
    const size_t n = (size_t)L.n_chunks * L.KB;
    L.pack.resize(n);
    L.dst.resize(n);
    L.range.assign((size_t)L.n_chunks * 4, 0);
    for (int ch = 0; ch < L.n_chunks; ch++) {
        int xlo = 1 << 30, xhi = -(1 << 30), my = 0, mz = 0;
        for (int t = 0; t < 512; t++) {
            const std::vector<Vec> &un = units[(size_t)ch * 512 + t];
            for (int r = 0; r < R; r++) {
                const Vec &v = un[r];
                const size_t pos = (size_t)ch * L.KB + (size_t)r * 512 + t;
                L.pack[pos] = (uint32_t)(v.nx + 128) | ((uint32_t)(v.ny + 128) << 8) | ((uint32_t)v.nz << 16) |
                              (v.sw ? kPackSwap : 0u);
                L.dst[pos] = v.d;
                xlo = std::min(xlo, v.nx);
                xhi = std::max(xhi, v.nx);
                my = std::max(my, std::abs(v.ny));
                mz = std::max(mz, v.nz);
            }
        }
        L.range[4 * ch] = xlo;
        L.range[4 * ch + 1] = xhi;
        // bit 15: the plan's waves of singles and pairs hold fewer slots
        L.range[4 * ch + 2] = my | (mz << 16) | (short_waves ? (1 << 15) : 0);
        const int n_act = std::min(512, n_units - ch * 512);
        // two bits per wave: its group code (group size 1 << code)
        unsigned codes = 0;
        int a_by_g[4] = {0, 0, 0, 0};  // active waves by code
        for (int w = 0; w < 8 && w * 64 < n_act; w++) {
            const int g = unit_g[(size_t)ch * 512 + w * 64];
            const int cd = pair ? gs_code(g) : 0;
            codes |= (unsigned)cd << (2 * w);
            a_by_g[cd]++;
        }
        L.range[4 * ch + 3] = (int)((unsigned)n_act | (codes << 16));
        L.active_waves += (n_act + 63) / 64;
        // The chunk's time per atom follows its busiest SIMD: waves w and
        // w + 4 share one, and each wave's lanes issue its slots x
        // (c - 4 + 4/GS) VALU operations per atom.
        {
            double load[8] = {0, 0, 0, 0, 0, 0, 0, 0}, busiest = 0.0;
            for (int w = 0; w < 8 && w * 64 < n_act; w++) {
                const int gs = unit_g[(size_t)ch * 512 + w * 64];
                load[w] = slots_of(gs) * ((c - 4.0) + 4.0 / gs);
            }
            for (int q = 0; q < 4; q++) busiest = std::max(busiest, load[q] + load[q + 4]);
            // the per-tile table build, spread over every SIMD: about 95
            // VALU per (axis, group of 8 entries) item and 1,400 for the
            // anchors, per atom
            const int xa = xlo >= 0 ? xlo : (xhi <= 0 ? -xhi : 0), xb = std::max(std::abs(xlo), std::abs(xhi));
            const int items = (xb >> 3) - (xa >> 3) + 1 + (my >> 3) + 1 + (mz >> 3) + 1;
            L.simd_cost += 4.0 * 64.0 * busiest + 95.0 * items + 1400.0;
        }
        // Work of the waves that run the sums, in slots.  A block's 8 waves
        // share 4 SIMDs two by two, so a chunk with a active waves takes as
        // long as ceil(a / 4) * 4 of them.
        const int a = (n_act + 63) / 64;
        if (a > 0) {
            const double scale = (double)(((a + 3) / 4) * 4) / a;
            L.gs8_slots += (long)(a_by_g[3] * 64.0 * slots_of(8) * scale);
            L.quad_slots += (long)(a_by_g[2] * 64.0 * slots_of(4) * scale);
            L.paired_slots += (long)(a_by_g[1] * 64.0 * slots_of(2) * scale);
            L.unpaired_slots += (long)(a_by_g[0] * 64.0 * slots_of(1) * scale);
        }
        // conflict estimate of the chunk's reads
        for (int r = 0; r < R; r++)
            for (int h = 0; h < 512; h += kLanesPerPass) {
                std::vector<int> ys, zs, xs;
                for (int t = h; t < h + kLanesPerPass; t++) {
                    const Vec &v = units[(size_t)ch * 512 + t][r];
                    ys.push_back(E.shared(v));
                    zs.push_back(E.own(v));
                    if (r == 0) xs.push_back(v.nx);
                }
                if (r % unit_g[(size_t)ch * 512 + h] == 0) L.conflicts += passes_minus_one(ys);
                L.conflicts += passes_minus_one(zs);
                if (r == 0) L.conflicts += passes_minus_one(xs);
            }
    }
    return L;
}

}  // namespace amdat

static int amdat_internal_plan_layout_body(const int *nvec, int K, int R, int pair, uint32_t *pack, int *dst,
                                          int *range, int *n_chunks, long *conflicts)
{
    if (!nvec || K < 1 || (R != 2 && R != 4 && R != 8) || !n_chunks || pair < 0 || pair > 5) return AMDAT_EINVAL;
    if (pair && R < 4) return AMDAT_EINVAL;
    int M = 0;
    for (int i = 0; i < 3 * K; i++) M = std::max(M, std::abs(nvec[i]));
    if (table_half_width(M) < 0) return AMDAT_ERANGE;
    amdat::PlanLayout L = amdat::plan_layout(nvec, K, R, pair, table_half_width(M));
    *n_chunks = L.n_chunks;
    if (conflicts) *conflicts = L.conflicts;
    if (pack) {
        std::copy(L.pack.begin(), L.pack.end(), pack);
        std::copy(L.dst.begin(), L.dst.end(), dst);
        std::copy(L.range.begin(), L.range.end(), range);
    }
    return AMDAT_OK;
}

extern "C" int amdat_internal_plan_layout(const int *nvec, int K, int R, int pair, uint32_t *pack, int *dst,
                                          int *range, int *n_chunks, long *conflicts)
{
    AMDAT_GUARD("amdat_internal_plan_layout", amdat_internal_plan_layout_body(nvec, K, R, pair, pack, dst, range, n_chunks, conflicts))
}
