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
#include <string>
#include <vector>

#include "../../include/amdat_hip.h"
#include "abi_internal.h"

using namespace amdat;

extern "C" {

int amdat_version(void) { return 1 * 10000 + 0 * 100 + 0; }

int amdat_ts_n_timesteps(int nexp, int nes)
{
    if (nexp < 1 || nes < 0) return set_error(AMDAT_EINVAL, "time scheme: n_exponentials must be >= 1 and n_exponential_steps >= 0");
    return nes * nexp + 1;
}

int amdat_ts_n_timegaps(int nexp, int nes, int frt)
{
    if (nexp < 1 || nes < 0) return set_error(AMDAT_EINVAL, "time scheme: n_exponentials must be >= 1 and n_exponential_steps >= 0");
    return nes + nexp - 1 + (frt ? 0 : 1);
}

// System::create_timelist, src/system.cpp:3542-3568.  exp_base is a float
// member; pow(float, int) is evaluated in double.
int amdat_ts_timelist(int nexp, int nes, float base, int first_exp, float dt, float *out, int cap)
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

// System::displacement_times, src/system.cpp:3763-3784.
int amdat_ts_displacement_times(int nexp, int nes, int frt, const float *tl, float *out, int cap)
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

// System::displacement_list, src/system.cpp:4090-4179.  The pair order is
// the reference's: in-block gaps by gap then block, cross-block gaps by gap,
// block offset, block; then the (0, last) pair when frt == 0.
int amdat_ts_pairs(int nexp, int nes, int frt, int fullblock, int limit, int *tg, int *t0, int *t1,
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

// System::displacement_list(analysis, timegap, false), src/system.cpp:
// 4349-4398: the pairs of one time gap with the first time of each block
// only (in-block gaps: every block; cross-block gaps: blocks that reach;
// the last gap: frames 0 and n_timesteps - 1, which with frt = 1 is time
// gap n_timegaps, one past the full enumeration).  limit > 0 keeps the
// first limit pairs.
int amdat_ts_gap_pairs(int nexp, int nes, int frt, int limit, int time_index, int *tg, int *t0, int *t1, int cap,
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

// System::timegap_weighting, src/system.cpp:4599-4628 (cross-block entries
// carry the reference's +1).
int amdat_ts_weighting(int nexp, int nes, int frt, int fullblock, int limit, int *w, int cap)
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

int amdat_shells_read(const char *stem, int first_bin, int n_bins, int *shell_off, int *nvec,
                      int cap_vectors, int *n_vectors)
{
    return read_shell_files(stem, "xyz", first_bin, n_bins, shell_off, nvec, cap_vectors, n_vectors, nullptr);
}

int amdat_shells_read_plane(const char *stem, const char *plane, int first_bin, int n_bins, int *shell_off,
                            int *nvec, int cap_vectors, int *n_vectors, int *denominator)
{
    return read_shell_files(stem, plane, first_bin, n_bins, shell_off, nvec, cap_vectors, n_vectors, denominator);
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
int amdat_shells_generate(int first_bin, int n_bins, int max_per_shell, uint64_t seed,
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

}  // extern "C"

// ---- block layout of k-vector plans -------------------------------------

namespace amdat {

namespace {

constexpr int kLanesPerPass = 32;  // a wave's 8-byte LDS reads are served 32 lanes at a time
constexpr int kBankPairs = 32;     // 64 banks of 4 bytes = 32 eight-byte entries

struct Vec {
    int nx, ny, nz, d;
    bool sw = false;  // paired on nz: the shared read is z, the own read y
};

// Table entries a slot reads besides x: the one its pair shares (y, or z
// when swapped) and its own (z, or y).  Their bank pairs follow the entry
// index: every lane of a read addresses the same atom row.
struct Entries {
    int oy, oz;
    int shared(const Vec &v) const { return v.sw ? oz + v.nz : oy + v.ny; }
    int own(const Vec &v) const { return v.sw ? oy + v.ny : oz + v.nz; }
};

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
// ny, read by their first slot) are placed whole.
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

}  // namespace

PlanLayout plan_layout(const int *nvec, int K, int R, int mode, int MM)
{
    const Entries E{table_oy(MM), table_oz(MM)};
    const bool pair = mode != 0, mixed = mode == 2, quad = mode == 3;
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
    // one ny, dealt R / 4 per unit.  Every wave holds one kind of unit: in
    // modes 2 and 3 the kinds fill separate waves of 64 units, emitted in nx
    // order as they fill, so a chunk's waves of every kind cover a narrow nx
    // range.
    std::vector<std::vector<Vec>> units_q, units_p, units_s;
    // modes 2, 3: waves emitted as they fill, with their group size
    std::vector<std::vector<Vec>> waved;
    std::vector<char> wave_g;
    size_t sent_q = 0, sent_p = 0, sent_s = 0;
    auto emit = [&](std::vector<std::vector<Vec>> &src, size_t &sent, char g, bool flush) {
        while (src.size() - sent >= 64 || (flush && src.size() > sent)) {
            const size_t take = std::min<size_t>(64, src.size() - sent);
            for (size_t i = 0; i < take; i++) waved.push_back(src[sent + i]);
            for (size_t i = take; i < 64; i++)
                waved.push_back(std::vector<Vec>(R, Vec{src[sent + take - 1][0].nx, 0, 0, -1}));
            wave_g.push_back(g);
            sent += take;
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
    for (int g0 = 0; g0 < K;) {
        int g1 = g0;
        while (g1 < K && vs[g1].nx == vs[g0].nx) g1++;
        const int nx = vs[g0].nx;
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
    // unit_g[u]: group size of unit u's wave (1 unpaired, 2 pairs, 4 quads)
    std::vector<std::vector<Vec>> units;
    std::vector<char> unit_g;
    if (mixed || quad) {
        emit(units_q, sent_q, 4, true);
        emit(units_p, sent_p, 2, true);
        emit(units_s, sent_s, 1, true);
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
    for (size_t h0 = 0; h0 < units.size(); h0 += kLanesPerPass) order_half_wave(units, h0, R, unit_g[h0], E);
    // emit
    const size_t n = (size_t)L.n_chunks * L.KB;
    L.pack.resize(n);
    L.dst.resize(n);
    L.range.assign((size_t)L.n_chunks * 4, 0);
    for (int c = 0; c < L.n_chunks; c++) {
        int xlo = 1 << 30, xhi = -(1 << 30), my = 0, mz = 0;
        for (int t = 0; t < 512; t++) {
            const std::vector<Vec> &un = units[(size_t)c * 512 + t];
            for (int r = 0; r < R; r++) {
                const Vec &v = un[r];
                const size_t pos = (size_t)c * L.KB + (size_t)r * 512 + t;
                L.pack[pos] = (uint32_t)(v.nx + 128) | ((uint32_t)(v.ny + 128) << 8) | ((uint32_t)v.nz << 16) |
                              (v.sw ? kPackSwap : 0u);
                L.dst[pos] = v.d;
                xlo = std::min(xlo, v.nx);
                xhi = std::max(xhi, v.nx);
                my = std::max(my, std::abs(v.ny));
                mz = std::max(mz, v.nz);
            }
        }
        L.range[4 * c] = xlo;
        L.range[4 * c + 1] = xhi;
        L.range[4 * c + 2] = my | (mz << 16);
        const int n_act = std::min(512, n_units - c * 512);
        // bit w of the mask: wave w holds pairs (modes 1, 2) or quads (mode 3)
        int mask = 0, aq = 0, ap = 0;
        for (int w = 0; w < 8 && w * 64 < n_act; w++) {
            const int g = unit_g[(size_t)c * 512 + w * 64];
            if (g == 4) aq++;
            if (g == 2) ap++;
            if (g == (quad ? 4 : 2)) mask |= 1 << w;
        }
        L.range[4 * c + 3] = n_act | (mask << 16);
        // Work of the waves that run the sums, in slots.  A block's 8 waves
        // share 4 SIMDs two by two, so a chunk with a active waves takes as
        // long as ceil(a / 4) * 4 of them.
        const int a = (n_act + 63) / 64;
        if (a > 0) {
            const double scale = (double)(((a + 3) / 4) * 4) / a;
            L.quad_slots += (long)(aq * 64.0 * R * scale);
            L.paired_slots += (long)(ap * 64.0 * R * scale);
            L.unpaired_slots += (long)((a - ap - aq) * 64.0 * R * scale);
        }
        // conflict estimate of the chunk's reads
        for (int r = 0; r < R; r++)
            for (int h = 0; h < 512; h += kLanesPerPass) {
                std::vector<int> ys, zs, xs;
                for (int t = h; t < h + kLanesPerPass; t++) {
                    const Vec &v = units[(size_t)c * 512 + t][r];
                    ys.push_back(E.shared(v));
                    zs.push_back(E.own(v));
                    if (r == 0) xs.push_back(v.nx);
                }
                if (r % unit_g[(size_t)c * 512 + h] == 0) L.conflicts += passes_minus_one(ys);
                L.conflicts += passes_minus_one(zs);
                if (r == 0) L.conflicts += passes_minus_one(xs);
            }
    }
    return L;
}

}  // namespace amdat

extern "C" int amdat_internal_plan_layout(const int *nvec, int K, int R, int pair, uint32_t *pack, int *dst,
                                          int *range, int *n_chunks, long *conflicts)
{
    if (!nvec || K < 1 || (R != 2 && R != 4 && R != 8) || !n_chunks || pair < 0 || pair > 3) return AMDAT_EINVAL;
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
