// abi.hip -- C ABI of the device path (include/amdat_hip.h): contexts,
// device-resident trajectories, shell tables, pair tables and densities,
// and the entry points that launch the kernels of kernels.hip.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/amdat_hip.h"
#include "abi_internal.h"
#include "kernels.h"

namespace amdat {

namespace {
thread_local std::string g_last_error;
}

int set_error(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

void clear_error() { g_last_error.clear(); }

namespace {

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_error(e_ == hipErrorOutOfMemory ? AMDAT_ENOMEM : AMDAT_EDEVICE,  \
                             std::string(#expr) + ": " + hipGetErrorString(e_));        \
    } while (0)

// Owning device buffer.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { reset(); }
    void reset()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count)
    {
        reset();
        if (count == 0) return hipSuccess;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e == hipSuccess) n = count;
        else p = nullptr;
        return e;
    }
    hipError_t upload(const T *host, size_t count, hipStream_t st)
    {
        hipError_t e = alloc(count);
        if (e != hipSuccess || count == 0) return e;
        e = hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return e;
        return hipStreamSynchronize(st);
    }
};

}  // namespace

// Device scratch of amdat_msd_ngp_noalloc: the per-(pair, atom block)
// partial sums and the per-gap results of one call.  Each GPU the library
// runs on holds its own copy until the process ends.  Host code never
// passes this name to a kernel or a copy: it stands for a host-side shadow,
// not a device address; scratch_address() gives the address on a device.
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::unordered_map<int, void *> g_scratch_addr;  // device -> that device's copy of scratch

// Address of `scratch` on `device`, which must be the current device;
// looked up once per device and kept.
int scratch_address(int device, void **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    if (it != g_scratch_addr.end()) {
        *out = it->second;
        return AMDAT_OK;
    }
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(AMDAT_EDEVICE, std::string("scratch address lookup: ") + hipGetErrorString(e));
    g_scratch_addr[device] = p;
    *out = p;
    return AMDAT_OK;
}
}  // namespace
}  // namespace amdat

using namespace amdat;

struct amdat_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    void *scratch = nullptr;  // this device's copy of amdat::scratch
    // Grow-only device scratch (ctx_scratch), for entry points that
    // synchronise their stream before returning: one call's kernels are done
    // with it before the next call writes it.
    DevBuf<char> ws[6];
};

namespace {
template <typename T>
hipError_t ctx_scratch(amdat_ctx *ctx, int slot, size_t count, T **out)
{
    DevBuf<char> &b = ctx->ws[slot];
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (b.n < bytes) {
        hipError_t e = b.alloc(bytes);
        if (e != hipSuccess) return e;
    }
    *out = reinterpret_cast<T *>(b.p);
    return hipSuccess;
}
}  // namespace

struct amdat_traj {
    TrajDev dev;
    std::vector<float> box;  // host copy [T][3]
    std::vector<float> lo;   // box lower bounds [T][3] (custom files; zeros otherwise)
    DevBuf<float> x, y, z, dbox;
    DevBuf<int16_t> ix, iy, iz;
};

struct amdat_list {
    int n = 0;        // entries over all list times
    int n_atoms = 0;
    DevBuf<int> idx;  // empty = all atoms in order (static lists only)
    // Time-varying lists (src/trajectory_list.h:22-30): n_times > 0 lists,
    // entries off[c] .. off[c+1]-1 of idx at list time c, and the list time
    // conv[t] of each frame (src/exptime_trajectory_list.cpp:52); a frame
    // whose list time lies outside [0, n_times) has no atoms.  n_times == 0
    // is a static list.
    int n_times = 0;
    std::vector<int> conv, off;
    // Pair-keyed lists (src/bin_dynamics_analysis.h:120-141): list time c
    // belongs to the frame pair (t0, t1) that pkey maps to c, and conv is
    // empty; a pair with no key has no atoms.
    bool paired = false;
    std::unordered_map<int64_t, int> pkey;
};

// A cell table [T][N] kept on the device (-1: the atom is in no bin).
struct amdat_cells {
    int N = 0, T = 0, n_cells = 0;
    DevBuf<int> cell;
};

struct KPlanOwned {
    KPlan kp;
    int mode = 0;  // plan_layout mode the plan was built in
    DevBuf<uint32_t> kpack;
    DevBuf<int> kdst;
    DevBuf<int4> range;
};

// Serial number of a table handle: cached batch jobs compare (address,
// serial), so a new handle at a freed handle's address is never mistaken
// for it.
static uint64_t next_serial()
{
    static std::atomic<uint64_t> c{1};
    return c.fetch_add(1);
}

struct amdat_shells {
    uint64_t serial = next_serial();
    int n_shells = 0;
    std::vector<int> off, nvec;
    std::map<std::tuple<int, int, int, int>, std::unique_ptr<KPlanOwned>> plans;  // (first, last, R, mode)
    std::map<std::tuple<int, int, int>, std::pair<int, int>> choice;               // (first, last, c) -> (R, mode)
    int hint_R = 0, hint_pair = 0;  // amdat_shells_plan_hint (0: automatic)
    DevBuf<int> d_nvec;             // the integer vectors on the device, uploaded on first use
    DevBuf<int> d_off;              // the shell offsets on the device, uploaded on first use
};

struct amdat_pairs {
    uint64_t serial = next_serial();
    int n = 0, ntg = 0;
    std::vector<int> tg, t0, t1;
    std::vector<int> group_off, order;  // pairs grouped by time gap, reference order within a gap
    DevBuf<int> d_t0, d_t1, d_group_off, d_order;
};

struct amdat_rho {
    int n_frames = 0, K = 0, n_list = 0, T = 0;
    int first_shell = 0, n_shells = 0;
    std::vector<int> frames;     // frame index per row
    std::vector<int> frame_row;  // row per frame, -1 absent
    std::vector<int> shell_off;  // relative to the first computed vector
    std::vector<int> row_n;      // atoms summed into each row (the list at that frame)
    DevBuf<double2> d;
    DevBuf<int> d_shell_off;
};

// Parameters of one batch job (amdat_batch_run); equal parameters reuse
// the cached plan and captured graph.
struct JobParams {
    const amdat_shells *sk_sh = nullptr;
    int sk_first = 0, sk_last = -1;
    std::vector<int> frames;
    const amdat_shells *fs_sh = nullptr;
    std::vector<int> fs_first, fs_last;
    const amdat_pairs *pairs = nullptr;
    bool msd = false;
    std::vector<float> ell;
    uint64_t sk_serial = 0, fs_serial = 0, pairs_serial = 0;
    bool operator==(const JobParams &o) const
    {
        return sk_sh == o.sk_sh && sk_first == o.sk_first && sk_last == o.sk_last && frames == o.frames &&
               fs_sh == o.fs_sh && fs_first == o.fs_first && fs_last == o.fs_last && pairs == o.pairs &&
               msd == o.msd && ell == o.ell && sk_serial == o.sk_serial && fs_serial == o.fs_serial &&
               pairs_serial == o.pairs_serial;
    }
};

// Launch settings of a batch job (the autotuner's knobs).
struct JobChoice {
    int RS = 0;  // S(k): vectors per thread
    int SM = 0;  // S(k): plan mode (0 unpaired, 1 paired, 4 grouped, 5 grouped with short waves)
    int FG = 0;  // S(k): frames per item (0: from the item target)
    int RF = 0;  // Fs: vectors per thread
    int PC = 0;  // Fs: pairs per item (0: all pairs of a time gap)
    int AF = 1;  // Fs: atom parts per pair chunk
};

// What the autotuner kept for one job, with its timings of the cost-model
// settings and of the kept ones (whole job, all trajectories).
struct TuneResult {
    JobChoice ch;
    float model_ms = 0.0f, tuned_ms = 0.0f;
};

// A planned job: device tables, queue, outputs and (optionally) the
// instantiated graph of its launch sequence.
struct JobCache {
    JobParams prm;
    JobChoice ch;  // as planned (FG resolved)
    int MM = 0;
    bool tuned = false;
    float model_ms = 0.0f, tuned_ms = 0.0f;
    int n = 0, ntg = 0, nsh_sk = 0, G = 0, max_nsh = 0, n_pc = 0;
    std::vector<int> out_off;              // Fs offsets per trajectory
    std::vector<double> sk_count, fs_count, msd_count;
    std::vector<char> msd_item;            // [n*ntg]: an MSD item wrote (b, g)
    DevBuf<BatchTraj> sk_bt, fs_bt;
    DevBuf<int4> items;
    DevBuf<int> frames, sk_soff, fs_soff, fs_base, fs_nsh, fs_oo;
    DevBuf<uint32_t> fs_kpack;
    DevBuf<int> fs_kdst, fs_pcoff, fs_rowoff;
    DevBuf<int4> fs_range;
    DevBuf<double> sk_rows, fs_rows, sk_norm, fs_norm, S, Fs, s2, s4;
    BatchJob job;
    JobLaunch L;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~JobCache()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

struct amdat_batch {
    int n = 0, T = 0;
    std::vector<const amdat_traj *> trajs;
    std::vector<const amdat_list *> lists;
    DevBuf<int> counter;  // work-queue counter of the persistent launches
    std::unique_ptr<JobCache> job;
    std::vector<std::pair<JobParams, TuneResult>> tuned;  // autotuned settings per job
    bool forced = false;  // amdat_batch_use_settings
    JobChoice force;      // zero fields: the cost model's
};

namespace {

hipStream_t stream_of(amdat_ctx *c) { return c->stream ? c->stream : c->own; }

// Vectors per thread R (k-vectors per block: KB = 512 R) for K vectors.
// Each (atom, chunk) costs KB padded updates at c VALU ops per lane plus a
// fixed part (tile staging, table build, barriers, folds) measured at about
// 12,000 lane-ops: a 14,543-vector S(k) at R = 2 ran 1.63x slower than at
// R = 8.  c = 9 for rho (complex sums), 6 for Fs (real part only).
int choose_R(int K, int M, double c = 9.0)
{
    (void)M;
    static const int Rs[] = {2, 4, 8};
    int best = 2;
    double best_cost = 1e300;
    for (int R : Rs) {
        const int KB = kPhaseBlock * R;
        const int nch = (K + KB - 1) / KB;
        const double cost = (double)nch * ((double)KB * c + 12000.0);
        if (cost < best_cost * 0.98) { best_cost = cost; best = R; }
    }
    return best;
}

int shell_range_M(const amdat_shells *sh, int first, int last)
{
    int M = 0;
    for (int v = sh->off[first]; v < sh->off[last + 1]; v++)
        for (int c = 0; c < 3; c++) M = std::max(M, std::abs(sh->nvec[3 * v + c]));
    return M;
}

// Modelled cost of a layout per atom (PlanLayout::simd_cost): per chunk,
// the busiest SIMD (its two waves' slots x (c - 4 + 4/GS) VALU operations
// per lane, so uneven group sizes on one SIMD count) and the table build
// and anchors of the chunk's sections.
double layout_cost(const PlanLayout &L, double c)
{
    (void)c;
    return L.simd_cost;
}

// R_force > 0 fixes the vectors per thread and the layout mode to
// mode_force (batched launches need one R).  Otherwise the plan is the
// cheapest by layout_cost of R = 2, 4, 8 unpaired and, when allow_pair
// (callers whose kernels use pairs), R = 4, 8 paired (mode 1).
// c: VALU ops per update of the kernel the plan serves (see choose_R).
int build_plan(amdat_ctx *ctx, amdat_shells *sh, int first, int last, KPlanOwned **out, int R_force = 0,
               double c = 9.0, bool allow_pair = true, int mode_force = 0)
{
    const int v0 = sh->off[first], v1 = sh->off[last + 1];
    const int K = v1 - v0;
    int M = 0;
    for (int v = v0; v < v1; v++)
        for (int c = 0; c < 3; c++) M = std::max(M, std::abs(sh->nvec[3 * v + c]));
    if (M > 127) return set_error(AMDAT_ERANGE, "k-vector component beyond +-127");
    int R = R_force > 0 ? R_force : choose_R(K, M, c);
    int mode = R_force > 0 ? mode_force : 0;  // plan_layout mode: 0 unpaired, 1 paired with padding,
                                              // 2 paired and unpaired waves, 3 quads, 4 grouped
    PlanLayout lay;
    bool have_layout = false;
    if (R_force <= 0 && sh->hint_R > 0) {
        R = sh->hint_R;
        mode = allow_pair ? sh->hint_pair : 0;
    } else if (R_force <= 0 && K > 0) {
        const auto ckey = std::make_tuple(first, last, (int)(c * 16.0) * 2 + (allow_pair ? 1 : 0));
        auto ch = sh->choice.find(ckey);
        if (ch != sh->choice.end()) {
            R = ch->second.first;
            mode = ch->second.second;
        } else {
            double best = 1e300;
            int best_waves = 0;
            // Mode 2 is left to amdat_shells_plan_hint: its unpaired waves
            // gather the partnerless vectors of many nx groups, whose table
            // reads spread over the banks (64,000-atom workload: rho 542 ms
            // in mode 2 against 464 ms in mode 1, although it issues fewer
            // VALU ops).
            const int cand[8][2] = {{2, 0}, {4, 0}, {8, 0}, {4, 1}, {8, 1}, {4, 4}, {8, 4}, {8, 5}};
            for (const auto &cd : cand) {
                if (cd[1] && !allow_pair) continue;
                // Mode 5 evens out the waves' work at the cost of more
                // chunks; it pays where the kernel waits on LDS as well as
                // VALU (Fs: 238 -> 229 ms on the 64,000-atom workload), not
                // for rho, whose VALU issue is saturated (375 against 377
                // ms, with 2% more VALU operations and more bank conflicts)
                if (cd[1] == 5 && c > 7.0) continue;
                PlanLayout L =
                    plan_layout(sh->nvec.data() + 3 * (size_t)v0, K, cd[0], cd[1], plan_table_width(M), c, false);
                const double cost = layout_cost(L, c);
                // an exact tie goes to the layout that keeps at least as
                // many waves running the sums (a block with few has little
                // to hide its latencies with), then to the later candidate:
                // the scale form's 5 Fs shells (1,500 vectors) tie at R = 4
                // and 8 unpaired and R = 8 mode 5, which ran 9.5, 11.2 and
                // 9.2 s with 7, 4 and 7 such waves
                const bool tie = std::fabs(cost - best) <= 1e-9 * best && L.active_waves >= best_waves;
                if (cost < best * 0.98 || tie) {
                    best = cost;
                    best_waves = L.active_waves;
                    lay = std::move(L);
                    have_layout = cd[1] < 4;  // a grouped plan is placed once chosen
                    R = cd[0];
                    mode = cd[1];
                }
            }
            sh->choice[ckey] = {R, mode};
        }
    }
    auto key = std::make_tuple(first, last, R, mode);
    auto it = sh->plans.find(key);
    if (it != sh->plans.end()) { *out = it->second.get(); return AMDAT_OK; }

    auto P = std::make_unique<KPlanOwned>();
    KPlan &kp = P->kp;
    P->mode = mode;
    if (K == 0) {
        // empty shells (the 2D tables have some): no chunks, every sum is 0
        kp.R = R;
        kp.KB = kPhaseBlock * R;
        kp.MM = plan_table_width(0);
        *out = P.get();
        sh->plans[key] = std::move(P);
        return AMDAT_OK;
    }
    if (!have_layout) lay = plan_layout(sh->nvec.data() + 3 * (size_t)v0, K, R, mode, plan_table_width(M), c);
    kp.K = K;
    kp.R = R;
    kp.KB = lay.KB;
    kp.n_chunks = lay.n_chunks;
    kp.MM = plan_table_width(M);
    kp.pair = mode != 0 ? 1 : 0;
    hipStream_t st = stream_of(ctx);
    HIP_TRY(P->kpack.upload(lay.pack.data(), lay.pack.size(), st));
    HIP_TRY(P->kdst.upload(lay.dst.data(), lay.dst.size(), st));
    HIP_TRY(P->range.upload((const int4 *)lay.range.data(), lay.n_chunks, st));
    kp.kpack = P->kpack.p;
    kp.kdst = P->kdst.p;
    kp.range = P->range.p;
    *out = P.get();
    sh->plans[key] = std::move(P);
    return AMDAT_OK;
}

// Plan cost parameters of the kernels that read a plan, by precision:
// the table kernels (FP32 tables, FP64 tables) gain from grouped slots
// (c VALU per unpaired update: rho 9, Fs 6); the per-update sincos
// kernels (plain FP64, exact-phase FP32) do not, and pay as much for a
// padding slot as for a vector, so they take ungrouped plans with few
// padding slots (c: tens of operations per update).
struct PlanCost {
    double c;
    bool grouped;
};
PlanCost plan_cost(int precision, bool fs)
{
    if (precision == AMDAT_PREC_FP64 || precision == AMDAT_PREC_FP32X) return {40.0, false};
    return {fs ? 6.0 : 9.0, true};
}

// Atom splits so that the launch holds enough blocks to fill the chip,
// while each split still streams at least 8 tiles of atoms.
int choose_splits(long blocks_without_split, int n_list)
{
    const long target = 2048;
    int s = 1;
    if (blocks_without_split < target) s = (int)((target + blocks_without_split - 1) / blocks_without_split);
    const int max_s = std::max(1, n_list / 256);
    return std::max(1, std::min(s, max_s));
}

const int *list_ptr(const amdat_list *l) { return l && l->idx.p ? l->idx.p : nullptr; }
int list_n(const amdat_traj *tr, const amdat_list *l) { return l ? l->n : tr->dev.N; }
bool list_timed(const amdat_list *l) { return l && l->n_times > 0; }

// List time of frame t (0 for a static list, -1 where a time-varying list
// holds no atoms).
int list_time(const amdat_list *l, int t)
{
    if (!list_timed(l)) return 0;
    const int c = (t >= 0 && t < (int)l->conv.size()) ? l->conv[t] : -1;
    return (c >= 0 && c < l->n_times) ? c : -1;
}

// The atoms of a list at one list time: device pointer (nullptr = all atoms
// in order) and count.
struct ListSeg {
    const int *p;
    int n;
};
ListSeg list_seg(const amdat_traj *tr, const amdat_list *l, int c)
{
    if (!list_timed(l)) return {list_ptr(l), list_n(tr, l)};
    if (c < 0) return {nullptr, 0};
    const int n = l->off[c + 1] - l->off[c];
    return {n > 0 ? l->idx.p + l->off[c] : nullptr, n};
}
ListSeg list_at(const amdat_traj *tr, const amdat_list *l, int t) { return list_seg(tr, l, list_time(l, t)); }

// List time of the frame pair (t0, t1): the pair's key for a pair-keyed
// list, else the list time of t0 (src/trajectory_list.cpp:202-211).
int64_t pair_key(int t0, int t1) { return ((int64_t)t0 << 32) | (uint32_t)t1; }
int pair_time(const amdat_list *l, int t0, int t1)
{
    if (!l || !l->paired) return list_time(l, t0);
    const auto it = l->pkey.find(pair_key(t0, t1));
    return it == l->pkey.end() ? -1 : it->second;
}
ListSeg list_at_pair(const amdat_traj *tr, const amdat_list *l, int t0, int t1)
{
    return list_seg(tr, l, pair_time(l, t0, t1));
}
// Entry points over frames take lists per frame: static or time-varying.
bool frame_list_ok(const amdat_list *l) { return !l || !l->paired; }

// Atom blocks per pair of the MSD kernel: 1,024 atoms per block, at most 64.
int msd_blocks(int n_list) { return std::max(1, std::min(64, (n_list + kBlock * 4 - 1) / (kBlock * 4))); }

// A list indexes the atoms of the trajectory it was built for: one built
// for another atom count would send indices past this trajectory's frames.
int check_list_traj(const amdat_list *l, const amdat_traj *tr, const char *fn)
{
    if (l && l->n_atoms != tr->dev.N)
        return set_error(AMDAT_EINVAL, std::string(fn) + ": list built for a trajectory of another size");
    return AMDAT_OK;
}

// Pairs regrouped by the list time of their t0, for per-list-time launches:
// pos[i] is pair i's row in the regrouped order (stable within a list
// time), seg[g] = {list time, first row, row count} per group, and
// t0/t1/tg hold the regrouped pairs.
// Launch plan of a pair-based entry point (Fs, MSD/NGP, displacements):
// rows [row, row + len) of the per-pair partial buffers sum over the atoms
// of seg; t0/t1/tg are the device pairs in row order and order maps the
// reduction's reference-order positions to rows.  A static list is one
// group over the pairs in their own order.  A time-varying list (the list
// at the list time of t0, src/trajectory_list.cpp:202-211) or a pair-keyed
// list is one group as well, over the list's entries, with a per-pair
// segment table: pair p takes entries [pseg[p].x, pseg[p].x + pseg[p].y)
// (one launch for every pair, whatever its list time or key).  count[g]
// receives the atom-pair count of each time gap
// (src/mean_square_displacement.cpp:157; src/incoherent_scattering_
// function.cpp:283).
struct PairLaunch {
    struct Group {
        ListSeg seg;
        int row, len;
    };
    std::vector<Group> groups;
    const int *t0 = nullptr, *t1 = nullptr, *tg = nullptr, *order = nullptr;
    const int2 *pseg = nullptr;
    int max_n = 0;
    bool any_empty = false;
    DevBuf<int> d_tg;
    DevBuf<int2> d_pseg;
};
int plan_pair_launch(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *l, const amdat_pairs *pairs,
                     bool need_tg, PairLaunch &pl, double *count)
{
    hipStream_t st = stream_of(ctx);
    for (int g = 0; g < pairs->ntg; g++) count[g] = 0.0;
    for (int i = 0; i < pairs->n; i++)
        count[pairs->tg[i]] += (double)list_at_pair(tr, l, pairs->t0[i], pairs->t1[i]).n;
    pl.t0 = pairs->d_t0.p;
    pl.t1 = pairs->d_t1.p;
    pl.order = pairs->d_order.p;
    if (need_tg) HIP_TRY(pl.d_tg.upload(pairs->tg.data(), pairs->tg.size(), st));
    pl.tg = pl.d_tg.p;
    if (!list_timed(l)) {
        const ListSeg sg = list_seg(tr, l, 0);
        pl.groups.push_back({sg, 0, pairs->n});
        pl.max_n = sg.n;
        pl.any_empty = sg.n == 0;
        return AMDAT_OK;
    }
    std::vector<int2> seg(pairs->n, make_int2(0, 0));
    for (int i = 0; i < pairs->n; i++) {
        const int c = pair_time(l, pairs->t0[i], pairs->t1[i]);
        if (c >= 0 && c < l->n_times) seg[i] = make_int2(l->off[c], l->off[c + 1] - l->off[c]);
        pl.max_n = std::max(pl.max_n, seg[i].y);
        if (seg[i].y == 0) pl.any_empty = true;
    }
    HIP_TRY(pl.d_pseg.upload(seg.data(), seg.size(), st));
    pl.pseg = pl.d_pseg.p;
    pl.groups.push_back({ListSeg{l->idx.p, pl.max_n}, 0, pairs->n});
    return AMDAT_OK;
}

}  // namespace

extern "C" {

const char *amdat_last_error(void) { return g_last_error.c_str(); }

static int amdat_ctx_create_body(int device, amdat_ctx **out)
{
    if (!out) return set_error(AMDAT_EINVAL, "amdat_ctx_create: null output");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_error(AMDAT_EINVAL, "amdat_ctx_create: no such device");
    HIP_TRY(hipSetDevice(device));
    auto c = std::make_unique<amdat_ctx>();
    c->device = device;
    int rc = scratch_address(device, &c->scratch);
    if (rc) return rc;
    HIP_TRY(hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking));
    *out = c.release();
    return AMDAT_OK;
}

int amdat_ctx_create(int device, amdat_ctx **out)
{
    AMDAT_GUARD("amdat_ctx_create", amdat_ctx_create_body(device, out))
}

static int amdat_ctx_destroy_body(amdat_ctx *ctx)
{
    if (!ctx) return AMDAT_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    delete ctx;
    return AMDAT_OK;
}

int amdat_ctx_destroy(amdat_ctx *ctx)
{
    AMDAT_GUARD("amdat_ctx_destroy", amdat_ctx_destroy_body(ctx))
}

static int amdat_ctx_set_stream_body(amdat_ctx *ctx, void *stream)
{
    if (!ctx) return set_error(AMDAT_EINVAL, "null context");
    ctx->stream = (hipStream_t)stream;
    return AMDAT_OK;
}

int amdat_ctx_set_stream(amdat_ctx *ctx, void *stream)
{
    AMDAT_GUARD("amdat_ctx_set_stream", amdat_ctx_set_stream_body(ctx, stream))
}

static int amdat_ctx_synchronize_body(amdat_ctx *ctx)
{
    if (!ctx) return set_error(AMDAT_EINVAL, "null context");
    HIP_TRY(hipStreamSynchronize(stream_of(ctx)));
    return AMDAT_OK;
}

int amdat_ctx_synchronize(amdat_ctx *ctx)
{
    AMDAT_GUARD("amdat_ctx_synchronize", amdat_ctx_synchronize_body(ctx))
}

static int amdat_shells_upload_body(amdat_ctx *ctx, int n_shells, const int *shell_off, const int *nvec, amdat_shells **out)
{
    if (!ctx || !shell_off || !out || n_shells < 1) return set_error(AMDAT_EINVAL, "amdat_shells_upload: bad arguments");
    const int K = shell_off[n_shells];
    if (shell_off[0] != 0 || K < 0 || (K > 0 && !nvec)) return set_error(AMDAT_EINVAL, "amdat_shells_upload: bad offsets");
    for (int s = 0; s < n_shells; s++)
        if (shell_off[s + 1] < shell_off[s]) return set_error(AMDAT_EINVAL, "amdat_shells_upload: offsets not ascending");
    auto sh = std::make_unique<amdat_shells>();
    sh->n_shells = n_shells;
    sh->off.assign(shell_off, shell_off + n_shells + 1);
    sh->nvec.assign(nvec, nvec + 3 * (size_t)K);
    for (int v = 0; v < 3 * K; v++)
        if (std::abs(sh->nvec[v]) > 127) return set_error(AMDAT_ERANGE, "amdat_shells_upload: component beyond +-127");
    *out = sh.release();
    return AMDAT_OK;
}

int amdat_shells_upload(amdat_ctx *ctx, int n_shells, const int *shell_off, const int *nvec, amdat_shells **out)
{
    AMDAT_GUARD("amdat_shells_upload", amdat_shells_upload_body(ctx, n_shells, shell_off, nvec, out))
}

static int amdat_shells_plan_hint_body(amdat_shells *sh, int R, int pair)
{
    clear_error();
    if (!sh) return set_error(AMDAT_EINVAL, "amdat_shells_plan_hint: null table");
    if (R != 0 && R != 2 && R != 4 && R != 8) return set_error(AMDAT_EINVAL, "amdat_shells_plan_hint: R must be 0, 2, 4 or 8");
    if (pair < 0 || pair > 5) return set_error(AMDAT_EINVAL, "amdat_shells_plan_hint: pair must be 0 to 5");
    if (pair == 5 && R != 8) return set_error(AMDAT_EINVAL, "amdat_shells_plan_hint: mode 5 needs R = 8");
    if (pair && R != 4 && R != 8) return set_error(AMDAT_EINVAL, "amdat_shells_plan_hint: pairs need R = 4 or 8");
    sh->hint_R = R;
    sh->hint_pair = pair;
    return AMDAT_OK;
}

int amdat_shells_plan_hint(amdat_shells *sh, int R, int pair)
{
    AMDAT_GUARD("amdat_shells_plan_hint", amdat_shells_plan_hint_body(sh, R, pair))
}

static int amdat_shells_free_body(amdat_ctx *ctx_unused, amdat_shells *sh)
{
    (void)ctx_unused;
    delete sh;
    return AMDAT_OK;
}

int amdat_shells_free(amdat_ctx *ctx_unused, amdat_shells *sh)
{
    AMDAT_GUARD("amdat_shells_free", amdat_shells_free_body(ctx_unused, sh))
}

static int traj_finish(amdat_ctx *ctx, amdat_traj *t, int n_atoms, int n_frames, const float *box)
{
    t->dev.N = n_atoms;
    t->dev.T = n_frames;
    t->box.assign(box, box + 3 * (size_t)n_frames);
    if (t->lo.size() != t->box.size()) t->lo.assign(t->box.size(), 0.0f);
    HIP_TRY(t->dbox.upload(t->box.data(), t->box.size(), stream_of(ctx)));
    t->dev.box = t->dbox.p;
    return AMDAT_OK;
}

static int amdat_traj_upload_body(amdat_ctx *ctx, int n_atoms, int n_frames, const float *x, const float *y, const float *z,
                      const int16_t *ix, const int16_t *iy, const int16_t *iz, const float *box, amdat_traj **out)
{
    if (!ctx || !out || n_atoms < 1 || n_frames < 1 || !x || !y || !z || !ix || !iy || !iz || !box)
        return set_error(AMDAT_EINVAL, "amdat_traj_upload: bad arguments");
    HIP_TRY(hipSetDevice(ctx->device));
    auto t = std::make_unique<amdat_traj>();
    const size_t n = (size_t)n_atoms * n_frames;
    hipStream_t st = stream_of(ctx);
    HIP_TRY(t->x.upload(x, n, st));
    HIP_TRY(t->y.upload(y, n, st));
    HIP_TRY(t->z.upload(z, n, st));
    HIP_TRY(t->ix.upload(ix, n, st));
    HIP_TRY(t->iy.upload(iy, n, st));
    HIP_TRY(t->iz.upload(iz, n, st));
    t->dev.x = t->x.p; t->dev.y = t->y.p; t->dev.z = t->z.p;
    t->dev.ix = t->ix.p; t->dev.iy = t->iy.p; t->dev.iz = t->iz.p;
    int rc = traj_finish(ctx, t.get(), n_atoms, n_frames, box);
    if (rc) return rc;
    *out = t.release();
    return AMDAT_OK;
}

int amdat_traj_upload(amdat_ctx *ctx, int n_atoms, int n_frames, const float *x, const float *y, const float *z,
                      const int16_t *ix, const int16_t *iy, const int16_t *iz, const float *box, amdat_traj **out)
{
    AMDAT_GUARD("amdat_traj_upload", amdat_traj_upload_body(ctx, n_atoms, n_frames, x, y, z, ix, iy, iz, box, out))
}

static int amdat_traj_wrap_device_body(amdat_ctx *ctx, int n_atoms, int n_frames, const float *d_x, const float *d_y,
                           const float *d_z, const int16_t *d_ix, const int16_t *d_iy, const int16_t *d_iz,
                           const float *box_host, amdat_traj **out)
{
    if (!ctx || !out || n_atoms < 1 || n_frames < 1 || !d_x || !d_y || !d_z || !d_ix || !d_iy || !d_iz || !box_host)
        return set_error(AMDAT_EINVAL, "amdat_traj_wrap_device: bad arguments");
    auto t = std::make_unique<amdat_traj>();
    t->dev.x = d_x; t->dev.y = d_y; t->dev.z = d_z;
    t->dev.ix = d_ix; t->dev.iy = d_iy; t->dev.iz = d_iz;
    int rc = traj_finish(ctx, t.get(), n_atoms, n_frames, box_host);
    if (rc) return rc;
    *out = t.release();
    return AMDAT_OK;
}

int amdat_traj_wrap_device(amdat_ctx *ctx, int n_atoms, int n_frames, const float *d_x, const float *d_y,
                           const float *d_z, const int16_t *d_ix, const int16_t *d_iy, const int16_t *d_iz,
                           const float *box_host, amdat_traj **out)
{
    AMDAT_GUARD("amdat_traj_wrap_device", amdat_traj_wrap_device_body(ctx, n_atoms, n_frames, d_x, d_y, d_z, d_ix, d_iy, d_iz, box_host, out))
}

static int amdat_traj_free_body(amdat_ctx *ctx_unused, amdat_traj *tr)
{
    (void)ctx_unused;
    delete tr;
    return AMDAT_OK;
}

int amdat_traj_free(amdat_ctx *ctx_unused, amdat_traj *tr)
{
    AMDAT_GUARD("amdat_traj_free", amdat_traj_free_body(ctx_unused, tr))
}

static int amdat_traj_shape_body(const amdat_traj *tr, int *n_atoms, int *n_frames, float *box)
{
    if (!tr) return set_error(AMDAT_EINVAL, "amdat_traj_shape: null trajectory");
    if (n_atoms) *n_atoms = tr->dev.N;
    if (n_frames) *n_frames = tr->dev.T;
    if (box) std::copy(tr->box.begin(), tr->box.end(), box);
    return AMDAT_OK;
}

int amdat_traj_shape(const amdat_traj *tr, int *n_atoms, int *n_frames, float *box)
{
    AMDAT_GUARD("amdat_traj_shape", amdat_traj_shape_body(tr, n_atoms, n_frames, box))
}

static int amdat_traj_bounds_body(const amdat_traj *tr, float *lo)
{
    if (!tr || !lo) return set_error(AMDAT_EINVAL, "amdat_traj_bounds: null argument");
    std::copy(tr->lo.begin(), tr->lo.end(), lo);
    return AMDAT_OK;
}

int amdat_traj_bounds(const amdat_traj *tr, float *lo)
{
    AMDAT_GUARD("amdat_traj_bounds", amdat_traj_bounds_body(tr, lo))
}

}  // extern "C"

namespace {

// Rectangular cells into dc ([T][N] on the device).
int cells_rect(amdat_ctx *ctx, const amdat_traj *tr, int n_xbins, int n_ybins, int n_zbins, const float *lo,
               const float *hi, DevBuf<int> &dc, const char *fn)
{
    if (!ctx || !tr || !lo || !hi) return set_error(AMDAT_EINVAL, std::string(fn) + ": null argument");
    if (n_xbins < 1 || n_ybins < 1 || n_zbins < 1 || (long)n_xbins * n_ybins * n_zbins > (1L << 30))
        return set_error(AMDAT_EINVAL, std::string(fn) + ": bin counts must be >= 1");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = stream_of(ctx);
    const int N = tr->dev.N, T = tr->dev.T;
    DevBuf<float> dlo, dhi;
    HIP_TRY(dlo.upload(lo, 3 * (size_t)T, st));
    HIP_TRY(dhi.upload(hi, 3 * (size_t)T, st));
    HIP_TRY(dc.alloc((size_t)N * T));
    HIP_TRY(launch_bins_assign(tr->dev, n_xbins, n_ybins, n_zbins, dlo.p, dhi.p, dc.p, st));
    HIP_TRY(hipStreamSynchronize(st));
    return AMDAT_OK;
}

}  // namespace

extern "C" {

static int amdat_bins_assign_body(amdat_ctx *ctx, const amdat_traj *tr, int n_xbins, int n_ybins, int n_zbins, const float *lo,
                      const float *hi, int *cell)
{
    clear_error();
    if (!cell) return set_error(AMDAT_EINVAL, "amdat_bins_assign: null argument");
    DevBuf<int> dc;
    int rc = cells_rect(ctx, tr, n_xbins, n_ybins, n_zbins, lo, hi, dc, "amdat_bins_assign");
    if (rc) return rc;
    hipStream_t st = stream_of(ctx);
    HIP_TRY(hipMemcpyAsync(cell, dc.p, sizeof(int) * dc.n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return AMDAT_OK;
}

int amdat_bins_assign(amdat_ctx *ctx, const amdat_traj *tr, int n_xbins, int n_ybins, int n_zbins, const float *lo,
                      const float *hi, int *cell)
{
    AMDAT_GUARD("amdat_bins_assign", amdat_bins_assign_body(ctx, tr, n_xbins, n_ybins, n_zbins, lo, hi, cell))
}

static int amdat_cells_assign_body(amdat_ctx *ctx, const amdat_traj *tr, int n_xbins, int n_ybins, int n_zbins, const float *lo,
                       const float *hi, amdat_cells **out)
{
    clear_error();
    if (!out) return set_error(AMDAT_EINVAL, "amdat_cells_assign: null output");
    auto c = std::make_unique<amdat_cells>();
    int rc = cells_rect(ctx, tr, n_xbins, n_ybins, n_zbins, lo, hi, c->cell, "amdat_cells_assign");
    if (rc) return rc;
    c->N = tr->dev.N;
    c->T = tr->dev.T;
    c->n_cells = n_xbins * n_ybins * n_zbins;
    *out = c.release();
    return AMDAT_OK;
}

int amdat_cells_assign(amdat_ctx *ctx, const amdat_traj *tr, int n_xbins, int n_ybins, int n_zbins, const float *lo,
                       const float *hi, amdat_cells **out)
{
    AMDAT_GUARD("amdat_cells_assign", amdat_cells_assign_body(ctx, tr, n_xbins, n_ybins, n_zbins, lo, hi, out))
}

}  // extern "C"

namespace {
// Page-locked host buffer (hipHostMalloc), freed on scope exit.
struct PinnedBuf {
    void *p = nullptr;
    ~PinnedBuf()
    {
        if (p) (void)hipHostFree(p);
    }
};
}  // namespace

// Per frame, the entries of a frame list: {first entry of idx, count}; the
// index array (null = all atoms in order) and the largest count.
static const int *frame_segs(const amdat_traj *tr, const amdat_list *l, std::vector<int2> &seg, int &max_n)
{
    const int T = tr->dev.T;
    seg.assign(T, make_int2(0, 0));
    max_n = 0;
    for (int t = 0; t < T; t++) {
        if (!list_timed(l)) {
            seg[t] = make_int2(0, list_n(tr, l));
        } else {
            const int c = list_time(l, t);
            if (c >= 0) seg[t] = make_int2(l->off[c], l->off[c + 1] - l->off[c]);
        }
        max_n = std::max(max_n, seg[t].y);
    }
    return list_ptr(l);
}

namespace {

// Distance cells into dc ([T][N] on the device); messages name fn.
int cells_dist(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *binned, int kind, const float *point,
               int axis, int direction, const amdat_list *cluster, float thickness, int n_bins, DevBuf<int> &dc,
               const char *fn)
{
    const std::string F(fn);
    if (!ctx || !tr) return set_error(AMDAT_EINVAL, F + ": null argument");
    if (kind < AMDAT_DIST_POINT || kind > AMDAT_DIST_LIST)
        return set_error(AMDAT_EINVAL, F + ": unknown kind");
    if (n_bins < 1 || !(thickness > 0.0f))
        return set_error(AMDAT_EINVAL, F + ": need n_bins >= 1 and thickness > 0");
    if (!frame_list_ok(binned) || !frame_list_ok(cluster))
        return set_error(AMDAT_EINVAL, F + ": a pair-keyed list applies to frame pairs only");
    if ((binned && binned->n_atoms != tr->dev.N) || (cluster && cluster->n_atoms != tr->dev.N))
        return set_error(AMDAT_EINVAL, F + ": list built for another trajectory size");
    const int N = tr->dev.N, T = tr->dev.T;
    float p[3] = {0.0f, 0.0f, 0.0f};
    if (kind == AMDAT_DIST_POINT) {
        if (!point) return set_error(AMDAT_EINVAL, F + ": null point");
        p[0] = point[0], p[1] = point[1], p[2] = point[2];
        // src/trajectory_list_bins.cpp:1358-1362 (Coordinate::operator>)
        for (int t = 0; t < T; t++)
            if (p[0] > tr->box[3 * t] || p[1] > tr->box[3 * t + 1] || p[2] > tr->box[3 * t + 2])
                return set_error(AMDAT_ERANGE, F + ": point chosen does not lie within the box dimensions");
    } else if (kind == AMDAT_DIST_PLANE) {
        if (!point || axis < 0 || axis > 2 || direction < 0 || direction > 2)
            return set_error(AMDAT_EINVAL, F + ": plane needs a position, axis 0-2 and direction 0-2");
        p[0] = point[0];
        for (int t = 0; t < T; t++)
            if (p[0] > tr->box[3 * t + axis])
                return set_error(AMDAT_ERANGE, F + ": plane chosen does not lie within the box dimensions");
    }
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = stream_of(ctx);
    std::vector<int2> seg, cseg;
    int max_n = 0, max_c = 0;
    const int *idx = frame_segs(tr, binned, seg, max_n);
    DevBuf<int2> dseg, dcseg;
    HIP_TRY(dseg.upload(seg.data(), seg.size(), st));
    HIP_TRY(dc.alloc((size_t)N * T));
    HIP_TRY(hipMemsetAsync(dc.p, 0xFF, sizeof(int) * (size_t)N * T, st));
    if (kind == AMDAT_DIST_LIST) {
        const int *cidx = frame_segs(tr, cluster, cseg, max_c);
        HIP_TRY(dcseg.upload(cseg.data(), cseg.size(), st));
        HIP_TRY(launch_cluster_cells(tr->dev, dseg.p, idx, max_n, dcseg.p, cidx, thickness, n_bins, dc.p, st));
    } else {
        HIP_TRY(launch_distance_cells(tr->dev, dseg.p, idx, max_n, kind, axis, direction, p, thickness, n_bins, dc.p,
                                      st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    return AMDAT_OK;
}

}  // namespace

extern "C" {

static int amdat_bins_distance_body(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *binned, int kind,
                        const float *point, int axis, int direction, const amdat_list *cluster, float thickness,
                        int n_bins, int *cell)
{
    clear_error();
    if (!cell) return set_error(AMDAT_EINVAL, "amdat_bins_distance: null argument");
    DevBuf<int> dc;
    int rc = cells_dist(ctx, tr, binned, kind, point, axis, direction, cluster, thickness, n_bins, dc,
                        "amdat_bins_distance");
    if (rc) return rc;
    hipStream_t st = stream_of(ctx);
    HIP_TRY(hipMemcpyAsync(cell, dc.p, sizeof(int) * dc.n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return AMDAT_OK;
}

int amdat_bins_distance(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *binned, int kind,
                        const float *point, int axis, int direction, const amdat_list *cluster, float thickness,
                        int n_bins, int *cell)
{
    AMDAT_GUARD("amdat_bins_distance", amdat_bins_distance_body(ctx, tr, binned, kind, point, axis, direction, cluster, thickness, n_bins, cell))
}

static int amdat_cells_distance_body(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *binned, int kind,
                         const float *point, int axis, int direction, const amdat_list *cluster, float thickness,
                         int n_bins, amdat_cells **out)
{
    clear_error();
    if (!out) return set_error(AMDAT_EINVAL, "amdat_cells_distance: null output");
    auto c = std::make_unique<amdat_cells>();
    int rc = cells_dist(ctx, tr, binned, kind, point, axis, direction, cluster, thickness, n_bins, c->cell,
                        "amdat_cells_distance");
    if (rc) return rc;
    c->N = tr->dev.N;
    c->T = tr->dev.T;
    c->n_cells = n_bins;
    *out = c.release();
    return AMDAT_OK;
}

int amdat_cells_distance(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *binned, int kind,
                         const float *point, int axis, int direction, const amdat_list *cluster, float thickness,
                         int n_bins, amdat_cells **out)
{
    AMDAT_GUARD("amdat_cells_distance", amdat_cells_distance_body(ctx, tr, binned, kind, point, axis, direction, cluster, thickness, n_bins, out))
}

static int amdat_cells_download_body(amdat_ctx *ctx, const amdat_cells *c, int *cell)
{
    clear_error();
    if (!ctx || !c || !cell) return set_error(AMDAT_EINVAL, "amdat_cells_download: null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = stream_of(ctx);
    HIP_TRY(hipMemcpyAsync(cell, c->cell.p, sizeof(int) * c->cell.n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return AMDAT_OK;
}

int amdat_cells_download(amdat_ctx *ctx, const amdat_cells *c, int *cell)
{
    AMDAT_GUARD("amdat_cells_download", amdat_cells_download_body(ctx, c, cell))
}

static int amdat_cells_free_body(amdat_ctx *ctx, amdat_cells *c)
{
    if (ctx && c) (void)hipSetDevice(ctx->device);
    delete c;
    return AMDAT_OK;
}

int amdat_cells_free(amdat_ctx *ctx, amdat_cells *c)
{
    AMDAT_GUARD("amdat_cells_free", amdat_cells_free_body(ctx, c))
}

}  // extern "C"

namespace {

// Per-bin lists of `rows` (int4: frame of the cell, frame whose cell must
// match or -1, row of the list mask) for bins [first, first + n_bins) of the
// table: out[b]->off[r .. r+1] and idx receive the row's atoms of bin
// first + b in ascending order.  l gives the mask: NULL = every atom.
int cells_lists(amdat_ctx *ctx, const amdat_cells *c, const amdat_list *l, std::vector<int4> &rows, int first,
                int n_bins, std::vector<std::unique_ptr<amdat_list>> &out)
{
    hipStream_t st = stream_of(ctx);
    const int N = c->N, nr = (int)rows.size();
    // list mask rows: static list 0 / every atom -2 / no atom -1; time-varying list: its list time
    DevBuf<uint8_t> mask;
    const bool explicit_idx = l && l->n > 0 && l->idx.p;
    if (explicit_idx) {
        const int nt = list_timed(l) ? l->n_times : 1;
        std::vector<int> off = list_timed(l) ? l->off : std::vector<int>{0, l->n};
        DevBuf<int> doff;
        HIP_TRY(doff.upload(off.data(), off.size(), st));
        HIP_TRY(mask.alloc((size_t)nt * N));
        HIP_TRY(hipMemsetAsync(mask.p, 0, mask.n, st));
        HIP_TRY(launch_list_mask(l->idx.p, doff.p, nt, l->n, N, mask.p, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    for (int4 &r : rows) {
        if (!l) { r.z = -2; continue; }
        if (!list_timed(l)) { r.z = l->n == 0 ? -1 : (l->idx.p ? 0 : -2); continue; }
        const int ct = list_time(l, r.x);
        r.z = (ct < 0 || l->off[ct + 1] == l->off[ct]) ? -1 : ct;
    }
    DevBuf<int4> drows;
    HIP_TRY(drows.upload(rows.data(), rows.size(), st));
    out.clear();
    for (int b = 0; b < n_bins; b++) {
        out.push_back(std::make_unique<amdat_list>());
        out.back()->n_atoms = N;
        out.back()->off.assign(nr + 1, 0);
    }
    std::vector<int> cnt, pos;
    std::vector<int *> ptr;
    for (int w0 = 0; w0 < n_bins; w0 += kBinWindow) {
        const int nw = std::min(kBinWindow, n_bins - w0);
        DevBuf<int> dcnt, dpos;
        HIP_TRY(dcnt.alloc((size_t)nw * nr));
        HIP_TRY(hipMemsetAsync(dcnt.p, 0, sizeof(int) * dcnt.n, st));
        HIP_TRY(launch_bin_count(c->cell.p, N, drows.p, nr, mask.p, first + w0, nw, dcnt.p, st));
        cnt.resize((size_t)nw * nr);
        HIP_TRY(hipMemcpyAsync(cnt.data(), dcnt.p, sizeof(int) * cnt.size(), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        pos.resize(cnt.size());
        ptr.assign(nw, nullptr);
        for (int w = 0; w < nw; w++) {
            amdat_list &L = *out[w0 + w];
            for (int r = 0; r < nr; r++) {
                pos[(size_t)w * nr + r] = L.off[r];
                L.off[r + 1] = L.off[r] + cnt[(size_t)w * nr + r];
            }
            L.n = L.off[nr];
            if (L.n > 0) HIP_TRY(L.idx.alloc((size_t)L.n));
            ptr[w] = L.idx.p;
        }
        DevBuf<int *> dptr;
        HIP_TRY(dpos.upload(pos.data(), pos.size(), st));
        HIP_TRY(dptr.upload(ptr.data(), ptr.size(), st));
        HIP_TRY(launch_bin_write(c->cell.p, N, drows.p, nr, mask.p, first + w0, nw, dpos.p, dptr.p, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    return AMDAT_OK;
}

int cells_args(amdat_ctx *ctx, const amdat_cells *c, const amdat_list *l, int first_bin, int n_bins,
               amdat_list **lists, const char *fn)
{
    const std::string F(fn);
    if (!ctx || !c || !lists) return set_error(AMDAT_EINVAL, F + ": null argument");
    if (n_bins < 1 || first_bin < 0 || first_bin + n_bins > c->n_cells)
        return set_error(AMDAT_ERANGE, F + ": bin range outside the table");
    if (!frame_list_ok(l)) return set_error(AMDAT_EINVAL, F + ": a pair-keyed list applies to frame pairs only");
    if (l && l->n_atoms != c->N) return set_error(AMDAT_EINVAL, F + ": list built for another trajectory size");
    if (l && list_timed(l) && (int)l->conv.size() != c->T)
        return set_error(AMDAT_EINVAL, F + ": list built for a trajectory of another length");
    return AMDAT_OK;
}

}  // namespace

extern "C" {

static int amdat_cells_bin_lists_body(amdat_ctx *ctx, const amdat_cells *c, const amdat_list *l, int first_bin, int n_bins,
                          amdat_list **lists)
{
    clear_error();
    int rc = cells_args(ctx, c, l, first_bin, n_bins, lists, "amdat_cells_bin_lists");
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<int4> rows(c->T);
    for (int t = 0; t < c->T; t++) rows[t] = make_int4(t, -1, 0, 0);
    std::vector<std::unique_ptr<amdat_list>> out;
    rc = cells_lists(ctx, c, l, rows, first_bin, n_bins, out);
    if (rc) return rc;
    for (int b = 0; b < n_bins; b++) {
        out[b]->n_times = c->T;
        out[b]->conv.resize(c->T);
        std::iota(out[b]->conv.begin(), out[b]->conv.end(), 0);
        lists[b] = out[b].release();
    }
    return AMDAT_OK;
}

int amdat_cells_bin_lists(amdat_ctx *ctx, const amdat_cells *c, const amdat_list *l, int first_bin, int n_bins,
                          amdat_list **lists)
{
    AMDAT_GUARD("amdat_cells_bin_lists", amdat_cells_bin_lists_body(ctx, c, l, first_bin, n_bins, lists))
}

static int amdat_cells_pair_lists_body(amdat_ctx *ctx, const amdat_cells *c, const amdat_list *l, int n_keys, const int *t0,
                           const int *t1, int first_bin, int n_bins, amdat_list **lists)
{
    clear_error();
    int rc = cells_args(ctx, c, l, first_bin, n_bins, lists, "amdat_cells_pair_lists");
    if (rc) return rc;
    if (n_keys < 1 || !t0 || !t1) return set_error(AMDAT_EINVAL, "amdat_cells_pair_lists: bad frame pairs");
    std::unordered_map<int64_t, int> pkey;
    std::vector<int4> rows(n_keys);
    for (int k = 0; k < n_keys; k++) {
        if (t0[k] < 0 || t0[k] >= c->T || t1[k] < 0 || t1[k] >= c->T)
            return set_error(AMDAT_ERANGE, "amdat_cells_pair_lists: frame out of range");
        if (!pkey.emplace(pair_key(t0[k], t1[k]), k).second)
            return set_error(AMDAT_EINVAL, "amdat_cells_pair_lists: frame pair given twice");
        rows[k] = make_int4(t0[k], t1[k], 0, 0);
    }
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<std::unique_ptr<amdat_list>> out;
    rc = cells_lists(ctx, c, l, rows, first_bin, n_bins, out);
    if (rc) return rc;
    for (int b = 0; b < n_bins; b++) {
        out[b]->n_times = n_keys;
        out[b]->paired = true;
        out[b]->pkey = pkey;
        lists[b] = out[b].release();
    }
    return AMDAT_OK;
}

int amdat_cells_pair_lists(amdat_ctx *ctx, const amdat_cells *c, const amdat_list *l, int n_keys, const int *t0,
                           const int *t1, int first_bin, int n_bins, amdat_list **lists)
{
    AMDAT_GUARD("amdat_cells_pair_lists", amdat_cells_pair_lists_body(ctx, c, l, n_keys, t0, t1, first_bin, n_bins, lists))
}

static int amdat_traj_read_body(amdat_ctx *ctx, const char *format, const char *path, int n_atoms, int n_frames,
                    const float *box3, int threads, int *types, amdat_traj **out)
{
    clear_error();
    if (!ctx || !format || !path || !out || n_atoms < 1 || n_frames < 1)
        return set_error(AMDAT_EINVAL, "amdat_traj_read: bad arguments");
    const std::string fmt(format);
    if (fmt != "custom" && fmt != "xyz") return set_error(AMDAT_EINVAL, "amdat_traj_read: format must be custom or xyz");
    if (fmt == "xyz" && !box3) return set_error(AMDAT_EINVAL, "amdat_traj_read: xyz needs the box");
    HIP_TRY(hipSetDevice(ctx->device));
    MappedFile f;
    int rc = f.open(path);
    if (rc) return rc;
    std::vector<size_t> off;
    rc = fmt == "custom" ? index_custom(f, n_frames, threads, off) : index_xyz(f, n_atoms, n_frames, threads, off);
    if (rc) return rc;
    const int N = n_atoms, T = n_frames;
    const size_t n = (size_t)N * T;
    auto t = std::make_unique<amdat_traj>();
    HIP_TRY(t->x.alloc(n));
    HIP_TRY(t->y.alloc(n));
    HIP_TRY(t->z.alloc(n));
    HIP_TRY(t->ix.alloc(n));
    HIP_TRY(t->iy.alloc(n));
    HIP_TRY(t->iz.alloc(n));
    // chunks of about 64 MiB of parsed frames (or AMDAT_READ_CHUNK_FRAMES),
    // double-buffered in pinned memory
    int CF = (int)std::max<size_t>(1, std::min<size_t>((size_t)T, (64u << 20) / ((size_t)N * 18)));
    if (const char *e = std::getenv("AMDAT_READ_CHUNK_FRAMES")) CF = std::max(1, std::min(T, std::atoi(e)));
    const size_t cn = (size_t)CF * N;
    PinnedBuf stage[2];
    hipEvent_t done[2] = {nullptr, nullptr};
    for (int b = 0; b < 2; b++) {
        HIP_TRY(hipHostMalloc(&stage[b].p, cn * 18, hipHostMallocDefault));
        HIP_TRY(hipEventCreateWithFlags(&done[b], hipEventDisableTiming));
    }
    struct EvGuard {
        hipEvent_t *e;
        ~EvGuard()
        {
            for (int b = 0; b < 2; b++)
                if (e[b]) (void)hipEventDestroy(e[b]);
        }
    } evg{done};
    hipStream_t st = stream_of(ctx);
    std::vector<float> box((size_t)3 * T), lo((size_t)3 * T, 0.0f), prev_xyz;
    std::vector<int16_t> prev_img;
    for (int f0 = 0, c = 0; f0 < T; f0 += CF, c++) {
        const int f1 = std::min(T, f0 + CF), nf = f1 - f0, b = c & 1;
        if (c >= 2) HIP_TRY(hipEventSynchronize(done[b]));  // the copy out of this buffer has finished
        float *sx = (float *)stage[b].p, *sy = sx + cn, *sz = sy + cn;
        int16_t *six = (int16_t *)(sz + cn), *siy = six + cn, *siz = siy + cn;
        if (fmt == "custom") {
            rc = parse_custom(f, off, f0, f1, N, threads, sx, sy, sz, six, siy, siz, box.data() + 3 * (size_t)f0,
                              lo.data() + 3 * (size_t)f0, f0 == 0 ? types : nullptr);
        } else {
            rc = parse_xyz(f, off, f0, f1, N, threads, sx, sy, sz, f0 == 0 ? types : nullptr);
            if (!rc)
                rc = unwrap_images(N, nf, box3, sx, sy, sz, six, siy, siz, f0 ? prev_xyz.data() : nullptr,
                                   f0 ? prev_img.data() : nullptr, threads);
            if (!rc) {
                // the last frame of this chunk seeds the next one
                const size_t last = (size_t)(nf - 1) * N;
                prev_xyz.assign(sx + last, sx + last + N);
                prev_xyz.insert(prev_xyz.end(), sy + last, sy + last + N);
                prev_xyz.insert(prev_xyz.end(), sz + last, sz + last + N);
                prev_img.assign(six + last, six + last + N);
                prev_img.insert(prev_img.end(), siy + last, siy + last + N);
                prev_img.insert(prev_img.end(), siz + last, siz + last + N);
                for (int r = f0; r < f1; r++)
                    for (int k = 0; k < 3; k++) box[3 * (size_t)r + k] = box3[k];
            }
        }
        if (rc) {
            (void)hipStreamSynchronize(st);
            return rc;
        }
        const size_t o = (size_t)f0 * N, m = (size_t)nf * N;
        HIP_TRY(hipMemcpyAsync(t->x.p + o, sx, m * sizeof(float), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(t->y.p + o, sy, m * sizeof(float), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(t->z.p + o, sz, m * sizeof(float), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(t->ix.p + o, six, m * sizeof(int16_t), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(t->iy.p + o, siy, m * sizeof(int16_t), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(t->iz.p + o, siz, m * sizeof(int16_t), hipMemcpyHostToDevice, st));
        HIP_TRY(hipEventRecord(done[b], st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    t->dev.x = t->x.p; t->dev.y = t->y.p; t->dev.z = t->z.p;
    t->dev.ix = t->ix.p; t->dev.iy = t->iy.p; t->dev.iz = t->iz.p;
    t->lo = lo;
    rc = traj_finish(ctx, t.get(), N, T, box.data());
    if (rc) return rc;
    *out = t.release();
    return AMDAT_OK;
}

int amdat_traj_read(amdat_ctx *ctx, const char *format, const char *path, int n_atoms, int n_frames,
                    const float *box3, int threads, int *types, amdat_traj **out)
{
    AMDAT_GUARD("amdat_traj_read", amdat_traj_read_body(ctx, format, path, n_atoms, n_frames, box3, threads, types, out))
}

static int amdat_list_create_body(amdat_ctx *ctx, const amdat_traj *tr, int n, const int *indices, amdat_list **out)
{
    if (!ctx || !tr || !out || n < 0) return set_error(AMDAT_EINVAL, "amdat_list_create: bad arguments");
    auto l = std::make_unique<amdat_list>();
    l->n = n;
    l->n_atoms = tr->dev.N;
    if (indices) {
        for (int a = 0; a < n; a++)
            if (indices[a] < 0 || indices[a] >= tr->dev.N)
                return set_error(AMDAT_ERANGE, "amdat_list_create: atom index out of range");
        HIP_TRY(l->idx.upload(indices, (size_t)n, stream_of(ctx)));
    } else if (n != tr->dev.N && n != 0) {
        return set_error(AMDAT_EINVAL, "amdat_list_create: NULL indices needs n == n_atoms or n == 0");
    }
    *out = l.release();
    return AMDAT_OK;
}

int amdat_list_create(amdat_ctx *ctx, const amdat_traj *tr, int n, const int *indices, amdat_list **out)
{
    AMDAT_GUARD("amdat_list_create", amdat_list_create_body(ctx, tr, n, indices, out))
}

static int amdat_list_free_body(amdat_ctx *ctx_unused, amdat_list *l)
{
    (void)ctx_unused;
    delete l;
    return AMDAT_OK;
}

int amdat_list_free(amdat_ctx *ctx_unused, amdat_list *l)
{
    AMDAT_GUARD("amdat_list_free", amdat_list_free_body(ctx_unused, l))
}

static int amdat_list_create_timed_body(amdat_ctx *ctx, const amdat_traj *tr, int n_times, const int *time_conv,
                            const int *off, const int *indices, amdat_list **out)
{
    clear_error();
    if (!ctx || !tr || !out || n_times < 1 || !time_conv || !off)
        return set_error(AMDAT_EINVAL, "amdat_list_create_timed: bad arguments");
    if (off[0] != 0) return set_error(AMDAT_EINVAL, "amdat_list_create_timed: off[0] must be 0");
    for (int c = 0; c < n_times; c++)
        if (off[c + 1] < off[c]) return set_error(AMDAT_EINVAL, "amdat_list_create_timed: offsets decrease");
    const int n = off[n_times];
    if (n > 0 && !indices) return set_error(AMDAT_EINVAL, "amdat_list_create_timed: null indices");
    for (int a = 0; a < n; a++)
        if (indices[a] < 0 || indices[a] >= tr->dev.N)
            return set_error(AMDAT_ERANGE, "amdat_list_create_timed: atom index out of range");
    HIP_TRY(hipSetDevice(ctx->device));
    auto l = std::make_unique<amdat_list>();
    l->n = n;
    l->n_atoms = tr->dev.N;
    l->n_times = n_times;
    l->conv.assign(time_conv, time_conv + tr->dev.T);
    l->off.assign(off, off + n_times + 1);
    if (n > 0) HIP_TRY(l->idx.upload(indices, (size_t)n, stream_of(ctx)));
    *out = l.release();
    return AMDAT_OK;
}

int amdat_list_create_timed(amdat_ctx *ctx, const amdat_traj *tr, int n_times, const int *time_conv,
                            const int *off, const int *indices, amdat_list **out)
{
    AMDAT_GUARD("amdat_list_create_timed", amdat_list_create_timed_body(ctx, tr, n_times, time_conv, off, indices, out))
}

static int amdat_list_create_paired_body(amdat_ctx *ctx, const amdat_traj *tr, int n_keys, const int *t0, const int *t1,
                             const int *off, const int *indices, amdat_list **out)
{
    clear_error();
    if (!ctx || !tr || !out || n_keys < 1 || !t0 || !t1 || !off)
        return set_error(AMDAT_EINVAL, "amdat_list_create_paired: bad arguments");
    if (off[0] != 0) return set_error(AMDAT_EINVAL, "amdat_list_create_paired: off[0] must be 0");
    auto l = std::make_unique<amdat_list>();
    for (int c = 0; c < n_keys; c++) {
        if (off[c + 1] < off[c]) return set_error(AMDAT_EINVAL, "amdat_list_create_paired: offsets decrease");
        if (t0[c] < 0 || t0[c] >= tr->dev.T || t1[c] < 0 || t1[c] >= tr->dev.T)
            return set_error(AMDAT_ERANGE, "amdat_list_create_paired: frame out of range");
        if (!l->pkey.emplace(pair_key(t0[c], t1[c]), c).second)
            return set_error(AMDAT_EINVAL, "amdat_list_create_paired: frame pair given twice");
    }
    const int n = off[n_keys];
    if (n > 0 && !indices) return set_error(AMDAT_EINVAL, "amdat_list_create_paired: null indices");
    for (int a = 0; a < n; a++)
        if (indices[a] < 0 || indices[a] >= tr->dev.N)
            return set_error(AMDAT_ERANGE, "amdat_list_create_paired: atom index out of range");
    HIP_TRY(hipSetDevice(ctx->device));
    l->n = n;
    l->n_atoms = tr->dev.N;
    l->n_times = n_keys;
    l->paired = true;
    l->off.assign(off, off + n_keys + 1);
    if (n > 0) HIP_TRY(l->idx.upload(indices, (size_t)n, stream_of(ctx)));
    *out = l.release();
    return AMDAT_OK;
}

int amdat_list_create_paired(amdat_ctx *ctx, const amdat_traj *tr, int n_keys, const int *t0, const int *t1,
                             const int *off, const int *indices, amdat_list **out)
{
    AMDAT_GUARD("amdat_list_create_paired", amdat_list_create_paired_body(ctx, tr, n_keys, t0, t1, off, indices, out))
}

static int amdat_list_shape_body(const amdat_list *l, int *n_times, int *n_entries)
{
    if (!l) return set_error(AMDAT_EINVAL, "amdat_list_shape: null list");
    if (n_times) *n_times = l->n_times;
    if (n_entries) *n_entries = l->n;
    return AMDAT_OK;
}

int amdat_list_shape(const amdat_list *l, int *n_times, int *n_entries)
{
    AMDAT_GUARD("amdat_list_shape", amdat_list_shape_body(l, n_times, n_entries))
}

static int amdat_list_download_body(amdat_ctx *ctx, const amdat_list *l, int *time_conv, int *off, int *indices)
{
    clear_error();
    if (!ctx || !l) return set_error(AMDAT_EINVAL, "amdat_list_download: null argument");
    if (list_timed(l)) {
        if (time_conv && !l->paired) std::copy(l->conv.begin(), l->conv.end(), time_conv);
        if (off) std::copy(l->off.begin(), l->off.end(), off);
    }
    if (indices && l->n > 0) {
        if (l->idx.p) {
            HIP_TRY(hipSetDevice(ctx->device));
            hipStream_t st = stream_of(ctx);
            HIP_TRY(hipMemcpyAsync(indices, l->idx.p, sizeof(int) * (size_t)l->n, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
        } else {
            std::iota(indices, indices + l->n, 0);
        }
    }
    return AMDAT_OK;
}

int amdat_list_download(amdat_ctx *ctx, const amdat_list *l, int *time_conv, int *off, int *indices)
{
    AMDAT_GUARD("amdat_list_download", amdat_list_download_body(ctx, l, time_conv, off, indices))
}

static int amdat_list_select_displacement_body(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *in, int n_exponentials,
                                   int n_exponential_steps, int frt, int limit, int time_index, int mode, float lo,
                                   float hi, amdat_list **out)
{
    clear_error();
    if (!ctx || !tr || !out) return set_error(AMDAT_EINVAL, "amdat_list_select_displacement: null argument");
    const int nexp = n_exponentials, nes = n_exponential_steps, f = frt ? 1 : 0;
    if (nexp < 1 || nes < 1 || limit < 0)
        return set_error(AMDAT_EINVAL, "amdat_list_select_displacement: bad time scheme");
    if (mode < AMDAT_SELECT_FAST || mode > AMDAT_SELECT_BETWEEN)
        return set_error(AMDAT_EINVAL, "amdat_list_select_displacement: unknown mode");
    if (in && in->n_atoms != tr->dev.N)
        return set_error(AMDAT_EINVAL, "amdat_list_select_displacement: list built for another trajectory size");
    if (!frame_list_ok(in))
        return set_error(AMDAT_EINVAL, "amdat_list_select_displacement: a pair-keyed list applies to frame pairs only");
    const long T_scheme = (long)nes * nexp + 1;
    if (T_scheme > tr->dev.T)
        return set_error(AMDAT_ERANGE, "amdat_list_select_displacement: trajectory shorter than the time scheme");
    // Pairs of System::displacement_list(analysis, timegap, false), as
    // Fast_Particles::analyze calls it (src/fast_particles.cpp:117-122).
    int nq = 0;
    int rc = amdat_ts_gap_pairs(nexp, nes, f, limit, time_index, nullptr, nullptr, nullptr, 0, &nq);
    if (rc) return rc;
    std::vector<int> gtg(nq), p0(nq), p1(nq);
    rc = amdat_ts_gap_pairs(nexp, nes, f, limit, time_index, gtg.data(), p0.data(), p1.data(), nq, &nq);
    if (rc) return rc;
    // Block of each pair (src/fast_particles.cpp:133) and its input atoms
    // (the input list at the list time of t0).
    std::vector<int> blk(nq);
    std::vector<int4> seg(nq);
    int max_n = 0;
    for (int q = 0; q < nq; q++) {
        blk[q] = (int)(((float)p0[q] - (float)f) / (float)nes);
        if (blk[q] < 0 || blk[q] >= nexp)
            return set_error(AMDAT_ERANGE, "amdat_list_select_displacement: pair outside the exponential blocks");
        const int c = list_time(in, p0[q]);
        int first = -1, n = 0;
        if (!list_timed(in)) {
            n = list_n(tr, in);
            first = in && in->idx.p ? 0 : -1;
        } else if (c >= 0) {
            first = in->off[c];
            n = in->off[c + 1] - in->off[c];
        }
        seg[q] = make_int4(p0[q], p1[q], first, n);
        max_n = std::max(max_n, n);
    }
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = stream_of(ctx);
    const int n_tiles = (max_n + kBlock - 1) / kBlock;
    std::vector<int> cnt((size_t)nq * n_tiles, 0), base((size_t)nq * n_tiles, 0);
    const int *in_idx = in ? in->idx.p : nullptr;
    DevBuf<int4> d_seg;
    DevBuf<int> d_cnt, d_base;
    if (n_tiles > 0) {
        HIP_TRY(d_seg.upload(seg.data(), seg.size(), st));
        HIP_TRY(d_cnt.alloc(cnt.size()));
        HIP_TRY(launch_select(tr->dev, d_seg.p, nq, in_idx, mode, lo, hi, n_tiles, d_cnt.p, nullptr, nullptr, st));
        HIP_TRY(hipMemcpyAsync(cnt.data(), d_cnt.p, sizeof(int) * cnt.size(), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    // Output order: blocks in order, pairs of one block in pair order,
    // atoms in input order (addtrajectory appends, src/trajectory_list.cpp:
    // 509-514).
    std::vector<int> off(nexp + 1, 0);
    long total = 0;
    for (int b = 0; b < nexp; b++) {
        for (int q = 0; q < nq; q++) {
            if (blk[q] != b) continue;
            for (int t = 0; t < n_tiles; t++) {
                base[(size_t)q * n_tiles + t] = (int)total;
                total += cnt[(size_t)q * n_tiles + t];
            }
        }
        off[b + 1] = (int)total;
    }
    auto l = std::make_unique<amdat_list>();
    l->n = (int)total;
    l->n_atoms = tr->dev.N;
    l->n_times = nexp;
    l->off = off;
    l->conv.resize(tr->dev.T);
    // Exptime_Trajectory_List time conversion (src/exptime_trajectory_list.cpp:87-91)
    for (int t = 0; t < tr->dev.T; t++) l->conv[t] = (int)((float)(t - f) / (float)nes);
    if (total > 0) {
        HIP_TRY(l->idx.alloc((size_t)total));
        HIP_TRY(d_base.upload(base.data(), base.size(), st));
        HIP_TRY(launch_select(tr->dev, d_seg.p, nq, in_idx, mode, lo, hi, n_tiles, nullptr, d_base.p, l->idx.p, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    *out = l.release();
    return AMDAT_OK;
}

int amdat_list_select_displacement(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *in, int n_exponentials,
                                   int n_exponential_steps, int frt, int limit, int time_index, int mode, float lo,
                                   float hi, amdat_list **out)
{
    AMDAT_GUARD("amdat_list_select_displacement", amdat_list_select_displacement_body(ctx, tr, in, n_exponentials, n_exponential_steps, frt, limit, time_index, mode, lo, hi, out))
}

static int amdat_pairs_upload_body(amdat_ctx *ctx, int n_pairs, const int *tg, const int *t0, const int *t1, int n_timegaps,
                       amdat_pairs **out)
{
    if (!ctx || !out || n_pairs < 0 || n_timegaps < 1 || (n_pairs > 0 && (!tg || !t0 || !t1)))
        return set_error(AMDAT_EINVAL, "amdat_pairs_upload: bad arguments");
    auto p = std::make_unique<amdat_pairs>();
    p->n = n_pairs;
    p->ntg = n_timegaps;
    p->tg.assign(tg, tg + n_pairs);
    p->t0.assign(t0, t0 + n_pairs);
    p->t1.assign(t1, t1 + n_pairs);
    for (int i = 0; i < n_pairs; i++) {
        if (tg[i] < 0 || tg[i] >= n_timegaps) return set_error(AMDAT_ERANGE, "amdat_pairs_upload: time gap out of range");
        if (t0[i] < 0 || t1[i] < 0) return set_error(AMDAT_ERANGE, "amdat_pairs_upload: negative frame index");
    }
    p->group_off.assign(n_timegaps + 1, 0);
    for (int i = 0; i < n_pairs; i++) p->group_off[tg[i] + 1]++;
    for (int g = 0; g < n_timegaps; g++) p->group_off[g + 1] += p->group_off[g];
    p->order.resize(n_pairs);
    std::vector<int> fill(p->group_off.begin(), p->group_off.end() - 1);
    for (int i = 0; i < n_pairs; i++) p->order[fill[tg[i]]++] = i;
    hipStream_t st = stream_of(ctx);
    HIP_TRY(p->d_t0.upload(t0, n_pairs, st));
    HIP_TRY(p->d_t1.upload(t1, n_pairs, st));
    HIP_TRY(p->d_group_off.upload(p->group_off.data(), p->group_off.size(), st));
    HIP_TRY(p->d_order.upload(p->order.data(), p->order.size(), st));
    *out = p.release();
    return AMDAT_OK;
}

int amdat_pairs_upload(amdat_ctx *ctx, int n_pairs, const int *tg, const int *t0, const int *t1, int n_timegaps,
                       amdat_pairs **out)
{
    AMDAT_GUARD("amdat_pairs_upload", amdat_pairs_upload_body(ctx, n_pairs, tg, t0, t1, n_timegaps, out))
}

static int amdat_pairs_free_body(amdat_ctx *ctx_unused, amdat_pairs *p)
{
    (void)ctx_unused;
    delete p;
    return AMDAT_OK;
}

int amdat_pairs_free(amdat_ctx *ctx_unused, amdat_pairs *p)
{
    AMDAT_GUARD("amdat_pairs_free", amdat_pairs_free_body(ctx_unused, p))
}

static int check_pairs_frames(const amdat_pairs *p, int T)
{
    for (int i = 0; i < p->n; i++)
        if (p->t0[i] >= T || p->t1[i] >= T) return set_error(AMDAT_ERANGE, "pair frame index beyond the trajectory");
    return AMDAT_OK;
}

static int amdat_rho_compute_body(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list, const amdat_shells *sh_c,
                      int first_shell, int last_shell, float length_scale, const int *frames, int n_frames,
                      int precision, amdat_rho **out)
{
    clear_error();
    if (!ctx || !tr || !sh_c || !out) return set_error(AMDAT_EINVAL, "amdat_rho_compute: null argument");
    amdat_shells *sh = const_cast<amdat_shells *>(sh_c);
    if (first_shell < 0 || last_shell < first_shell || last_shell >= sh->n_shells)
        return set_error(AMDAT_ERANGE, "amdat_rho_compute: shell range outside the table");
    if (!(length_scale > 0.0f)) return set_error(AMDAT_EINVAL, "amdat_rho_compute: length scale must be > 0");
    if (!frame_list_ok(list))
        return set_error(AMDAT_EINVAL, "amdat_rho_compute: a pair-keyed list applies to frame pairs only");
    if (int rc0 = check_list_traj(list, tr, "amdat_rho_compute")) return rc0;
    if (precision != AMDAT_PREC_FP32C && precision != AMDAT_PREC_FP64 && precision != AMDAT_PREC_FP32X &&
        precision != AMDAT_PREC_FP64T)
        return set_error(AMDAT_EINVAL, "amdat_rho_compute: unknown precision");
    HIP_TRY(hipSetDevice(ctx->device));
    const int T = tr->dev.T;
    std::vector<int> fr;
    if (frames) {
        if (n_frames < 1) return set_error(AMDAT_EINVAL, "amdat_rho_compute: empty frame list");
        fr.assign(frames, frames + n_frames);
        for (int f : fr)
            if (f < 0 || f >= T) return set_error(AMDAT_ERANGE, "amdat_rho_compute: frame out of range");
    } else {
        fr.resize(T);
        std::iota(fr.begin(), fr.end(), 0);
    }
    KPlanOwned *plan = nullptr;
    const PlanCost pc = plan_cost(precision, false);
    int rc = build_plan(ctx, sh, first_shell, last_shell, &plan, 0, pc.c, pc.grouped);
    if (rc) return rc;
    const KPlan &kp = plan->kp;
    const int nl = list_n(tr, list);
    auto r = std::make_unique<amdat_rho>();
    r->n_frames = (int)fr.size();
    r->K = kp.K;
    r->n_list = nl;
    r->T = T;
    r->first_shell = first_shell;
    r->n_shells = last_shell - first_shell + 1;
    r->frames = fr;
    r->frame_row.assign(T, -1);
    for (int i = 0; i < r->n_frames; i++) r->frame_row[fr[i]] = i;
    r->row_n.resize(r->n_frames);
    for (int i = 0; i < r->n_frames; i++) r->row_n[i] = list_at(tr, list, fr[i]).n;
    r->shell_off.resize(r->n_shells + 1);
    for (int s = 0; s <= r->n_shells; s++) r->shell_off[s] = sh->off[first_shell + s] - sh->off[first_shell];
    hipStream_t st = stream_of(ctx);
    HIP_TRY(r->d_shell_off.upload(r->shell_off.data(), r->shell_off.size(), st));
    HIP_TRY(r->d.alloc((size_t)r->n_frames * kp.K));
    if (kp.K > 0) {
        DevBuf<int> dfr;
        HIP_TRY(dfr.upload(fr.data(), fr.size(), st));
        const double inv_ell = 1.0 / (double)length_scale;
        // Consecutive rows with one list time share a launch (a static list
        // is one run); launches hold at most fb frames (grid limit).
        const int fb = 65535;
        struct Run {
            int row, len;
            ListSeg seg;
            int splits;
        };
        std::vector<Run> runs;
        for (int i = 0; i < r->n_frames;) {
            const int c = list_time(list, fr[i]);
            int j = i + 1;
            while (j < r->n_frames && list_time(list, fr[j]) == c) j++;
            const ListSeg sg = list_seg(tr, list, c);
            runs.push_back({i, j - i, sg, sg.n > 0 ? choose_splits((long)kp.n_chunks * (j - i), sg.n) : 1});
            i = j;
        }
        size_t part_n = 0;
        for (const Run &u : runs)
            if (u.seg.n > 0 && u.splits > 1)
                part_n = std::max(part_n, (size_t)std::min(fb, u.len) * u.splits * kp.K);
        DevBuf<double2> part;
        if (part_n) HIP_TRY(part.alloc(part_n));
        for (const Run &u : runs) {
            double2 *rows = r->d.p + (size_t)u.row * kp.K;
            if (u.seg.n == 0) {
                HIP_TRY(hipMemsetAsync(rows, 0, sizeof(double2) * (size_t)u.len * kp.K, st));
                continue;
            }
            for (int f0 = 0; f0 < u.len; f0 += fb) {
                const int nf = std::min(fb, u.len - f0);
                double2 *dst = rows + (size_t)f0 * kp.K;
                HIP_TRY(launch_rho(precision, tr->dev, u.seg.p, u.seg.n, dfr.p + u.row + f0, nf, kp, inv_ell,
                                   u.splits, u.splits > 1 ? part.p : dst, st));
                if (u.splits > 1) HIP_TRY(launch_split_reduce2(part.p, nf, u.splits, kp.K, dst, st));
            }
        }
        HIP_TRY(hipStreamSynchronize(st));
    }
    *out = r.release();
    return AMDAT_OK;
}

int amdat_rho_compute(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list, const amdat_shells *sh_c,
                      int first_shell, int last_shell, float length_scale, const int *frames, int n_frames,
                      int precision, amdat_rho **out)
{
    AMDAT_GUARD("amdat_rho_compute", amdat_rho_compute_body(ctx, tr, list, sh_c, first_shell, last_shell, length_scale, frames, n_frames, precision, out))
}

static int amdat_rho_download_body(amdat_ctx *ctx, const amdat_rho *rho, double *out, size_t n_doubles)
{
    if (!ctx || !rho || !out) return set_error(AMDAT_EINVAL, "amdat_rho_download: null argument");
    const size_t need = (size_t)rho->n_frames * rho->K * 2;
    if (n_doubles < need) return set_error(AMDAT_EINVAL, "amdat_rho_download: output too small");
    HIP_TRY(hipMemcpyAsync(out, rho->d.p, need * sizeof(double), hipMemcpyDeviceToHost, stream_of(ctx)));
    HIP_TRY(hipStreamSynchronize(stream_of(ctx)));
    return AMDAT_OK;
}

int amdat_rho_download(amdat_ctx *ctx, const amdat_rho *rho, double *out, size_t n_doubles)
{
    AMDAT_GUARD("amdat_rho_download", amdat_rho_download_body(ctx, rho, out, n_doubles))
}

static int amdat_rho_shape_body(const amdat_rho *rho, int *n_frames, int *n_vectors)
{
    if (!rho) return set_error(AMDAT_EINVAL, "amdat_rho_shape: null argument");
    if (n_frames) *n_frames = rho->n_frames;
    if (n_vectors) *n_vectors = rho->K;
    return AMDAT_OK;
}

int amdat_rho_shape(const amdat_rho *rho, int *n_frames, int *n_vectors)
{
    AMDAT_GUARD("amdat_rho_shape", amdat_rho_shape_body(rho, n_frames, n_vectors))
}

static int amdat_rho_free_body(amdat_ctx *ctx_unused, amdat_rho *rho)
{
    (void)ctx_unused;
    delete rho;
    return AMDAT_OK;
}

int amdat_rho_free(amdat_ctx *ctx_unused, amdat_rho *rho)
{
    AMDAT_GUARD("amdat_rho_free", amdat_rho_free_body(ctx_unused, rho))
}

static int amdat_sk_body(amdat_ctx *ctx, const amdat_rho *r1, const amdat_rho *r2, double *S)
{
    clear_error();
    if (!ctx || !r1 || !r2 || !S) return set_error(AMDAT_EINVAL, "amdat_sk: null argument");
    if (r1->K != r2->K || r1->n_frames != r2->n_frames || r1->frames != r2->frames || r1->shell_off != r2->shell_off)
        return set_error(AMDAT_EINVAL, "amdat_sk: densities cover different frames or vectors");
    hipStream_t st = stream_of(ctx);
    const int nf = r1->n_frames;
    std::vector<int> rows(nf), goff = {0, nf};
    std::iota(rows.begin(), rows.end(), 0);
    double norm = 0.0;
    for (int i = 0; i < nf; i++) norm += (double)r1->row_n[i];
    if (norm == 0.0 || r1->K == 0) {
        // no atoms over the frames: S = 0 (src/structure_factor.cpp:275-285)
        std::fill(S, S + r1->n_shells, 0.0);
        return AMDAT_OK;
    }
    DevBuf<int> d_rows, d_goff;
    DevBuf<double> d_norm, d_out;
    HIP_TRY(d_rows.upload(rows.data(), rows.size(), st));
    HIP_TRY(d_goff.upload(goff.data(), goff.size(), st));
    HIP_TRY(d_norm.upload(&norm, 1, st));
    HIP_TRY(d_out.alloc(r1->n_shells));
    HIP_TRY(launch_corr_shell(r1->d.p, r2->d.p, r1->K, d_goff.p, d_rows.p, d_rows.p, 1, r1->d_shell_off.p,
                              r1->d_shell_off.p + 1,
                              r1->n_shells, d_norm.p, d_out.p, st));
    HIP_TRY(hipMemcpyAsync(S, d_out.p, sizeof(double) * r1->n_shells, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return AMDAT_OK;
}

int amdat_sk(amdat_ctx *ctx, const amdat_rho *r1, const amdat_rho *r2, double *S)
{
    AMDAT_GUARD("amdat_sk", amdat_sk_body(ctx, r1, r2, S))
}

}  // extern "C"

namespace {
// F(k,t) over columns of vectors: column c sums vectors vbeg[c] ..
// vend[c]-1 of the densities (device arrays; null = the densities' shells
// col0 .. col0+n_cols-1).  Pairs in reference order within each time gap;
// normalised by the atoms of the first density at each pair's t0.
int fkt_columns(amdat_ctx *ctx, const amdat_rho *r1, const amdat_rho *r2, const amdat_pairs *pairs, int n_cols,
                const int *vbeg, const int *vend, int col0, double *F, double *count, const char *fn)
{
    if (r1->K != r2->K || r1->shell_off != r2->shell_off || r1->T != r2->T)
        return set_error(AMDAT_EINVAL, std::string(fn) + ": densities cover different vectors");
    int rc = check_pairs_frames(pairs, r1->T);
    if (rc) return rc;
    hipStream_t st = stream_of(ctx);
    const int ntg = pairs->ntg;
    std::vector<int> row0(pairs->n), row1(pairs->n);
    std::vector<double> norm(ntg, 0.0);
    for (int q = 0; q < pairs->n; q++) {
        const int i = pairs->order[q];
        row0[q] = r1->frame_row[pairs->t0[i]];
        row1[q] = r2->frame_row[pairs->t1[i]];
        if (row0[q] < 0 || row1[q] < 0)
            return set_error(AMDAT_ERANGE, std::string(fn) + ": a pair uses a frame the density lacks");
        norm[pairs->tg[i]] += (double)r1->row_n[row0[q]];
    }
    for (int g = 0; g < ntg; g++) count[g] = norm[g];
    if (pairs->n == 0 || n_cols == 0) {
        std::fill(F, F + (size_t)ntg * n_cols, 0.0);
        return AMDAT_OK;
    }
    DevBuf<int> d_r0, d_r1;
    DevBuf<double> d_norm, d_out;
    HIP_TRY(d_r0.upload(row0.data(), row0.size(), st));
    HIP_TRY(d_r1.upload(row1.data(), row1.size(), st));
    HIP_TRY(d_norm.upload(norm.data(), norm.size(), st));
    HIP_TRY(d_out.alloc((size_t)ntg * n_cols));
    const int *b = vbeg ? vbeg : r1->d_shell_off.p + col0;
    const int *e = vbeg ? vend : r1->d_shell_off.p + col0 + 1;
    HIP_TRY(launch_corr_shell(r1->d.p, r2->d.p, r1->K, pairs->d_group_off.p, d_r0.p, d_r1.p, ntg, b, e, n_cols,
                              d_norm.p, d_out.p, st));
    HIP_TRY(hipMemcpyAsync(F, d_out.p, sizeof(double) * ntg * n_cols, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return AMDAT_OK;
}
}  // namespace

extern "C" {

static int amdat_fkt_body(amdat_ctx *ctx, const amdat_rho *r1, const amdat_rho *r2, const amdat_pairs *pairs, int wn1, int wn2,
              double *F, double *count)
{
    clear_error();
    if (!ctx || !r1 || !r2 || !pairs || !F || !count) return set_error(AMDAT_EINVAL, "amdat_fkt: null argument");
    if (wn1 < 0 || wn2 < wn1 || wn2 >= r1->n_shells)
        return set_error(AMDAT_ERANGE, "amdat_fkt: shell range outside the densities");
    HIP_TRY(hipSetDevice(ctx->device));
    return fkt_columns(ctx, r1, r2, pairs, wn2 - wn1 + 1, nullptr, nullptr, wn1, F, count, "amdat_fkt");
}

int amdat_fkt(amdat_ctx *ctx, const amdat_rho *r1, const amdat_rho *r2, const amdat_pairs *pairs, int wn1, int wn2,
              double *F, double *count)
{
    AMDAT_GUARD("amdat_fkt", amdat_fkt_body(ctx, r1, r2, pairs, wn1, wn2, F, count))
}

static int amdat_fkt_vectors_body(amdat_ctx *ctx, const amdat_rho *r1, const amdat_rho *r2, const amdat_pairs *pairs,
                      int n_cols, const int *v_first, const int *v_count, double *F, double *count)
{
    clear_error();
    if (!ctx || !r1 || !r2 || !pairs || !F || !count || n_cols < 0 || (n_cols > 0 && (!v_first || !v_count)))
        return set_error(AMDAT_EINVAL, "amdat_fkt_vectors: null argument");
    std::vector<int> b(n_cols), e(n_cols);
    for (int c = 0; c < n_cols; c++) {
        if (v_first[c] < 0 || v_count[c] < 0 || (long)v_first[c] + v_count[c] > r1->K)
            return set_error(AMDAT_ERANGE, "amdat_fkt_vectors: vector range outside the densities");
        b[c] = v_first[c];
        e[c] = v_first[c] + v_count[c];
    }
    HIP_TRY(hipSetDevice(ctx->device));
    DevBuf<int> d_b, d_e;
    hipStream_t st = stream_of(ctx);
    HIP_TRY(d_b.upload(b.data(), b.size(), st));
    HIP_TRY(d_e.upload(e.data(), e.size(), st));
    return fkt_columns(ctx, r1, r2, pairs, n_cols, d_b.p, d_e.p, 0, F, count, "amdat_fkt_vectors");
}

int amdat_fkt_vectors(amdat_ctx *ctx, const amdat_rho *r1, const amdat_rho *r2, const amdat_pairs *pairs,
                      int n_cols, const int *v_first, const int *v_count, double *F, double *count)
{
    AMDAT_GUARD("amdat_fkt_vectors", amdat_fkt_vectors_body(ctx, r1, r2, pairs, n_cols, v_first, v_count, F, count))
}

static int amdat_rho_sum_body(amdat_ctx *ctx, const amdat_rho *a, const amdat_rho *b, amdat_rho **out)
{
    clear_error();
    if (!ctx || !a || !b || !out) return set_error(AMDAT_EINVAL, "amdat_rho_sum: null argument");
    if (a->K != b->K || a->frames != b->frames || a->shell_off != b->shell_off || a->T != b->T ||
        a->first_shell != b->first_shell)
        return set_error(AMDAT_EINVAL, "amdat_rho_sum: densities cover different frames or vectors");
    HIP_TRY(hipSetDevice(ctx->device));
    auto r = std::make_unique<amdat_rho>();
    r->n_frames = a->n_frames;
    r->K = a->K;
    r->n_list = a->n_list + b->n_list;
    r->T = a->T;
    r->first_shell = a->first_shell;
    r->n_shells = a->n_shells;
    r->frames = a->frames;
    r->frame_row = a->frame_row;
    r->shell_off = a->shell_off;
    r->row_n.resize(a->row_n.size());
    for (size_t i = 0; i < r->row_n.size(); i++) r->row_n[i] = a->row_n[i] + b->row_n[i];
    hipStream_t st = stream_of(ctx);
    HIP_TRY(r->d_shell_off.upload(r->shell_off.data(), r->shell_off.size(), st));
    const size_t n = (size_t)r->n_frames * r->K;
    HIP_TRY(r->d.alloc(n));
    HIP_TRY(launch_rho_add(a->d.p, b->d.p, n, r->d.p, st));
    HIP_TRY(hipStreamSynchronize(st));
    *out = r.release();
    return AMDAT_OK;
}

int amdat_rho_sum(amdat_ctx *ctx, const amdat_rho *a, const amdat_rho *b, amdat_rho **out)
{
    AMDAT_GUARD("amdat_rho_sum", amdat_rho_sum_body(ctx, a, b, out))
}

static int amdat_fs_body(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list, const amdat_shells *sh_c, int first_shell,
             int last_shell, float length_scale, const amdat_pairs *pairs, int precision, double *Fs, double *count)
{
    clear_error();
    if (!ctx || !tr || !sh_c || !pairs || !Fs || !count) return set_error(AMDAT_EINVAL, "amdat_fs: null argument");
    amdat_shells *sh = const_cast<amdat_shells *>(sh_c);
    if (first_shell < 0 || last_shell < first_shell || last_shell >= sh->n_shells)
        return set_error(AMDAT_ERANGE, "amdat_fs: shell range outside the table");
    if (!(length_scale > 0.0f)) return set_error(AMDAT_EINVAL, "amdat_fs: length scale must be > 0");
    if (precision != AMDAT_PREC_FP32C && precision != AMDAT_PREC_FP64 && precision != AMDAT_PREC_FP64T)
        return set_error(AMDAT_EINVAL, "amdat_fs: unknown precision");
    int rc = check_list_traj(list, tr, "amdat_fs");
    if (rc) return rc;
    rc = check_pairs_frames(pairs, tr->dev.T);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    KPlanOwned *plan = nullptr;
    const PlanCost pc = plan_cost(precision, true);
    rc = build_plan(ctx, sh, first_shell, last_shell, &plan, 0, pc.c, pc.grouped);
    if (rc) return rc;
    const KPlan &kp = plan->kp;
    const int ntg = pairs->ntg, nsh = last_shell - first_shell + 1;
    std::vector<double> norm(ntg, 0.0);
    PairLaunch pl;
    rc = plan_pair_launch(ctx, tr, list, pairs, false, pl, norm.data());
    if (rc) return rc;
    for (int g = 0; g < ntg; g++) count[g] = norm[g];
    if (pairs->n == 0 || pl.max_n == 0 || kp.K == 0) {
        std::fill(Fs, Fs + (size_t)ntg * nsh, 0.0);
        return AMDAT_OK;
    }
    hipStream_t st = stream_of(ctx);
    const int n_splits = choose_splits((long)kp.n_chunks * pairs->n, pl.max_n);
    DevBuf<double> V, d_norm, d_out;
    DevBuf<int> d_soff;
    std::vector<int> soff(nsh + 1);
    for (int s = 0; s <= nsh; s++) soff[s] = sh->off[first_shell + s] - sh->off[first_shell];
    HIP_TRY(d_soff.upload(soff.data(), soff.size(), st));
    HIP_TRY(d_norm.upload(norm.data(), norm.size(), st));
    HIP_TRY(V.alloc((size_t)pairs->n * n_splits * kp.K));
    if (pl.any_empty) HIP_TRY(hipMemsetAsync(V.p, 0, V.n * sizeof(double), st));
    const double inv_ell = 1.0 / (double)length_scale;
    const int pb = 65535;
    for (const PairLaunch::Group &gr : pl.groups) {
        if (gr.seg.n == 0) continue;
        for (int p0 = gr.row; p0 < gr.row + gr.len; p0 += pb) {
            const int np = std::min(pb, gr.row + gr.len - p0);
            HIP_TRY(launch_fs(precision, tr->dev, gr.seg.p, gr.seg.n, pl.pseg ? pl.pseg + p0 : nullptr, pl.t0 + p0,
                              pl.t1 + p0, np, kp, inv_ell,
                              n_splits, V.p + (size_t)p0 * n_splits * kp.K, st));
        }
    }
    HIP_TRY(d_out.alloc((size_t)ntg * nsh));
    HIP_TRY(launch_fs_shell(V.p, n_splits, kp.K, pairs->d_group_off.p, pl.order, ntg, d_soff.p, nsh, d_norm.p,
                            d_out.p, st));
    HIP_TRY(hipMemcpyAsync(Fs, d_out.p, sizeof(double) * ntg * nsh, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return AMDAT_OK;
}

int amdat_fs(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list, const amdat_shells *sh_c, int first_shell,
             int last_shell, float length_scale, const amdat_pairs *pairs, int precision, double *Fs, double *count)
{
    AMDAT_GUARD("amdat_fs", amdat_fs_body(ctx, tr, list, sh_c, first_shell, last_shell, length_scale, pairs, precision, Fs, count))
}

static int amdat_msd_ngp_body(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list, const amdat_pairs *pairs, double *s2,
                  double *s4, double *count)
{
    clear_error();
    if (!ctx || !tr || !pairs || !s2 || !s4 || !count) return set_error(AMDAT_EINVAL, "amdat_msd_ngp: null argument");
    int rc = check_list_traj(list, tr, "amdat_msd_ngp");
    if (rc) return rc;
    rc = check_pairs_frames(pairs, tr->dev.T);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    const int ntg = pairs->ntg;
    PairLaunch pl;
    rc = plan_pair_launch(ctx, tr, list, pairs, false, pl, count);
    if (rc) return rc;
    if (pairs->n == 0 || pl.max_n == 0) {
        std::fill(s2, s2 + ntg, 0.0);
        std::fill(s4, s4 + ntg, 0.0);
        return AMDAT_OK;
    }
    hipStream_t st = stream_of(ctx);
    const int nb = msd_blocks(pl.max_n);
    DevBuf<double> part, d2, d4;
    HIP_TRY(part.alloc((size_t)pairs->n * nb * 2));
    if (pl.any_empty) HIP_TRY(hipMemsetAsync(part.p, 0, part.n * sizeof(double), st));
    const int pb = 65535;
    for (const PairLaunch::Group &gr : pl.groups) {
        if (gr.seg.n == 0) continue;
        for (int p0 = gr.row; p0 < gr.row + gr.len; p0 += pb) {
            const int np = std::min(pb, gr.row + gr.len - p0);
            HIP_TRY(launch_msd(tr->dev, gr.seg.p, gr.seg.n, pl.pseg ? pl.pseg + p0 : nullptr, pl.t0 + p0, pl.t1 + p0,
                               np, nb,
                               part.p + (size_t)p0 * nb * 2, st));
        }
    }
    HIP_TRY(d2.alloc(ntg));
    HIP_TRY(d4.alloc(ntg));
    HIP_TRY(launch_msd_reduce(part.p, nb, pairs->d_group_off.p, pl.order, ntg, d2.p, d4.p, st));
    HIP_TRY(hipMemcpyAsync(s2, d2.p, sizeof(double) * ntg, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s4, d4.p, sizeof(double) * ntg, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return AMDAT_OK;
}

int amdat_msd_ngp(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list, const amdat_pairs *pairs, double *s2,
                  double *s4, double *count)
{
    AMDAT_GUARD("amdat_msd_ngp", amdat_msd_ngp_body(ctx, tr, list, pairs, s2, s4, count))
}

static int amdat_msd_ngp_noalloc_body(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list, const amdat_pairs *pairs,
                          double *s2, double *s4, double *count)
{
    clear_error();
    if (!ctx || !tr || !pairs || !s2 || !s4 || !count)
        return set_error(AMDAT_EINVAL, "amdat_msd_ngp_noalloc: null argument");
    if (list && (list->n_times > 0 || list->paired))
        return set_error(AMDAT_EINVAL, "amdat_msd_ngp_noalloc: static lists only (a time-varying or pair-keyed "
                                       "list needs per-call tables: use amdat_msd_ngp)");
    int rc = check_list_traj(list, tr, "amdat_msd_ngp_noalloc");
    if (rc) return rc;
    rc = check_pairs_frames(pairs, tr->dev.T);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    const int ntg = pairs->ntg, n = list_n(tr, list), np = pairs->n;
    for (int g = 0; g < ntg; g++) count[g] = 0.0;
    for (int i = 0; i < np; i++) count[pairs->tg[i]] += (double)n;
    std::fill(s2, s2 + ntg, 0.0);
    std::fill(s4, s4 + ntg, 0.0);
    if (np == 0 || n == 0) return AMDAT_OK;
    // the same atom blocks as amdat_msd_ngp, so the sums are the same bits
    const int nb = msd_blocks(n);
    const size_t n_part = (size_t)np * nb * 2;
    if ((n_part + 2 * (size_t)ntg) * sizeof(double) > kNoAllocScratchBytes)
        return set_error(AMDAT_ERANGE, "amdat_msd_ngp_noalloc: " + std::to_string(np) + " pairs x " +
                                           std::to_string(nb) + " atom blocks need more than the " +
                                           std::to_string(kNoAllocScratchBytes >> 20) + " MiB scratch");
    if (!ctx->scratch) {
        rc = scratch_address(ctx->device, &ctx->scratch);
        if (rc) return rc;
    }
    double *part = static_cast<double *>(ctx->scratch);
    double *d2 = part + n_part, *d4 = d2 + ntg;
    hipStream_t st = stream_of(ctx);
    // every (pair, block) partial is written by launch_msd before the
    // reduction reads it: nothing needs clearing
    const int pb = 65535;
    for (int p0 = 0; p0 < np; p0 += pb) {
        const int m = std::min(pb, np - p0);
        HIP_TRY(launch_msd(tr->dev, list_ptr(list), n, nullptr, pairs->d_t0.p + p0, pairs->d_t1.p + p0, m, nb,
                           part + (size_t)p0 * nb * 2, st));
    }
    HIP_TRY(launch_msd_reduce(part, nb, pairs->d_group_off.p, pairs->d_order.p, ntg, d2, d4, st));
    HIP_TRY(hipMemcpyAsync(s2, d2, sizeof(double) * ntg, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s4, d4, sizeof(double) * ntg, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return AMDAT_OK;
}

int amdat_msd_ngp_noalloc(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list, const amdat_pairs *pairs,
                          double *s2, double *s4, double *count)
{
    AMDAT_GUARD("amdat_msd_ngp_noalloc", amdat_msd_ngp_noalloc_body(ctx, tr, list, pairs, s2, s4, count))
}

}  // extern "C"

namespace {

// Displacement sums per time gap and, with n_bins > 0, a histogram per time
// gap (hist_mode 0: |dr| below n_bins; 1: |dr|^2 with overflow bin).
int displacements_impl(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list, const amdat_pairs *pairs,
                       int n_bins, int hist_mode, float bin_size, double *sum_d, double *sum_d2,
                       unsigned long long *hist, double *count)
{
    int rc = check_list_traj(list, tr, "amdat_displacements");
    if (rc) return rc;
    rc = check_pairs_frames(pairs, tr->dev.T);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    const int ntg = pairs->ntg;
    std::fill(sum_d, sum_d + 3 * (size_t)ntg, 0.0);
    std::fill(sum_d2, sum_d2 + 3 * (size_t)ntg, 0.0);
    if (n_bins > 0) std::fill(hist, hist + (size_t)ntg * n_bins, 0ull);
    PairLaunch pl;
    rc = plan_pair_launch(ctx, tr, list, pairs, true, pl, count);
    if (rc) return rc;
    if (pairs->n == 0 || pl.max_n == 0) return AMDAT_OK;
    hipStream_t st = stream_of(ctx);
    const int nb = std::max(1, std::min(64, (pl.max_n + kBlock * 4 - 1) / (kBlock * 4)));
    DevBuf<double> part, out;
    DevBuf<unsigned long long> dh;
    HIP_TRY(part.alloc((size_t)pairs->n * nb * 6));
    if (pl.any_empty) HIP_TRY(hipMemsetAsync(part.p, 0, part.n * sizeof(double), st));
    if (n_bins > 0) {
        HIP_TRY(dh.alloc((size_t)ntg * n_bins));
        HIP_TRY(hipMemsetAsync(dh.p, 0, dh.n * sizeof(unsigned long long), st));
    }
    const int pb = 65535;
    for (const PairLaunch::Group &gr : pl.groups) {
        if (gr.seg.n == 0) continue;
        for (int p0 = gr.row; p0 < gr.row + gr.len; p0 += pb) {
            const int np = std::min(pb, gr.row + gr.len - p0);
            HIP_TRY(launch_disp(tr->dev, gr.seg.p, gr.seg.n, pl.pseg ? pl.pseg + p0 : nullptr, pl.t0 + p0, pl.t1 + p0,
                                pl.tg + p0, np, nb, n_bins,
                                hist_mode, bin_size, part.p + (size_t)p0 * nb * 6, dh.p, st));
        }
    }
    HIP_TRY(out.alloc((size_t)ntg * 6));
    HIP_TRY(launch_disp_reduce(part.p, nb, pairs->d_group_off.p, pl.order, ntg, out.p, st));
    std::vector<double> h((size_t)ntg * 6);
    HIP_TRY(hipMemcpyAsync(h.data(), out.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    if (n_bins > 0)
        HIP_TRY(hipMemcpyAsync(hist, dh.p, sizeof(unsigned long long) * dh.n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int g = 0; g < ntg; g++)
        for (int c = 0; c < 3; c++) {
            sum_d[3 * g + c] = h[6 * (size_t)g + c];
            sum_d2[3 * g + c] = h[6 * (size_t)g + 3 + c];
        }
    return AMDAT_OK;
}

}  // namespace

extern "C" {

static int amdat_displacements_body(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list, const amdat_pairs *pairs,
                        int n_bins, float max_range, double *sum_d, double *sum_d2, unsigned long long *hist,
                        double *count)
{
    clear_error();
    if (!ctx || !tr || !pairs || !sum_d || !sum_d2 || !count)
        return set_error(AMDAT_EINVAL, "amdat_displacements: null argument");
    if (n_bins < 0 || n_bins > 16384 || (n_bins > 0 && (!hist || !(max_range > 0.0f))))
        return set_error(AMDAT_EINVAL, "amdat_displacements: bins must be 0..16384 with a histogram and max_range > 0");
    // bin size of Space_Time_Correlation_Function (src/van_hove_self.cpp:56)
    const float bin_size = n_bins > 0 ? max_range / (float)n_bins : 1.0f;
    return displacements_impl(ctx, tr, list, pairs, n_bins, 0, bin_size, sum_d, sum_d2, hist, count);
}

int amdat_displacements(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list, const amdat_pairs *pairs,
                        int n_bins, float max_range, double *sum_d, double *sum_d2, unsigned long long *hist,
                        double *count)
{
    AMDAT_GUARD("amdat_displacements", amdat_displacements_body(ctx, tr, list, pairs, n_bins, max_range, sum_d, sum_d2, hist, count))
}

static int amdat_displacement_dist_body(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list, const amdat_pairs *pairs,
                            int n_bins, float max_value, unsigned long long *counts, double *sum_d2,
                            double *sum_d4, double *count)
{
    clear_error();
    if (!ctx || !tr || !pairs || !counts || !sum_d2 || !sum_d4 || !count)
        return set_error(AMDAT_EINVAL, "amdat_displacement_dist: null argument");
    if (n_bins < 1 || n_bins > 16383 || !(max_value > 0.0f))
        return set_error(AMDAT_EINVAL, "amdat_displacement_dist: bins must be 1..16383 and max_value > 0");
    const int ntg = pairs->ntg, nb = n_bins + 1;
    // binsize = maxvalue / bins (src/debyewaller_dist.cpp:20)
    const float bin_size = max_value / (float)n_bins;
    std::vector<double> sd(3 * (size_t)ntg), sd2(3 * (size_t)ntg), cnt(ntg), s2(ntg), s4(ntg), c2(ntg);
    std::vector<unsigned long long> h((size_t)ntg * nb);
    int rc = displacements_impl(ctx, tr, list, pairs, nb, 1, bin_size, sd.data(), sd2.data(), h.data(), cnt.data());
    if (rc) return rc;
    rc = amdat_msd_ngp(ctx, tr, list, pairs, s2.data(), s4.data(), c2.data());
    if (rc) return rc;
    std::fill(counts, counts + nb, 0ull);
    *sum_d2 = *sum_d4 = *count = 0.0;
    for (int g = 0; g < ntg; g++) {
        for (int b = 0; b < nb; b++) counts[b] += h[(size_t)g * nb + b];
        *sum_d2 += s2[g];
        *sum_d4 += s4[g];
        *count += cnt[g];
    }
    return AMDAT_OK;
}

int amdat_displacement_dist(amdat_ctx *ctx, const amdat_traj *tr, const amdat_list *list, const amdat_pairs *pairs,
                            int n_bins, float max_value, unsigned long long *counts, double *sum_d2,
                            double *sum_d4, double *count)
{
    AMDAT_GUARD("amdat_displacement_dist", amdat_displacement_dist_body(ctx, tr, list, pairs, n_bins, max_value, counts, sum_d2, sum_d4, count))
}

static int amdat_sk_loss_grad_body(amdat_ctx *ctx, int n_atoms, const float *d_x, const float *d_y, const float *d_z,
                       const amdat_shells *sh_c, int first_shell, int last_shell, float length_scale,
                       const double *weights, const double *target, double *loss, double *S, double *d_gx,
                       double *d_gy, double *d_gz)
{
    clear_error();
    if (!ctx || !sh_c || !d_x || !d_y || !d_z || !weights || !target || !loss || n_atoms < 1)
        return set_error(AMDAT_EINVAL, "amdat_sk_loss_grad: bad arguments");
    if ((d_gx || d_gy || d_gz) && !(d_gx && d_gy && d_gz))
        return set_error(AMDAT_EINVAL, "amdat_sk_loss_grad: give all three gradient outputs or none");
    amdat_shells *sh = const_cast<amdat_shells *>(sh_c);
    if (first_shell < 0 || last_shell < first_shell || last_shell >= sh->n_shells)
        return set_error(AMDAT_ERANGE, "amdat_sk_loss_grad: shell range outside the table");
    if (!(length_scale > 0.0f)) return set_error(AMDAT_EINVAL, "amdat_sk_loss_grad: length scale must be > 0");
    HIP_TRY(hipSetDevice(ctx->device));
    // the plan of the exact-phase kernel, which computes a phase per slot
    KPlanOwned *plan = nullptr;
    const PlanCost pc = plan_cost(AMDAT_PREC_FP32X, false);
    int rc = build_plan(ctx, sh, first_shell, last_shell, &plan, 0, pc.c, pc.grouped);
    if (rc) return rc;
    const KPlan &kp = plan->kp;
    const int K = kp.K, nsh = last_shell - first_shell + 1, v0 = sh->off[first_shell];
    if (K == 0) {
        // no vectors: S = 0 on every shell, a constant loss, zero gradient
        double L = 0.0;
        for (int s = 0; s < nsh; s++) {
            if (S) S[s] = 0.0;
            L += weights[s] * target[s] * target[s];
        }
        *loss = L;
        if (d_gx) {
            HIP_TRY(hipMemsetAsync(d_gx, 0, sizeof(double) * n_atoms, stream_of(ctx)));
            HIP_TRY(hipMemsetAsync(d_gy, 0, sizeof(double) * n_atoms, stream_of(ctx)));
            HIP_TRY(hipMemsetAsync(d_gz, 0, sizeof(double) * n_atoms, stream_of(ctx)));
            HIP_TRY(hipStreamSynchronize(stream_of(ctx)));
        }
        return AMDAT_OK;
    }
    hipStream_t st = stream_of(ctx);
    TrajDev view;
    view.N = n_atoms;
    view.T = 1;
    view.x = d_x; view.y = d_y; view.z = d_z;
    const double inv_ell = 1.0 / (double)length_scale;
    // forward: rho for the frame (FP64 result); scratch from the context
    // (the call ends with a stream synchronisation)
    int *dfr = nullptr;
    HIP_TRY(ctx_scratch(ctx, 0, 1, &dfr));
    HIP_TRY(hipMemsetAsync(dfr, 0, sizeof(int), st));
    const int n_splits = choose_splits((long)kp.n_chunks, n_atoms);
    double2 *rho = nullptr;
    HIP_TRY(ctx_scratch(ctx, 1, (size_t)K * (n_splits > 1 ? n_splits + 1 : 1), &rho));
    double2 *part = rho + K;
    // one frame and few vectors per shell: no averaging over frames, so the
    // direct exact-phase kernel (about 1 ulp per term) is used
    HIP_TRY(launch_rho(AMDAT_PREC_FP32X, view, nullptr, n_atoms, dfr, 1, kp, inv_ell, n_splits,
                       n_splits > 1 ? part : rho, st));
    if (n_splits > 1) HIP_TRY(launch_split_reduce2(part, 1, n_splits, K, rho, st));
    // S per shell, the loss and the adjoint weights q_v = c_v rho_v on the
    // device (FP64, K values): no host round trip between the forward and
    // the backward kernels
    if (!sh->d_off.p) HIP_TRY(sh->d_off.upload(sh->off.data(), sh->off.size(), st));
    double *wt = nullptr, *res = nullptr;
    HIP_TRY(ctx_scratch(ctx, 4, (size_t)2 * nsh, &wt));
    HIP_TRY(ctx_scratch(ctx, 5, (size_t)1 + 2 * nsh, &res));
    HIP_TRY(hipMemcpyAsync(wt, weights, sizeof(double) * nsh, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(wt + nsh, target, sizeof(double) * nsh, hipMemcpyHostToDevice, st));
    float2 *dq = nullptr;
    HIP_TRY(ctx_scratch(ctx, 2, (size_t)K, &dq));
    HIP_TRY(launch_adjoint_weights(rho, sh->d_off.p + first_shell, nsh, n_atoms, wt, res, dq, st));
    if (d_gx) {
        // backward: the adjoint kernel recomputes every phase
        if (!sh->d_nvec.p) HIP_TRY(sh->d_nvec.upload(sh->nvec.data(), sh->nvec.size(), st));
        const int g_splits = sk_adjoint_splits(n_atoms, K);
        double *g_part = nullptr;
        if (g_splits > 1) HIP_TRY(ctx_scratch(ctx, 3, (size_t)g_splits * 3 * n_atoms, &g_part));
        HIP_TRY(launch_sk_adjoint(d_x, d_y, d_z, n_atoms, inv_ell, sh->d_nvec.p + 3 * (size_t)v0, dq, K, g_splits,
                                  g_part, d_gx, d_gy, d_gz, st));
    }
    std::vector<double> h((size_t)1 + nsh);
    HIP_TRY(hipMemcpyAsync(h.data(), res, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    *loss = h[0];
    if (S)
        for (int s = 0; s < nsh; s++) S[s] = h[1 + s];
    return AMDAT_OK;
}

int amdat_sk_loss_grad(amdat_ctx *ctx, int n_atoms, const float *d_x, const float *d_y, const float *d_z,
                       const amdat_shells *sh_c, int first_shell, int last_shell, float length_scale,
                       const double *weights, const double *target, double *loss, double *S, double *d_gx,
                       double *d_gy, double *d_gz)
{
    AMDAT_GUARD("amdat_sk_loss_grad", amdat_sk_loss_grad_body(ctx, n_atoms, d_x, d_y, d_z, sh_c, first_shell, last_shell, length_scale, weights, target, loss, S, d_gx, d_gy, d_gz))
}

static int amdat_batch_create_body(amdat_ctx *ctx, int n_traj, const amdat_traj *const *trajs, const amdat_list *const *lists,
                       amdat_batch **out)
{
    clear_error();
    if (!ctx || !trajs || !out || n_traj < 1) return set_error(AMDAT_EINVAL, "amdat_batch_create: bad arguments");
    HIP_TRY(hipSetDevice(ctx->device));
    auto b = std::make_unique<amdat_batch>();
    b->n = n_traj;
    for (int i = 0; i < n_traj; i++) {
        const amdat_traj *t = trajs[i];
        const amdat_list *l = lists ? lists[i] : nullptr;
        if (!t) return set_error(AMDAT_EINVAL, "amdat_batch_create: null trajectory");
        if (i == 0) b->T = t->dev.T;
        else if (t->dev.T != b->T)
            return set_error(AMDAT_EINVAL, "amdat_batch_create: trajectories differ in frame count");
        if (l && l->n_atoms != t->dev.N)
            return set_error(AMDAT_EINVAL, "amdat_batch_create: list built for another trajectory size");
        if (list_timed(l))
            return set_error(AMDAT_EINVAL, "amdat_batch_create: the batched form takes static lists only");
        b->trajs.push_back(t);
        b->lists.push_back(l);
    }
    HIP_TRY(b->counter.alloc(1));
    *out = b.release();
    return AMDAT_OK;
}

int amdat_batch_create(amdat_ctx *ctx, int n_traj, const amdat_traj *const *trajs, const amdat_list *const *lists,
                       amdat_batch **out)
{
    AMDAT_GUARD("amdat_batch_create", amdat_batch_create_body(ctx, n_traj, trajs, lists, out))
}

static int amdat_batch_free_body(amdat_ctx *ctx_unused, amdat_batch *b)
{
    (void)ctx_unused;
    delete b;
    return AMDAT_OK;
}

int amdat_batch_free(amdat_ctx *ctx_unused, amdat_batch *b)
{
    AMDAT_GUARD("amdat_batch_free", amdat_batch_free_body(ctx_unused, b))
}

static BatchTraj batch_entry(const amdat_batch *b, int i, double inv_ell, int K, double *out)
{
    const TrajDev &d = b->trajs[i]->dev;
    BatchTraj e;
    e.x = d.x; e.y = d.y; e.z = d.z;
    e.ix = d.ix; e.iy = d.iy; e.iz = d.iz;
    e.box = d.box;
    e.list = list_ptr(b->lists[i]);
    e.N = d.N;
    e.n_list = list_n(b->trajs[i], b->lists[i]);
    e.K = K;
    e.inv_ell = inv_ell;
    e.out = out;
    return e;
}

// Work items in queue order: costly first, ties in (trajectory, chunk, group) order.
static void order_items(std::vector<int4> &items, const std::vector<double> &cost)
{
    std::vector<int> idx(items.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<int4> sorted(items.size());
    for (size_t i = 0; i < idx.size(); i++) sorted[i] = items[idx[i]];
    items.swap(sorted);
}

static int amdat_batch_sk_body(amdat_ctx *ctx, amdat_batch *batch, const amdat_shells *sh_c, int first_shell, int last_shell,
                   const float *length_scales, const int *frames, int n_frames, double *S, double *count)
{
    clear_error();
    if (!ctx || !batch || !sh_c || !length_scales || !S || !count)
        return set_error(AMDAT_EINVAL, "amdat_batch_sk: null argument");
    amdat_shells *sh = const_cast<amdat_shells *>(sh_c);
    if (first_shell < 0 || last_shell < first_shell || last_shell >= sh->n_shells)
        return set_error(AMDAT_ERANGE, "amdat_batch_sk: shell range outside the table");
    const int n = batch->n, T = batch->T;
    for (int i = 0; i < n; i++)
        if (!(length_scales[i] > 0.0f)) return set_error(AMDAT_EINVAL, "amdat_batch_sk: length scale must be > 0");
    std::vector<int> fr;
    if (frames) {
        if (n_frames < 1) return set_error(AMDAT_EINVAL, "amdat_batch_sk: empty frame list");
        fr.assign(frames, frames + n_frames);
        for (int f : fr)
            if (f < 0 || f >= T) return set_error(AMDAT_ERANGE, "amdat_batch_sk: frame out of range");
    } else {
        fr.resize(T);
        std::iota(fr.begin(), fr.end(), 0);
    }
    HIP_TRY(hipSetDevice(ctx->device));
    // the plan the single-trajectory S(k) path chooses (grouped slots allowed)
    KPlanOwned *plan = nullptr;
    int rc = build_plan(ctx, sh, first_shell, last_shell, &plan, 0, 9.0, true);
    if (rc) return rc;
    const KPlan &kp = plan->kp;
    const int K = kp.K, nsh = last_shell - first_shell + 1, F = (int)fr.size();
    // frame groups: enough items to balance the queue (AMDAT_BATCH_ITEMS, default 16384)
    long target = 16384;
    if (const char *e = std::getenv("AMDAT_BATCH_ITEMS")) target = std::max(1L, std::atol(e));
    const long units = std::max(1L, (long)n * kp.n_chunks);
    int G = (int)std::max(1L, std::min((long)F, (target + units - 1) / units));
    const int FG = (F + G - 1) / G;
    G = (F + FG - 1) / FG;
    hipStream_t st = stream_of(ctx);
    DevBuf<double> rows, d_norm, d_S;
    HIP_TRY(rows.alloc((size_t)n * G * K));
    std::vector<BatchTraj> bt(n);
    std::vector<double> norm(n);
    std::vector<int4> items;
    std::vector<double> cost;
    for (int i = 0; i < n; i++) {
        bt[i] = batch_entry(batch, i, 1.0 / (double)length_scales[i], K, rows.p + (size_t)i * G * K);
        norm[i] = (double)F * (double)bt[i].n_list;
        count[i] = norm[i];
        if (bt[i].n_list == 0) continue;
        for (int c = 0; c < kp.n_chunks; c++)
            for (int g = 0; g < G; g++) {
                items.push_back(make_int4(i, c, g, 0));
                cost.push_back((double)bt[i].n_list * (double)(std::min(F, (g + 1) * FG) - g * FG));
            }
    }
    order_items(items, cost);
    std::vector<int> soff(nsh + 1);
    for (int s = 0; s <= nsh; s++) soff[s] = sh->off[first_shell + s] - sh->off[first_shell];
    DevBuf<BatchTraj> d_bt;
    DevBuf<int4> d_items;
    DevBuf<int> d_fr, d_soff;
    HIP_TRY(d_bt.upload(bt.data(), bt.size(), st));
    HIP_TRY(d_items.upload(items.data(), items.size(), st));
    HIP_TRY(d_fr.upload(fr.data(), fr.size(), st));
    HIP_TRY(d_soff.upload(soff.data(), soff.size(), st));
    HIP_TRY(d_norm.upload(norm.data(), norm.size(), st));
    HIP_TRY(d_S.alloc((size_t)n * nsh));
    HIP_TRY(launch_sk_batch(d_bt.p, d_items.p, (int)items.size(), batch->counter.p, d_fr.p, F, FG, kp, st));
    HIP_TRY(launch_sk_batch_reduce(d_bt.p, n, G, K, d_soff.p, nsh, d_norm.p, d_S.p, st));
    HIP_TRY(hipMemcpyAsync(S, d_S.p, sizeof(double) * n * nsh, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return AMDAT_OK;
}

int amdat_batch_sk(amdat_ctx *ctx, amdat_batch *batch, const amdat_shells *sh_c, int first_shell, int last_shell,
                   const float *length_scales, const int *frames, int n_frames, double *S, double *count)
{
    AMDAT_GUARD("amdat_batch_sk", amdat_batch_sk_body(ctx, batch, sh_c, first_shell, last_shell, length_scales, frames, n_frames, S, count))
}

static int amdat_batch_fs_body(amdat_ctx *ctx, amdat_batch *batch, const amdat_shells *sh_c, const int *first_shell,
                   const int *last_shell, const float *length_scales, const amdat_pairs *pairs, double *Fs,
                   double *count)
{
    clear_error();
    if (!ctx || !batch || !sh_c || !first_shell || !last_shell || !length_scales || !pairs || !Fs || !count)
        return set_error(AMDAT_EINVAL, "amdat_batch_fs: null argument");
    amdat_shells *sh = const_cast<amdat_shells *>(sh_c);
    const int n = batch->n, ntg = pairs->ntg;
    int rc = check_pairs_frames(pairs, batch->T);
    if (rc) return rc;
    int Kmax = 0, M = 0;
    for (int i = 0; i < n; i++) {
        if (first_shell[i] < 0 || last_shell[i] < first_shell[i] || last_shell[i] >= sh->n_shells)
            return set_error(AMDAT_ERANGE, "amdat_batch_fs: shell range outside the table");
        if (!(length_scales[i] > 0.0f)) return set_error(AMDAT_EINVAL, "amdat_batch_fs: length scale must be > 0");
        const int K = sh->off[last_shell[i] + 1] - sh->off[first_shell[i]];
        Kmax = std::max(Kmax, K);
        M = std::max(M, shell_range_M(sh, first_shell[i], last_shell[i]));
    }
    if (M > 127) return set_error(AMDAT_ERANGE, "amdat_batch_fs: k-vector component beyond +-127");
    HIP_TRY(hipSetDevice(ctx->device));
    // one R and one table width for the launch; per-trajectory plans concatenated
    const int R = choose_R(std::max(Kmax, 1), M, 6.0);
    std::vector<KPlanOwned *> plans(n);
    std::vector<int> chunk_base(n + 1, 0), row_off(n + 1, 0);
    for (int i = 0; i < n; i++) {
        rc = build_plan(ctx, sh, first_shell[i], last_shell[i], &plans[i], R);
        if (rc) return rc;
        chunk_base[i + 1] = chunk_base[i] + plans[i]->kp.n_chunks;
    }
    KPlan cat;
    cat.R = R;
    cat.KB = kPhaseBlock * R;
    cat.n_chunks = chunk_base[n];
    cat.MM = plan_table_width(M);
    cat.K = 0;
    hipStream_t st = stream_of(ctx);
    DevBuf<uint32_t> kpack;
    DevBuf<int> kdst;
    DevBuf<int4> range;
    HIP_TRY(kpack.alloc((size_t)cat.n_chunks * cat.KB));
    HIP_TRY(kdst.alloc((size_t)cat.n_chunks * cat.KB));
    HIP_TRY(range.alloc(cat.n_chunks));
    for (int i = 0; i < n; i++) {
        const KPlan &p = plans[i]->kp;
        const size_t o = (size_t)chunk_base[i] * cat.KB, m = (size_t)p.n_chunks * cat.KB;
        if (m == 0) continue;
        HIP_TRY(hipMemcpyAsync(kpack.p + o, p.kpack, m * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(kdst.p + o, p.kdst, m * sizeof(int), hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(range.p + chunk_base[i], p.range, p.n_chunks * sizeof(int4), hipMemcpyDeviceToDevice, st));
    }
    cat.kpack = kpack.p;
    cat.kdst = kdst.p;
    cat.range = range.p;
    // rows [gap][K_b] per trajectory; shells, norms and output offsets
    std::vector<int> soff, shell_base(n), nsh(n), out_off(n + 1, 0);
    std::vector<double> norm((size_t)n * ntg, 0.0);
    int max_nsh = 0;
    for (int i = 0; i < n; i++) {
        const int K = plans[i]->kp.K;
        row_off[i + 1] = row_off[i] + ntg * K;
        nsh[i] = last_shell[i] - first_shell[i] + 1;
        max_nsh = std::max(max_nsh, nsh[i]);
        shell_base[i] = (int)soff.size();
        for (int s = 0; s <= nsh[i]; s++) soff.push_back(sh->off[first_shell[i] + s] - sh->off[first_shell[i]]);
        out_off[i + 1] = out_off[i] + ntg * nsh[i];
        const int nl = list_n(batch->trajs[i], batch->lists[i]);
        for (int q = 0; q < pairs->n; q++) norm[(size_t)i * ntg + pairs->tg[q]] += (double)nl;
    }
    for (size_t q = 0; q < norm.size(); q++) count[q] = norm[q];
    DevBuf<double> rows, d_norm, d_out;
    HIP_TRY(rows.alloc((size_t)row_off[n]));
    std::vector<BatchTraj> bt(n);
    std::vector<int4> items;
    std::vector<double> cost;
    for (int i = 0; i < n; i++) {
        bt[i] = batch_entry(batch, i, 1.0 / (double)length_scales[i], plans[i]->kp.K, rows.p + row_off[i]);
        if (bt[i].n_list == 0) continue;
        for (int c = 0; c < plans[i]->kp.n_chunks; c++)
            for (int g = 0; g < ntg; g++) {
                const int np = pairs->group_off[g + 1] - pairs->group_off[g];
                if (np == 0) continue;
                items.push_back(make_int4(i, chunk_base[i] + c, g, 0));
                cost.push_back((double)bt[i].n_list * np);
            }
    }
    order_items(items, cost);
    DevBuf<BatchTraj> d_bt;
    DevBuf<int4> d_items;
    DevBuf<int> d_soff, d_base, d_nsh, d_oo;
    HIP_TRY(d_bt.upload(bt.data(), bt.size(), st));
    HIP_TRY(d_items.upload(items.data(), items.size(), st));
    HIP_TRY(d_soff.upload(soff.data(), soff.size(), st));
    HIP_TRY(d_base.upload(shell_base.data(), shell_base.size(), st));
    HIP_TRY(d_nsh.upload(nsh.data(), nsh.size(), st));
    HIP_TRY(d_oo.upload(out_off.data(), out_off.size(), st));
    HIP_TRY(d_norm.upload(norm.data(), norm.size(), st));
    HIP_TRY(d_out.alloc((size_t)out_off[n]));
    HIP_TRY(launch_fs_batch(d_bt.p, d_items.p, (int)items.size(), batch->counter.p, pairs->d_t0.p, pairs->d_t1.p,
                            pairs->d_group_off.p, pairs->d_order.p, cat, st));
    HIP_TRY(launch_fs_batch_reduce(d_bt.p, n, d_soff.p, d_base.p, d_nsh.p, max_nsh, ntg, nullptr, d_norm.p, d_oo.p,
                                   d_out.p, st));
    HIP_TRY(hipMemcpyAsync(Fs, d_out.p, sizeof(double) * out_off[n], hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return AMDAT_OK;
}

int amdat_batch_fs(amdat_ctx *ctx, amdat_batch *batch, const amdat_shells *sh_c, const int *first_shell,
                   const int *last_shell, const float *length_scales, const amdat_pairs *pairs, double *Fs,
                   double *count)
{
    AMDAT_GUARD("amdat_batch_fs", amdat_batch_fs_body(ctx, batch, sh_c, first_shell, last_shell, length_scales, pairs, Fs, count))
}

static int amdat_batch_msd_ngp_body(amdat_ctx *ctx, amdat_batch *batch, const amdat_pairs *pairs, double *s2, double *s4,
                        double *count)
{
    clear_error();
    if (!ctx || !batch || !pairs || !s2 || !s4 || !count)
        return set_error(AMDAT_EINVAL, "amdat_batch_msd_ngp: null argument");
    int rc = check_pairs_frames(pairs, batch->T);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    const int n = batch->n, ntg = pairs->ntg, P = pairs->n;
    std::vector<BatchTraj> bt(n);
    std::vector<int2> items;
    for (int i = 0; i < n; i++) {
        bt[i] = batch_entry(batch, i, 1.0, 0, nullptr);
        for (int g = 0; g < ntg; g++) count[(size_t)i * ntg + g] = 0.0;
        for (int q = 0; q < P; q++) count[(size_t)i * ntg + pairs->tg[q]] += (double)bt[i].n_list;
        if (bt[i].n_list == 0) continue;
        for (int q = 0; q < P; q++) items.push_back(make_int2(i, q));
    }
    hipStream_t st = stream_of(ctx);
    DevBuf<BatchTraj> d_bt;
    DevBuf<int2> d_items;
    DevBuf<double> part, d2, d4;
    HIP_TRY(d_bt.upload(bt.data(), bt.size(), st));
    HIP_TRY(d_items.upload(items.data(), items.size(), st));
    HIP_TRY(part.alloc(std::max<size_t>(1, (size_t)n * P * 2)));
    HIP_TRY(hipMemsetAsync(part.p, 0, part.n * sizeof(double), st));
    const int pb = 1 << 20;  // launch in slices of at most 2^20 blocks
    for (size_t i0 = 0; i0 < items.size(); i0 += pb) {
        const int ni = (int)std::min<size_t>(pb, items.size() - i0);
        HIP_TRY(launch_msd_batch(d_bt.p, d_items.p + i0, ni, pairs->d_t0.p, pairs->d_t1.p, P, part.p, st));
    }
    HIP_TRY(d2.alloc(std::max(1, n * ntg)));
    HIP_TRY(d4.alloc(std::max(1, n * ntg)));
    HIP_TRY(launch_msd_batch_reduce(part.p, P, pairs->d_group_off.p, pairs->d_order.p, ntg, n, d2.p, d4.p, st));
    HIP_TRY(hipMemcpyAsync(s2, d2.p, sizeof(double) * n * ntg, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s4, d4.p, sizeof(double) * n * ntg, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return AMDAT_OK;
}

int amdat_batch_msd_ngp(amdat_ctx *ctx, amdat_batch *batch, const amdat_pairs *pairs, double *s2, double *s4,
                        double *count)
{
    AMDAT_GUARD("amdat_batch_msd_ngp", amdat_batch_msd_ngp_body(ctx, batch, pairs, s2, s4, count))
}

}  // extern "C"

namespace {

// Enqueue a planned job: persistent launch, then the per-shell reductions.
hipError_t enqueue_job_sequence(JobCache &jc, hipStream_t st)
{
    hipError_t e = enqueue_job(jc.L, jc.job, st);
    if (e == hipSuccess && jc.nsh_sk > 0)
        e = launch_sk_batch_reduce(jc.sk_bt.p, jc.n, jc.G, jc.job.sk.K, jc.sk_soff.p, jc.nsh_sk, jc.sk_norm.p, jc.S.p,
                                   st);
    if (e == hipSuccess && jc.max_nsh > 0)
        e = launch_fs_batch_reduce(jc.fs_bt.p, jc.n, jc.fs_soff.p, jc.fs_base.p, jc.fs_nsh.p, jc.max_nsh, jc.ntg,
                                   jc.fs_rowoff.p, jc.fs_norm.p, jc.fs_oo.p, jc.Fs.p, st);
    return e;
}

// Frames per S(k) item when the settings leave it open: enough items
// (trajectories x chunks x frame groups) to balance the queue, about
// AMDAT_BATCH_ITEMS (default 16384).
int default_frames_per_item(int n_traj, int n_chunks, int F)
{
    long target = 16384;
    if (const char *e = std::getenv("AMDAT_BATCH_ITEMS")) target = std::max(1L, std::atol(e));
    const long units = std::max(1L, (long)n_traj * n_chunks);
    const int G = (int)std::max(1L, std::min((long)F, (target + units - 1) / units));
    return (F + G - 1) / G;
}

// Pair chunks of at most PC pairs (0: whole gaps) over the pairs' reference
// order: offsets into order[] per chunk, and per gap its first chunk.
void pair_chunks(const amdat_pairs *pairs, int PC, std::vector<int> &pc_off, std::vector<int> &gap_pc)
{
    const int ntg = pairs->ntg;
    pc_off.assign(1, 0);
    gap_pc.assign(ntg + 1, 0);
    for (int g = 0; g < ntg; g++) {
        gap_pc[g] = (int)pc_off.size() - 1;
        const int b = pairs->group_off[g], e = pairs->group_off[g + 1];
        const int step = PC > 0 ? PC : std::max(1, e - b);
        for (int q = b; q < e; q += step) pc_off.push_back(std::min(e, q + step));
    }
    gap_pc[ntg] = (int)pc_off.size() - 1;
}

// Plan a job for the trajectories with use[b] != 0 (others get no items)
// with the launch settings ch.
int plan_job(amdat_ctx *ctx, amdat_batch *batch, const JobParams &prm, const std::vector<char> &use,
             const JobChoice &ch, JobCache &jc)
{
    const int n = batch->n;
    hipStream_t st = stream_of(ctx);
    jc.prm = prm;
    jc.n = n;
    jc.ch = ch;
    jc.ch.AF = std::max(1, ch.AF);
    jc.ntg = prm.pairs ? prm.pairs->ntg : 0;
    std::vector<int> nl(n);
    for (int b = 0; b < n; b++) nl[b] = list_n(batch->trajs[b], batch->lists[b]);
    std::vector<int4> items;
    std::vector<double> cost;
    int M = 0;
    // S(k): one plan, frame groups
    KPlan skp;
    if (prm.sk_sh) {
        amdat_shells *sh = const_cast<amdat_shells *>(prm.sk_sh);
        KPlanOwned *plan = nullptr;
        int rc = build_plan(ctx, sh, prm.sk_first, prm.sk_last, &plan, ch.RS, 9.0, true, ch.SM);
        if (rc) return rc;
        skp = plan->kp;
        M = std::max(M, shell_range_M(sh, prm.sk_first, prm.sk_last));
        const int F = (int)prm.frames.size(), K = skp.K;
        jc.nsh_sk = prm.sk_last - prm.sk_first + 1;
        const int FG = ch.FG > 0 ? std::min(ch.FG, F) : default_frames_per_item(n, skp.n_chunks, F);
        const int G = (F + FG - 1) / FG;
        jc.ch.FG = FG;
        jc.G = G;
        HIP_TRY(jc.sk_rows.alloc((size_t)n * G * K));
        std::vector<BatchTraj> bt(n);
        std::vector<double> norm(n);
        jc.sk_count.assign(n, 0.0);
        for (int b = 0; b < n; b++) {
            bt[b] = batch_entry(batch, b, 1.0 / (double)prm.ell[b], K, jc.sk_rows.p + (size_t)b * G * K);
            norm[b] = (double)F * nl[b];
            jc.sk_count[b] = norm[b];
            if (!use[b] || nl[b] == 0) continue;
            for (int c = 0; c < skp.n_chunks; c++)
                for (int g = 0; g < G; g++) {
                    items.push_back(make_int4(b, c, g, 0));
                    cost.push_back((double)nl[b] * (std::min(F, (g + 1) * FG) - g * FG) * skp.KB);
                }
        }
        std::vector<int> soff(jc.nsh_sk + 1);
        for (int s = 0; s <= jc.nsh_sk; s++) soff[s] = sh->off[prm.sk_first + s] - sh->off[prm.sk_first];
        HIP_TRY(jc.sk_bt.upload(bt.data(), bt.size(), st));
        HIP_TRY(jc.frames.upload(prm.frames.data(), prm.frames.size(), st));
        HIP_TRY(jc.sk_soff.upload(soff.data(), soff.size(), st));
        HIP_TRY(jc.sk_norm.upload(norm.data(), norm.size(), st));
        HIP_TRY(jc.S.alloc((size_t)n * jc.nsh_sk));
        jc.job.frames = jc.frames.p;
        jc.job.n_frames = F;
        jc.job.group_size = FG;
        jc.job.sk_trajs = jc.sk_bt.p;
    }
    // Fs: per-trajectory plans, concatenated
    KPlan fsp;
    const amdat_pairs *pairs = prm.pairs;
    if (prm.fs_sh) {
        amdat_shells *sh = const_cast<amdat_shells *>(prm.fs_sh);
        const int ntg = jc.ntg;
        std::vector<KPlanOwned *> plans(n);
        std::vector<int> chunk_base(n + 1, 0), row_off(n + 1, 0);
        for (int b = 0; b < n; b++) {
            int rc = build_plan(ctx, sh, prm.fs_first[b], prm.fs_last[b], &plans[b], ch.RF);
            if (rc) return rc;
            M = std::max(M, shell_range_M(sh, prm.fs_first[b], prm.fs_last[b]));
            chunk_base[b + 1] = chunk_base[b] + plans[b]->kp.n_chunks;
        }
        fsp.R = ch.RF;
        fsp.KB = kPhaseBlock * ch.RF;
        fsp.n_chunks = chunk_base[n];
        HIP_TRY(jc.fs_kpack.alloc((size_t)fsp.n_chunks * fsp.KB));
        HIP_TRY(jc.fs_kdst.alloc((size_t)fsp.n_chunks * fsp.KB));
        HIP_TRY(jc.fs_range.alloc(fsp.n_chunks));
        for (int b = 0; b < n; b++) {
            const KPlan &p = plans[b]->kp;
            const size_t o = (size_t)chunk_base[b] * fsp.KB, m = (size_t)p.n_chunks * fsp.KB;
            if (m == 0) continue;
            HIP_TRY(hipMemcpyAsync(jc.fs_kpack.p + o, p.kpack, m * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
            HIP_TRY(hipMemcpyAsync(jc.fs_kdst.p + o, p.kdst, m * sizeof(int), hipMemcpyDeviceToDevice, st));
            HIP_TRY(hipMemcpyAsync(jc.fs_range.p + chunk_base[b], p.range, p.n_chunks * sizeof(int4),
                                   hipMemcpyDeviceToDevice, st));
        }
        fsp.kpack = jc.fs_kpack.p;
        fsp.kdst = jc.fs_kdst.p;
        fsp.range = jc.fs_range.p;
        // rows: pair chunk x atom part, a gap's rows consecutive
        const int AF = jc.ch.AF;
        std::vector<int> pc_off, gap_pc;
        pair_chunks(pairs, ch.PC, pc_off, gap_pc);
        jc.n_pc = (int)pc_off.size() - 1;
        std::vector<int> gap_rows(ntg + 1);
        for (int g = 0; g <= ntg; g++) gap_rows[g] = gap_pc[g] * AF;
        const int n_rows = jc.n_pc * AF;
        std::vector<int> soff, base(n), nsh(n);
        jc.out_off.assign(n + 1, 0);
        std::vector<double> norm((size_t)n * ntg, 0.0);
        for (int b = 0; b < n; b++) {
            row_off[b + 1] = row_off[b] + n_rows * plans[b]->kp.K;
            nsh[b] = prm.fs_last[b] - prm.fs_first[b] + 1;
            jc.max_nsh = std::max(jc.max_nsh, nsh[b]);
            base[b] = (int)soff.size();
            for (int s = 0; s <= nsh[b]; s++) soff.push_back(sh->off[prm.fs_first[b] + s] - sh->off[prm.fs_first[b]]);
            jc.out_off[b + 1] = jc.out_off[b] + ntg * nsh[b];
            for (int q = 0; q < pairs->n; q++) norm[(size_t)b * ntg + pairs->tg[q]] += (double)nl[b];
        }
        jc.fs_count = norm;
        HIP_TRY(jc.fs_rows.alloc((size_t)row_off[n]));
        std::vector<BatchTraj> bt(n);
        for (int b = 0; b < n; b++) {
            bt[b] = batch_entry(batch, b, 1.0 / (double)prm.ell[b], plans[b]->kp.K, jc.fs_rows.p + row_off[b]);
            if (!use[b] || nl[b] == 0) continue;
            for (int c = 0; c < plans[b]->kp.n_chunks; c++)
                for (int q = 0; q < jc.n_pc; q++)
                    for (int a = 0; a < AF; a++) {
                        const int np = pc_off[q + 1] - pc_off[q];
                        const long na = ((long)(a + 1) * nl[b]) / AF - ((long)a * nl[b]) / AF;
                        items.push_back(make_int4(b, chunk_base[b] + c, q * AF + a, 1));
                        cost.push_back((double)na * np * fsp.KB);
                    }
        }
        HIP_TRY(jc.fs_pcoff.upload(pc_off.data(), pc_off.size(), st));
        HIP_TRY(jc.fs_rowoff.upload(gap_rows.data(), gap_rows.size(), st));
        jc.job.fs_pc = jc.fs_pcoff.p;
        jc.job.fs_split = AF;
        HIP_TRY(jc.fs_bt.upload(bt.data(), bt.size(), st));
        HIP_TRY(jc.fs_soff.upload(soff.data(), soff.size(), st));
        HIP_TRY(jc.fs_base.upload(base.data(), base.size(), st));
        HIP_TRY(jc.fs_nsh.upload(nsh.data(), nsh.size(), st));
        HIP_TRY(jc.fs_oo.upload(jc.out_off.data(), jc.out_off.size(), st));
        HIP_TRY(jc.fs_norm.upload(norm.data(), norm.size(), st));
        HIP_TRY(jc.Fs.alloc((size_t)jc.out_off[n]));
        jc.job.fs_trajs = jc.fs_bt.p;
    }
    // MSD/NGP: one item per (trajectory, gap)
    if (prm.msd) {
        const int ntg = jc.ntg;
        jc.msd_count.assign((size_t)n * ntg, 0.0);
        jc.msd_item.assign((size_t)n * ntg, 0);
        std::vector<BatchTraj> bt(n);
        for (int b = 0; b < n; b++) {
            bt[b] = batch_entry(batch, b, 1.0, 0, nullptr);
            for (int q = 0; q < pairs->n; q++) jc.msd_count[(size_t)b * ntg + pairs->tg[q]] += (double)nl[b];
            if (!use[b] || nl[b] == 0) continue;
            for (int g = 0; g < ntg; g++) {
                const int np = pairs->group_off[g + 1] - pairs->group_off[g];
                if (np == 0) continue;
                items.push_back(make_int4(b, g, 0, 2));
                cost.push_back((double)nl[b] * np * 4.0);
                jc.msd_item[(size_t)b * ntg + g] = 1;
            }
        }
        if (!jc.sk_bt.p && !jc.fs_bt.p) HIP_TRY(jc.sk_bt.upload(bt.data(), bt.size(), st));
        jc.job.msd_trajs = jc.sk_bt.p ? jc.sk_bt.p : jc.fs_bt.p;
        HIP_TRY(jc.s2.alloc(std::max(1, n * ntg)));
        HIP_TRY(jc.s4.alloc(std::max(1, n * ntg)));
        jc.job.s2 = jc.s2.p;
        jc.job.s4 = jc.s4.p;
    }
    if (pairs) {
        jc.job.t0 = pairs->d_t0.p;
        jc.job.t1 = pairs->d_t1.p;
        jc.job.group_off = pairs->d_group_off.p;
        jc.job.order = pairs->d_order.p;
        jc.job.ntg = jc.ntg;
    }
    if (M > 127) return set_error(AMDAT_ERANGE, "amdat_batch_run: k-vector component beyond +-127");
    jc.MM = plan_table_width(M);
    skp.MM = fsp.MM = jc.MM;
    jc.job.sk = skp;
    jc.job.fs = fsp;
    order_items(items, cost);
    HIP_TRY(jc.items.upload(items.data(), items.size(), st));
    jc.job.items = jc.items.p;
    jc.job.n_items = (int)items.size();
    jc.job.counter = batch->counter.p;
    HIP_TRY(prepare_job(jc.ch.RS, jc.ch.RF, jc.MM, jc.ch.SM != 0, std::max(1, jc.job.n_items), &jc.L));
    return AMDAT_OK;
}

// Time one planned job (second of two runs), in ms.
int time_job(amdat_ctx *ctx, JobCache &jc, float *ms)
{
    hipStream_t st = stream_of(ctx);
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    hipError_t e = enqueue_job_sequence(jc, st);
    if (e == hipSuccess) e = hipEventRecord(a, st);
    if (e == hipSuccess) e = enqueue_job_sequence(jc, st);
    if (e == hipSuccess) e = hipEventRecord(b, st);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    HIP_TRY(e);
    return AMDAT_OK;
}

// Bytes of the Fs rows of a job with pair chunks of PC pairs and AF atom
// parts (FP64, every trajectory).
double fs_rows_bytes(const amdat_batch *batch, const JobParams &prm, int PC, int AF)
{
    if (!prm.fs_sh || !prm.pairs) return 0.0;
    std::vector<int> pc_off, gap_pc;
    pair_chunks(prm.pairs, PC, pc_off, gap_pc);
    double k = 0.0;
    for (int b = 0; b < batch->n; b++) k += prm.fs_sh->off[prm.fs_last[b] + 1] - prm.fs_sh->off[prm.fs_first[b]];
    return 8.0 * k * (double)(pc_off.size() - 1) * AF;
}

// Autotuner.  Starts from the cost model's settings `model` and keeps, for
// the S(k) items and the Fs items separately, the fastest of:
//  1. vectors per thread and plan mode: S(k) R = 2, 4, 8 unpaired, R = 4, 8
//     paired and grouped, timed over every 8th frame of every trajectory
//     (a sample of trajectories instead leaves too few items per block:
//     its time measures the queue's tail, not the per-update speed); Fs R =
//     2, 4, 8.
//  2. granularity (the queue's balance depends on the whole job): S(k)
//     frames per item x1/2, x1, x2; Fs pairs per item (whole gaps, 2, 1),
//     then atom parts per pair chunk (1, 2, 4).
// Each part is timed alone (S(k)-only or Fs-only jobs).  Finally the whole
// job (with MSD) is timed with the model's and the tuned settings and the
// faster is kept, so the tuned launch is never slower than the model's as
// measured here.  Every timing is the second of two runs.
int autotune_job(amdat_ctx *ctx, amdat_batch *batch, const JobParams &prm, const JobChoice &model, TuneResult *res)
{
    const int n = batch->n;
    std::vector<char> all(n, 1);
    auto time_of = [&](const JobParams &p, const std::vector<char> &use, const JobChoice &c, float *ms,
                       int *fg_out) -> int {
        JobCache jc;
        int rc = plan_job(ctx, batch, p, use, c, jc);
        if (rc) return rc;
        if (fg_out) *fg_out = jc.ch.FG;
        return time_job(ctx, jc, ms);
    };
    JobChoice best = model;
    const double max_rows = 8.0 * (1ull << 30);  // Fs rows of a candidate: at most 8 GiB
    if (prm.sk_sh) {
        JobParams p = prm;
        p.fs_sh = nullptr;
        p.msd = false;
        // per-update speed on every trajectory over every 8th frame: the
        // queue keeps its shape (item count, sizes, balance), at an eighth
        // of the work
        JobParams ps = p;
        if (ps.frames.size() >= 16) {
            std::vector<int> fr;
            for (size_t i = 0; i < ps.frames.size(); i += 8) fr.push_back(ps.frames[i]);
            ps.frames.swap(fr);
        }
        static const int cand[8][2] = {{2, 0}, {4, 0}, {8, 0}, {4, 1}, {8, 1}, {4, 4}, {8, 4}, {8, 5}};
        float t_best = 1e30f;
        for (const auto &cd : cand) {
            JobChoice c = best;
            c.RS = cd[0];
            c.SM = cd[1];
            c.FG = 0;
            float ms = 0.0f;
            int rc = time_of(ps, all, c, &ms, nullptr);
            if (rc) return rc;
            if (ms < t_best) { t_best = ms; best.RS = c.RS; best.SM = c.SM; }
        }
        // frames per item around the default for the chosen plan
        JobChoice c = best;
        c.FG = 0;
        int fg0 = 0;
        float ms = 0.0f;
        int rc = time_of(p, all, c, &ms, &fg0);
        if (rc) return rc;
        t_best = ms;
        best.FG = fg0;
        for (int fg : {fg0 / 2, fg0 * 2}) {
            if (fg < 1 || fg == fg0 || fg > (int)p.frames.size()) continue;
            c.FG = fg;
            rc = time_of(p, all, c, &ms, nullptr);
            if (rc) return rc;
            if (ms < t_best) { t_best = ms; best.FG = fg; }
        }
    }
    if (prm.fs_sh) {
        JobParams p = prm;
        p.sk_sh = nullptr;
        p.msd = false;
        float t_best = 1e30f;
        for (int R : {2, 4, 8}) {
            JobChoice c = best;
            c.RF = R;
            c.PC = 0;
            c.AF = 1;
            float ms = 0.0f;
            int rc = time_of(p, all, c, &ms, nullptr);
            if (rc) return rc;
            if (ms < t_best) { t_best = ms; best.RF = R; }
        }
        best.PC = 0;
        best.AF = 1;
        int max_np = 0;
        for (int g = 0; g < prm.pairs->ntg; g++)
            max_np = std::max(max_np, prm.pairs->group_off[g + 1] - prm.pairs->group_off[g]);
        t_best = 1e30f;
        for (int pc : {0, 2, 1}) {
            if (pc > 0 && pc >= max_np) continue;
            if (fs_rows_bytes(batch, p, pc, 1) > max_rows) continue;
            JobChoice c = best;
            c.PC = pc;
            float ms = 0.0f;
            int rc = time_of(p, all, c, &ms, nullptr);
            if (rc) return rc;
            if (ms < t_best) { t_best = ms; best.PC = pc; }
        }
        for (int af : {2, 4}) {
            if (fs_rows_bytes(batch, p, best.PC, af) > max_rows) continue;
            JobChoice c = best;
            c.AF = af;
            float ms = 0.0f;
            int rc = time_of(p, all, c, &ms, nullptr);
            if (rc) return rc;
            if (ms < t_best) { t_best = ms; best.AF = af; }
        }
    }
    // the whole job: model against tuned
    float t_model = 0.0f, t_tuned = 0.0f;
    int rc = time_of(prm, all, model, &t_model, nullptr);
    if (rc) return rc;
    rc = time_of(prm, all, best, &t_tuned, nullptr);
    if (rc) return rc;
    res->model_ms = t_model;
    if (t_tuned <= t_model) {
        res->ch = best;
        res->tuned_ms = t_tuned;
    } else {
        res->ch = model;
        res->tuned_ms = t_model;
    }
    return AMDAT_OK;
}

}  // namespace

extern "C" {

static int amdat_batch_run_body(amdat_ctx *ctx, amdat_batch *batch, const float *length_scales, const amdat_shells *sk_shells,
                    int sk_first, int sk_last, const int *frames, int n_frames, const amdat_shells *fs_shells,
                    const int *fs_first, const int *fs_last, const amdat_pairs *pairs, int want_msd, int flags,
                    double *S, double *S_count, double *Fs, double *Fs_count, double *s2, double *s4,
                    double *msd_count)
{
    clear_error();
    if (!ctx || !batch || !length_scales) return set_error(AMDAT_EINVAL, "amdat_batch_run: null argument");
    if (!sk_shells && !fs_shells && !want_msd) return set_error(AMDAT_EINVAL, "amdat_batch_run: nothing to compute");
    if ((fs_shells || want_msd) && !pairs) return set_error(AMDAT_EINVAL, "amdat_batch_run: Fs and MSD need pairs");
    if (sk_shells && (!S || !S_count)) return set_error(AMDAT_EINVAL, "amdat_batch_run: null S output");
    if (fs_shells && (!Fs || !Fs_count || !fs_first || !fs_last))
        return set_error(AMDAT_EINVAL, "amdat_batch_run: null Fs argument");
    if (want_msd && (!s2 || !s4 || !msd_count)) return set_error(AMDAT_EINVAL, "amdat_batch_run: null MSD output");
    const int n = batch->n, T = batch->T;
    JobParams prm;
    prm.ell.assign(length_scales, length_scales + n);
    for (float l : prm.ell)
        if (!(l > 0.0f)) return set_error(AMDAT_EINVAL, "amdat_batch_run: length scale must be > 0");
    if (sk_shells) {
        if (sk_first < 0 || sk_last < sk_first || sk_last >= sk_shells->n_shells)
            return set_error(AMDAT_ERANGE, "amdat_batch_run: S(k) shell range outside the table");
        prm.sk_sh = sk_shells;
        prm.sk_serial = sk_shells->serial;
        prm.sk_first = sk_first;
        prm.sk_last = sk_last;
        if (frames) {
            if (n_frames < 1) return set_error(AMDAT_EINVAL, "amdat_batch_run: empty frame list");
            prm.frames.assign(frames, frames + n_frames);
            for (int f : prm.frames)
                if (f < 0 || f >= T) return set_error(AMDAT_ERANGE, "amdat_batch_run: frame out of range");
        } else {
            prm.frames.resize(T);
            std::iota(prm.frames.begin(), prm.frames.end(), 0);
        }
    }
    if (pairs) {
        int rc = check_pairs_frames(pairs, T);
        if (rc) return rc;
        prm.pairs = pairs;
        prm.pairs_serial = pairs->serial;
    }
    int Kfs = 0, Mfs = 0;
    if (fs_shells) {
        prm.fs_sh = fs_shells;
        prm.fs_serial = fs_shells->serial;
        prm.fs_first.assign(fs_first, fs_first + n);
        prm.fs_last.assign(fs_last, fs_last + n);
        for (int b = 0; b < n; b++) {
            if (fs_first[b] < 0 || fs_last[b] < fs_first[b] || fs_last[b] >= fs_shells->n_shells)
                return set_error(AMDAT_ERANGE, "amdat_batch_run: Fs shell range outside the table");
            const int K = fs_shells->off[fs_last[b] + 1] - fs_shells->off[fs_first[b]];
            Kfs = std::max(Kfs, K);
            Mfs = std::max(Mfs, shell_range_M(fs_shells, fs_first[b], fs_last[b]));
        }
    }
    prm.msd = want_msd != 0;
    HIP_TRY(hipSetDevice(ctx->device));
    const bool use_graph = (flags & AMDAT_RUN_GRAPH) != 0;
    hipStream_t user = stream_of(ctx);
    // settings from the tuner only when asked for (and not fixed by the caller)
    const bool tune = (flags & AMDAT_RUN_TUNE) != 0 && !batch->forced;
    if (!batch->job || !(batch->job->prm == prm) || (use_graph && !batch->job->exec) || batch->job->tuned != tune) {
        batch->job.reset();
        // launch settings: from the cost model, or autotuned from there
        JobChoice model;
        if (sk_shells) {
            // the plan the single-trajectory S(k) path would choose (cost model
            // over R and plan mode); kept in the table's plan cache
            KPlanOwned *plan = nullptr;
            int rc = build_plan(ctx, const_cast<amdat_shells *>(sk_shells), sk_first, sk_last, &plan, 0, 9.0, true);
            if (rc) return rc;
            model.RS = plan->kp.R;
            model.SM = plan->mode;
            if (model.RS == 2 && model.SM != 0) model.SM = 0;  // no grouped R = 2 job kernel
        }
        if (fs_shells) model.RF = choose_R(std::max(Kfs, 1), Mfs, 6.0);
        if (!model.RS) model.RS = model.RF ? model.RF : 2;
        if (!model.RF) model.RF = model.RS;
        TuneResult tr{model, 0.0f, 0.0f};
        if (batch->forced) {
            const JobChoice &f = batch->force;
            if (f.RS) { tr.ch.RS = f.RS; tr.ch.SM = f.SM; }
            if (f.FG) tr.ch.FG = f.FG;
            if (f.RF) tr.ch.RF = f.RF;
            tr.ch.PC = f.PC;
            tr.ch.AF = std::max(1, f.AF);
        } else if (tune) {
            bool found = false;
            for (auto &t : batch->tuned)
                if (t.first == prm) { tr = t.second; found = true; }
            if (!found) {
                int rc = autotune_job(ctx, batch, prm, model, &tr);
                if (rc) return rc;
                batch->tuned.push_back({prm, tr});
            }
        }
        auto jc = std::make_unique<JobCache>();
        std::vector<char> use(n, 1);
        int rc = plan_job(ctx, batch, prm, use, tr.ch, *jc);
        if (rc) return rc;
        jc->tuned = tune;
        jc->model_ms = tr.model_ms;
        jc->tuned_ms = tr.tuned_ms;
        HIP_TRY(hipStreamSynchronize(user));
        if (use_graph) {
            // capture on the context's own stream (the caller's may be the null stream)
            HIP_TRY(hipStreamBeginCapture(ctx->own, hipStreamCaptureModeThreadLocal));
            hipError_t e = enqueue_job_sequence(*jc, ctx->own);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(ctx->own, &g);
            jc->graph = g;
            HIP_TRY(e);
            HIP_TRY(e2);
            HIP_TRY(hipGraphInstantiate(&jc->exec, g, nullptr, nullptr, 0));
        }
        batch->job = std::move(jc);
    }
    JobCache &jc = *batch->job;
    hipStream_t st = user;
    if (use_graph) {
        // order the replay after the caller's stream, then run it on the own stream
        hipEvent_t ev;
        HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        hipError_t e = hipEventRecord(ev, user);
        if (e == hipSuccess) e = hipStreamWaitEvent(ctx->own, ev, 0);
        (void)hipEventDestroy(ev);
        HIP_TRY(e);
        st = ctx->own;
        HIP_TRY(hipGraphLaunch(jc.exec, st));
    } else {
        HIP_TRY(enqueue_job_sequence(jc, st));
    }
    if (sk_shells) {
        HIP_TRY(hipMemcpyAsync(S, jc.S.p, sizeof(double) * n * jc.nsh_sk, hipMemcpyDeviceToHost, st));
        std::copy(jc.sk_count.begin(), jc.sk_count.end(), S_count);
    }
    if (fs_shells) {
        HIP_TRY(hipMemcpyAsync(Fs, jc.Fs.p, sizeof(double) * jc.out_off[n], hipMemcpyDeviceToHost, st));
        std::copy(jc.fs_count.begin(), jc.fs_count.end(), Fs_count);
    }
    if (want_msd) {
        HIP_TRY(hipMemcpyAsync(s2, jc.s2.p, sizeof(double) * n * jc.ntg, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(s4, jc.s4.p, sizeof(double) * n * jc.ntg, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    if (want_msd) {
        for (size_t i = 0; i < jc.msd_item.size(); i++) {
            if (!jc.msd_item[i]) s2[i] = s4[i] = 0.0;
            msd_count[i] = jc.msd_count[i];
        }
    }
    return AMDAT_OK;
}

int amdat_batch_run(amdat_ctx *ctx, amdat_batch *batch, const float *length_scales, const amdat_shells *sk_shells,
                    int sk_first, int sk_last, const int *frames, int n_frames, const amdat_shells *fs_shells,
                    const int *fs_first, const int *fs_last, const amdat_pairs *pairs, int want_msd, int flags,
                    double *S, double *S_count, double *Fs, double *Fs_count, double *s2, double *s4,
                    double *msd_count)
{
    AMDAT_GUARD("amdat_batch_run", amdat_batch_run_body(ctx, batch, length_scales, sk_shells, sk_first, sk_last, frames, n_frames, fs_shells, fs_first, fs_last, pairs, want_msd, flags, S, S_count, Fs, Fs_count, s2, s4, msd_count))
}

static int amdat_batch_job_info_body(const amdat_batch *batch, int *RS, int *RF, int *paired, int *n_items, int *graph)
{
    if (!batch || !batch->job) return set_error(AMDAT_EINVAL, "amdat_batch_job_info: no job has run");
    if (RS) *RS = batch->job->ch.RS;
    if (RF) *RF = batch->job->ch.RF;
    if (paired) *paired = batch->job->ch.SM != 0;
    if (n_items) *n_items = batch->job->job.n_items;
    if (graph) *graph = batch->job->exec != nullptr;
    return AMDAT_OK;
}

int amdat_batch_job_info(const amdat_batch *batch, int *RS, int *RF, int *paired, int *n_items, int *graph)
{
    AMDAT_GUARD("amdat_batch_job_info", amdat_batch_job_info_body(batch, RS, RF, paired, n_items, graph))
}

static int amdat_batch_use_settings_body(amdat_batch *batch, const amdat_job_settings *s)
{
    clear_error();
    if (!batch) return set_error(AMDAT_EINVAL, "amdat_batch_use_settings: null batch");
    if (!s) {
        batch->forced = false;
        batch->force = JobChoice{};
        batch->job.reset();
        return AMDAT_OK;
    }
    auto r_ok = [](int R) { return R == 0 || R == 2 || R == 4 || R == 8; };
    if (!r_ok(s->sk_R) || !r_ok(s->fs_R))
        return set_error(AMDAT_EINVAL, "amdat_batch_use_settings: vectors per thread must be 2, 4 or 8");
    if (s->sk_plan_mode != 0 && s->sk_plan_mode != 1 && s->sk_plan_mode != 4 && s->sk_plan_mode != 5)
        return set_error(AMDAT_EINVAL, "amdat_batch_use_settings: S(k) plan mode must be 0, 1, 4 or 5");
    if (s->sk_plan_mode != 0 && s->sk_R != 4 && s->sk_R != 8)
        return set_error(AMDAT_EINVAL, "amdat_batch_use_settings: paired and grouped S(k) plans need R = 4 or 8");
    if (s->sk_plan_mode == 5 && s->sk_R != 8)
        return set_error(AMDAT_EINVAL, "amdat_batch_use_settings: S(k) plan mode 5 needs R = 8");
    if (s->sk_frames_per_item < 0 || s->fs_pairs_per_item < 0 || s->fs_atom_parts < 0 || s->fs_atom_parts > 64)
        return set_error(AMDAT_EINVAL, "amdat_batch_use_settings: frames, pairs >= 0 and atom parts in 0..64");
    JobChoice f;
    f.RS = s->sk_R;
    f.SM = s->sk_R ? s->sk_plan_mode : 0;
    f.FG = s->sk_frames_per_item;
    f.RF = s->fs_R;
    f.PC = s->fs_pairs_per_item;
    f.AF = std::max(1, s->fs_atom_parts);
    batch->forced = true;
    batch->force = f;
    batch->job.reset();
    return AMDAT_OK;
}

int amdat_batch_use_settings(amdat_batch *batch, const amdat_job_settings *s)
{
    AMDAT_GUARD("amdat_batch_use_settings", amdat_batch_use_settings_body(batch, s))
}

static int amdat_batch_job_settings_body(const amdat_batch *batch, amdat_job_settings *out)
{
    if (!batch || !batch->job) return set_error(AMDAT_EINVAL, "amdat_batch_job_settings: no job has run");
    if (!out) return set_error(AMDAT_EINVAL, "amdat_batch_job_settings: null argument");
    const JobCache &jc = *batch->job;
    out->sk_R = jc.prm.sk_sh ? jc.ch.RS : 0;
    out->sk_plan_mode = jc.prm.sk_sh ? jc.ch.SM : 0;
    out->sk_frames_per_item = jc.prm.sk_sh ? jc.ch.FG : 0;
    out->sk_frame_groups = jc.prm.sk_sh ? jc.G : 0;
    out->fs_R = jc.prm.fs_sh ? jc.ch.RF : 0;
    out->fs_pairs_per_item = jc.prm.fs_sh ? jc.ch.PC : 0;
    out->fs_atom_parts = jc.prm.fs_sh ? jc.ch.AF : 0;
    out->fs_pair_chunks = jc.prm.fs_sh ? jc.n_pc : 0;
    out->atoms_per_tile = table_atoms_per_tile(jc.MM);
    out->table_half_width = jc.MM;
    out->n_items = jc.job.n_items;
    out->grid = jc.L.grid;
    out->graph = jc.exec != nullptr;
    out->tuned = jc.tuned;
    out->model_ms = jc.model_ms;
    out->tuned_ms = jc.tuned_ms;
    return AMDAT_OK;
}

int amdat_batch_job_settings(const amdat_batch *batch, amdat_job_settings *out)
{
    AMDAT_GUARD("amdat_batch_job_settings", amdat_batch_job_settings_body(batch, out))
}

}  // extern "C"
