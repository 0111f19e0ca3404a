// phase_device.h -- device code shared by the phase kernels (kernels.hip,
// batch_kernels.hip): table geometry, table build, the tile loop and the
// per-atom phase-fraction functors.
//
// Phase kernels (rho, Fs), FP32-table form (AMDAT_PREC_FP32C):
//   * a block of 512 threads owns one chunk of KB = 512*R k-vectors (R per
//     thread, kept in registers as LDS addresses) and one frame (rho) or
//     time-origin pair (Fs), and streams the atoms of its split through LDS
//     in tiles of A atoms;
//   * per tile, every atom gets separable phase tables
//         e_c(m) = exp(2*pi*i * m * u_c),  u_c = frac(r_c / l),
//     for m in [-M, M] on x and y and [0, M] on z (vectors are stored with
//     nz >= 0; a vector with nz < 0 is flipped and its result conjugated).
//     u_c is an exact 32-bit fixed-point fraction formed in FP64 from the
//     FP32 coordinate (for Fs from the FP64 displacement rebuilt from image
//     counts), so m*u_c mod 1 is exact integer arithmetic.  Entries are
//     e(8j + d) = e(8j) * e(d) with both factors from sincos_turns(), so
//     every entry carries one FP32 complex product over exact anchors;
//   * the hot loop is per (atom, vector): three LDS reads, all with a
//     per-lane base and a compile-time offset (the tile loop is fully
//     unrolled), and two complex products,
//         exp(i k.r) = e_x(nx) e_y(ny) e_z(nz),
//     accumulated with FMA into an FP32 partial over 32 atoms (rho) or 64
//     (Fs); the partials are added into an FP64 sum per slot (one convert
//     and one FP64 add), which the block writes.
// The plain-FP64 form evaluates exp(i k.r) with FP64 sincos per update and
// accumulates in FP64; it is the reference the FP32 form is measured against.
#pragma once

#include "kernels.h"
#include "phase_math.h"

namespace amdat {
namespace {

constexpr int kWaves = kPhaseBlock / 64;
// Prefetch depth (updates in flight) of the hot loops.  (Fs with 6 or 8 in
// flight ran no faster on the 64,000-atom workload: 244 and 241 ms against
// 243 ms.)
constexpr int kDepth = 4;

// Table geometry for half-width MM: the sections are padded to whole
// groups of 8 entries, H = 8 floor(MM/8) + 7 >= MM, so the table build
// writes every entry of a group unconditionally.  Row of W float2 per atom:
// x at [0, 2H], y at [2H+1, 4H+1], z (m >= 0 only) at [4H+2, 5H+2].
// A atoms per tile: the reads of atom a sit a*W*8 bytes past the lane's
// base, an immediate offset while below 64 KiB.  A = 32 when at most the
// last two atoms of a tile fall past that reach (the compiler rebases their
// reads), else 16 when all of them fit, else 8.
template <int MM>
struct Geom {
    static constexpr int H = 8 * (MM / 8) + 7;
    // at least 257 float2 per row: reads of neighbouring atoms then never
    // pair into one ds_read2 (8-bit offsets), which keeps register use low
    static constexpr int W = (5 * H + 3) > 257 ? (5 * H + 3) : 257;
    static constexpr int OX = H;
    static constexpr int OY = 3 * H + 1;
    static constexpr int OZ = 4 * H + 2;
    static constexpr int A = (29 * W * 8 < 65536) ? 32 : ((15 * W * 8 < 65536) ? 16 : 8);
    static constexpr int LOG2A = A == 32 ? 5 : (A == 16 ? 4 : 3);
    static constexpr size_t tab_bytes = (size_t)A * W * sizeof(float2);
};

__device__ __forceinline__ void unpack_k(uint32_t kp, int &nx, int &ny, int &nz)
{
    nx = (int)(kp & 0xFFu) - 128;
    ny = (int)((kp >> 8) & 0xFFu) - 128;
    nz = (int)((kp >> 16) & 0xFFu);
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

// Per-chunk range descriptor (int4): x = lowest nx, y = highest nx,
// z = max |ny| | max nz << 16, w = thread units holding vectors (waves past
// them skip the hot loop) | group codes << 16: two bits per wave w at
// 16 + 2w, group size GS = 1 << code (slots g*GS .. g*GS + GS - 1 of a
// thread share the lead slot's e_x e_y, or e_x e_z).
__device__ __forceinline__ int range_my(int4 r) { return r.z & 0x7FFF; }
__device__ __forceinline__ int range_mz(int4 r) { return r.z >> 16; }
// z bit 15: waves of single slots hold R/2 vectors per thread and waves of
// pairs 3R/4 (plan mode 5); the slots past them are dead
__device__ __forceinline__ bool range_short(int4 r) { return (r.z >> 15) & 1; }
__device__ __forceinline__ int range_units(int4 r) { return r.w & 0xFFFF; }
__device__ __forceinline__ int range_gs_code(int4 r, int wave) { return (int)(((unsigned)r.w >> (16 + 2 * wave)) & 3u); }

// Anchor stage of a tile: the fixed-point fractions sU[3A] (row 3i+c =
// atom i, axis c) and the exact small anchors e(d), d < 8, into
// anc[d][3A]; atoms past the split's end get zero anchors, so every entry
// the group stage forms for them is zero and the hot loop adds nothing for
// them.  Row r is handled by LANES consecutive lanes, each converting the
// row's coordinate (loaded one tile ahead by fetch_tile) and computing PER
// anchors d = lane_in_row * PER + q.  Needs no barrier before it: it runs
// after the previous tile's group stage has read anc and sU.
// Lanes are dealt to rows axis-major (lane e: axis c = e / (LANES A), atom
// i), so the axis is uniform within a wave: the coordinate array is chosen
// by a scalar and the loads of a wave are contiguous.
// PER is chosen so that at most half the block's lanes (4 waves; 3 for
// every table width here) take part: a wave runs the row's conversion once
// whatever its lanes' d, so fewer lanes with more anchors each issue fewer
// operations, and the stage no longer puts two waves' work on one SIMD
// (waves w and w + 4 share one).  On the 64,000-atom workload: rho
// (A = 16, PER 1 -> 2) 379 -> 372 ms, Fs (A = 32, PER 2 -> 4) 230 -> 222 ms.
template <int MM>
struct AnchorLayout {
    static constexpr int NR = 3 * Geom<MM>::A;  // table rows of a tile
    // anchors per lane
    static constexpr int PER = (NR * 8 <= kPhaseBlock / 2) ? 1 : ((NR * 8 <= kPhaseBlock) ? 2 : 4);
    static constexpr int LANES = 8 / PER;  // lanes per row
    static constexpr int N = NR * LANES;   // lanes with work
    static_assert(N <= kPhaseBlock, "one anchor item per lane");
    static_assert(Geom<MM>::A % (64 / LANES) == 0, "one axis per wave");
};

// Axis (wave-uniform) and atom of anchor lane e.
template <int MM>
__device__ __forceinline__ void anchor_row(int e, int &c, int &i)
{
    using AL = AnchorLayout<MM>;
    const int rr = e / AL::LANES;
    c = __builtin_amdgcn_readfirstlane(rr / Geom<MM>::A);
    i = rr - c * Geom<MM>::A;
}

// Issue the loads of the tile starting at a0 (the lanes of the anchor
// stage, one row each); the values are consumed by tile_anchors after the
// current tile's hot loop, so their latency is hidden behind it.
template <int MM, typename UF>
__device__ __forceinline__ void fetch_tile(const UF &uf, int a0, int a_end, typename UF::Raw &raw)
{
    using AL = AnchorLayout<MM>;
    const int e = threadIdx.x;
    if (e >= AL::N) return;  // whole waves
    int c, i;
    anchor_row<MM>(e, c, i);
    if (a0 + i < a_end) {
        // one branch per axis (c is wave-uniform), each reading its fields
        // by name: selecting a field by c would index the functor in
        // memory, a per-lane stack copy
        if (c == 0) raw = uf.template load<0>(a0 + i);
        else if (c == 1) raw = uf.template load<1>(a0 + i);
        else raw = uf.template load<2>(a0 + i);
    }
}

template <int MM, typename UF>
__device__ __forceinline__ void tile_anchors(const UF &uf, int a0, int a_end, const typename UF::Raw &raw,
                                             float2 *anc, uint32_t *sU)
{
    using AL = AnchorLayout<MM>;
    const int e = threadIdx.x;
    if (e >= AL::N) return;  // whole waves
    int c, i;
    anchor_row<MM>(e, c, i);
    const int row = 3 * i + c, q0 = (e % AL::LANES) * AL::PER;
    const bool valid = a0 + i < a_end;
    uint32_t u = 0u;
    if (valid) u = c == 0 ? uf.template convert<0>(raw) : (c == 1 ? uf.template convert<1>(raw) : uf.template convert<2>(raw));
#pragma unroll
    for (int q = 0; q < AL::PER; ++q) {
        const float2 a = sincos_turns(u * (uint32_t)(q0 + q));
        anc[(q0 + q) * AL::NR + row] = valid ? a : make_float2(0.0f, 0.0f);
    }
    if (q0 == 0) sU[row] = u;
}

// Group stage of a tile, over the entries the chunk reads only: x in
// [xlo, xhi], y in [-my, my], z in [0, mz].  One work item per (axis, j,
// atom), atom fastest, computes the exact anchor e(8j) and the entries
// e(8j + d) = e(8j) e(d), writing all eight at +m and, on x and y, their
// conjugates at -m (the sections are padded to whole groups, so no store
// needs a range test).  Neighbouring lanes work on neighbouring atoms,
// whose rows start W entries apart, so their stores and anchor reads fall
// in distinct banks.
// The items of a thread are the same for every tile of the block's chunk.
// Item it = tid + 512 n covers every group of the chunk's sections within
// two rounds (n < 2, checked by the static_assert below); its word packs
// bits 0-13 the index (float2) of entry +8j in the tile table, i W +
// section offset + 8j, bits 14-20 the row 3i + c, bits 21-25 j and bit 26
// (conjugates written at -m, on x and y), or is all ones for no item.  The
// first round's word is formed once per block and kept; the second round,
// which only chunks with wide sections reach, is formed per tile.
constexpr uint32_t kNoItem = 0xFFFFFFFFu;

// Block-uniform section counts of a chunk: groups on x from jx0, then y, z.
struct BuildRange {
    int jx0, njx, njy, njz;
    __device__ __forceinline__ int n_groups() const { return njx + njy + njz; }
};

__device__ __forceinline__ BuildRange build_range(int4 rng)
{
    const int xlo = rng.x, xhi = rng.y, my = range_my(rng), mz = range_mz(rng);
    // |m| needed on x: [xa, xb]
    const int xa = xlo >= 0 ? xlo : (xhi <= 0 ? -xhi : 0);
    const int xb = max(abs(xlo), abs(xhi));
    BuildRange br;
    br.jx0 = xa >> 3;
    br.njx = (xb >> 3) - br.jx0 + 1;
    br.njy = (my >> 3) + 1;
    br.njz = (mz >> 3) + 1;
    return br;
}

template <int MM>
__device__ __forceinline__ uint32_t build_item(const BuildRange &br, int it)
{
    using Gm = Geom<MM>;
    static_assert(3 * (Gm::H / 8 + 1) * Gm::A <= 2 * kPhaseBlock, "two rounds of build items");
    static_assert(Gm::A * Gm::W < (1 << 14) && 3 * Gm::A < (1 << 7) && Gm::H / 8 < (1 << 5), "item packing");
    if (it >= br.n_groups() << Gm::LOG2A) return kNoItem;
    const int i = it & (Gm::A - 1), k = it >> Gm::LOG2A;
    const bool cx = k < br.njx, cy = k < br.njx + br.njy;
    const int j = cx ? br.jx0 + k : (cy ? k - br.njx : k - br.njx - br.njy);
    const int off = cx ? Gm::OX : (cy ? Gm::OY : Gm::OZ);
    const int c = cx ? 0 : (cy ? 1 : 2);
    return (uint32_t)(i * Gm::W + off + 8 * j) | ((uint32_t)(3 * i + c) << 14) | ((uint32_t)j << 21) |
           (cy ? (1u << 26) : 0u);
}

template <int MM>
__device__ __forceinline__ void build_one(float2 *tab, const float2 *anc, const uint32_t *sU, uint32_t w)
{
    constexpr int NR = 3 * Geom<MM>::A;  // table rows of a tile
    // an empty asm that "changes" w: the addresses below are unpacked each
    // tile instead of being hoisted out of the tile loop, where they would
    // hold eight more registers through the hot loop
    asm volatile("" : "+v"(w));
    const int up = (int)(w & 0x3FFFu), row = (int)((w >> 14) & 0x7Fu), j = (int)((w >> 21) & 0x1Fu);
    const float2 a8 = sincos_turns(sU[row] * (uint32_t)(8 * j));
    // entries m = 8j + d (d = 0..7) and, on x and y, -m: all inside the
    // section (8j + 7 <= H), each written by this lane only; e(0) is
    // written last so it keeps a +0 imaginary part
    float2 v[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) v[d] = cmul(a8, anc[d * NR + row]);
    if (w & (1u << 26)) {
        float2 *dn = tab + up - 16 * j - 7;
#pragma unroll
        for (int d = 0; d < 8; ++d) dn[7 - d] = make_float2(v[d].x, -v[d].y);
    }
    float2 *u = tab + up;
#pragma unroll
    for (int d = 0; d < 8; ++d) u[d] = v[d];
}

template <int MM>
__device__ __forceinline__ void build_groups(float2 *tab, const float2 *anc, const uint32_t *sU,
                                             const BuildRange &br, uint32_t item0)
{
    if (item0 != kNoItem) build_one<MM>(tab, anc, sU, item0);
    // second round: whole waves skip it unless the chunk's sections need it
    const int it1 = threadIdx.x + kPhaseBlock;
    if (__builtin_amdgcn_readfirstlane((int)(threadIdx.x & ~63u)) + kPhaseBlock < (br.n_groups() << Geom<MM>::LOG2A)) {
        const uint32_t w1 = build_item<MM>(br, it1);
        if (w1 != kNoItem) build_one<MM>(tab, anc, sU, w1);
    }
}

// One tile of the hot loop: A atoms x R slots, accumulated into the FP32
// partials p0 (Re, or Fs) and p1 (Im).  Software pipeline over the
// flattened (atom, slot) sequence: the reads of update u + kDepth, and the
// x read of the next atom, are issued before update u is computed.
// GS slots in a group (1, 2 or 4; R is a multiple of GS) share the factor
// behind py, so only the group's first slot reads it and forms
// q = e_x e_y.
// NS (<= R, a multiple of GS): slots that hold vectors; the rest of the R
// are dead and neither read nor summed.
template <int R, int MM, bool FS, int GS, int NS = R>
__device__ __forceinline__ void hot_tile(const float2 *px, const float2 *const (&py)[R], const float2 *const (&pz)[R],
                                         float (&p0)[R], float (&p1)[R])
{
    using Gm = Geom<MM>;
    static_assert(R % GS == 0 && NS % GS == 0 && NS <= R, "slots per thread must hold whole groups");
    float2 by[kDepth], bz[kDepth];
#pragma unroll
    for (int u = 0; u < kDepth; ++u) {
        if (u % GS == 0) by[u] = py[u % NS][(u / NS) * Gm::W];
        else by[u] = make_float2(0.0f, 0.0f);
        bz[u] = pz[u % NS][(u / NS) * Gm::W];
    }
    float2 ex = px[0], ex_next = ex;
    float qr = 0.0f, qi = 0.0f;
#pragma clang loop unroll(full)
    for (int u = 0; u < Gm::A * NS; ++u) {
        const int a = u / NS, r = u % NS, slot = u % kDepth;
        const bool lead = r % GS == 0;  // this update forms q = e_x e_y
        const float2 ey = by[slot], ez = bz[slot];
        if (u + kDepth < Gm::A * NS) {
            const int v = u + kDepth;
            if (v % GS == 0) by[slot] = py[v % NS][(v / NS) * Gm::W];
            bz[slot] = pz[v % NS][(v / NS) * Gm::W];
        }
        if (r == 0 && a + 1 < Gm::A) ex_next = px[(a + 1) * Gm::W];
        __builtin_amdgcn_sched_barrier(0);
        if (lead) {
            qr = fmaf(ex.x, ey.x, -ex.y * ey.y);
            qi = fmaf(ex.x, ey.y, ex.y * ey.x);
        }
        p0[r] = fmaf(qr, ez.x, fmaf(-qi, ez.y, p0[r]));
        if constexpr (!FS) p1[r] = fmaf(qr, ez.y, fmaf(qi, ez.x, p1[r]));
        __builtin_amdgcn_sched_barrier(0);
        if (r == NS - 1) ex = ex_next;
    }
}

// LDS read bases of a thread's slots in the tile table at tab: px the
// shared e_x(nx); py the factor a group shares (e_y, or e_z for a group
// sharing nz); pz the slot's own factor.
template <int R, int MM>
__device__ __forceinline__ void slot_pointers(const uint32_t *__restrict__ kp, const float2 *tab, const float2 *&px,
                                              const float2 *(&py)[R], const float2 *(&pz)[R])
{
    using Gm = Geom<MM>;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int nx, ny, nz;
        const uint32_t k = kp[r * kPhaseBlock + threadIdx.x];
        unpack_k(k, nx, ny, nz);
        if (r == 0) px = tab + Gm::OX + nx;
        const bool sw = (k >> 24) & 1u;
        py[r] = sw ? tab + Gm::OZ + nz : tab + Gm::OY + ny;
        pz[r] = sw ? tab + Gm::OY + ny : tab + Gm::OZ + nz;
    }
}

// kpack: (nx + 128) | (ny + 128) << 8 | nz << 16, bit 24 = roles of y and z
// swapped (see below).
// Hot loop over the atom tiles of one block, shared by the single-trajectory
// and batched phase kernels.  FS = false: rho (complex sums into h0/l0 = Re,
// h1/l1 = Im); FS = true: Fs (real part only, into h0/l0).  uf(a, c) gives
// the fixed-point fraction of list position a on axis c.  The R vectors of
// a thread share nx: e_x(nx) is read once per atom.  GROUPED plans: each
// wave's group code gives its group size GS; the GS slots of a group also
// share ny (or nz, with the roles of y and z swapped), so q = e_x e_y is
// read and formed once per group and each slot adds q e_z: rho costs
// 4 + 4/GS VALU per update (8 for GS = 1, 4.5 for GS = 8), Fs 2 + 4/GS.
// The tile loop of a wave whose groups hold GS slots (GS = 1 in plans
// without groups): every wave of the block runs the same tiles and meets
// the same barriers; only the hot loop differs with GS.
template <int R, int MM, bool FS, int GS, int NS, typename UF>
__device__ __forceinline__ void phase_tiles_gs(const uint32_t *__restrict__ kp, int4 rng, int a_begin, int a_end,
                                               float2 *tab, float2 *anc, uint32_t *sU, const UF &uf, bool active,
                                               double (&s0)[R], double (&s1)[R])
{
    using Gm = Geom<MM>;
    const float2 *px, *py[R], *pz[R];
    slot_pointers<R, MM>(kp, tab, px, py, pz);
#pragma unroll
    for (int r = 0; r < R; ++r) s0[r] = s1[r] = 0.0;
    // Tile pipeline, two barriers per tile: the coordinates of tile t+1 are
    // loaded before tile t's hot loop and its anchors computed after it, so
    // neither the load latency nor the anchor stage waits at a barrier.
    const BuildRange br = build_range(rng);
    const uint32_t item0 = build_item<MM>(br, threadIdx.x);
    typename UF::Raw raw{};
    fetch_tile<MM>(uf, a_begin, a_end, raw);
    tile_anchors<MM>(uf, a_begin, a_end, raw, anc, sU);
    if (a_begin + Gm::A < a_end) fetch_tile<MM>(uf, a_begin + Gm::A, a_end, raw);
    // Groups of FT tiles; the FP32 partials of a group are added into the
    // FP64 sums once, after its last tile: every 32 atoms for rho
    // (per-vector sums are outputs), every 64 for Fs (averaged over pairs
    // and shells).
    constexpr int FT = (FS ? 64 : 32) / Gm::A > 0 ? (FS ? 64 : 32) / Gm::A : 1;
    for (int g0 = a_begin; g0 < a_end; g0 += FT * Gm::A) {
        float p0[R], p1[R];
#pragma unroll
        for (int r = 0; r < R; ++r) p0[r] = p1[r] = 0.0f;
        const int g_end = min(a_end, g0 + FT * Gm::A);
        for (int a0 = g0; a0 < g_end; a0 += Gm::A) {
            __syncthreads();  // anchors of tile a0 written; the previous tile's reads of tab done
            build_groups<MM>(tab, anc, sU, br, item0);
            __syncthreads();  // tables of tile a0 complete
            if (active) hot_tile<R, MM, FS, GS, NS>(px, py, pz, p0, p1);
            const int an = a0 + Gm::A;  // block-uniform
            if (an < a_end) {
                tile_anchors<MM>(uf, an, a_end, raw, anc, sU);
                if (an + Gm::A < a_end) fetch_tile<MM>(uf, an + Gm::A, a_end, raw);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            s0[r] += (double)p0[r];
            if constexpr (!FS) s1[r] += (double)p1[r];
        }
    }
}

// Hot loop over the atom tiles of one block (see phase_tiles_gs), with the
// wave's group size chosen once, outside the tile loop: the partial sums
// then stay in the same registers from tile to tile.
template <int R, int MM, bool FS, bool GROUPED, typename UF>
__device__ __forceinline__ void phase_tiles(const uint32_t *__restrict__ kp, int4 rng, int a_begin, int a_end,
                                            float2 *tab, float2 *anc, uint32_t *sU, const UF uf,
                                            double (&s0)[R], double (&s1)[R])
{
    // waves whose units all lie past the chunk's last real unit skip the
    // sums; grouped plans give each wave its group size (wave-uniform)
    const int wave0 = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * 64;
    const bool active = wave0 < range_units(rng);
    if constexpr (GROUPED) {
        const int gs_code = range_gs_code(rng, wave0 >> 6);
        if constexpr (R >= 8) {
            if (gs_code == 3) {
                phase_tiles_gs<R, MM, FS, 8, R>(kp, rng, a_begin, a_end, tab, anc, sU, uf, active, s0, s1);
                return;
            }
        }
        if constexpr (R >= 4) {
            if (gs_code == 2) {
                phase_tiles_gs<R, MM, FS, 4, R>(kp, rng, a_begin, a_end, tab, anc, sU, uf, active, s0, s1);
                return;
            }
        }
        if constexpr (R == 8) {
            // plan mode 5: singles and pairs in waves of fewer slots
            if (range_short(rng)) {
                if (gs_code == 1)
                    phase_tiles_gs<R, MM, FS, 2, 3 * R / 4>(kp, rng, a_begin, a_end, tab, anc, sU, uf, active, s0, s1);
                else
                    phase_tiles_gs<R, MM, FS, 1, R / 2>(kp, rng, a_begin, a_end, tab, anc, sU, uf, active, s0, s1);
                return;
            }
        }
        if (gs_code == 1) {
            phase_tiles_gs<R, MM, FS, 2, R>(kp, rng, a_begin, a_end, tab, anc, sU, uf, active, s0, s1);
            return;
        }
    }
    phase_tiles_gs<R, MM, FS, 1, R>(kp, rng, a_begin, a_end, tab, anc, sU, uf, active, s0, s1);
}

// Fixed-point fraction of a coordinate (rho): load<C>() fetches the FP32
// coordinate of list position a on axis C, convert<C>() forms the fraction.
struct CoordFrac {
    const float *X, *Y, *Z;
    const int *list;
    size_t foff;
    double inv_ell;
    struct Raw {
        float v;
    };
    template <int C>
    __device__ __forceinline__ Raw load(int a) const
    {
        const float *P = C == 0 ? X : (C == 1 ? Y : Z);
        return Raw{P[foff + (list ? list[a] : a)]};
    }
    template <int C>
    __device__ __forceinline__ uint32_t convert(const Raw &r) const
    {
        return turns_fixed((double)r.v * inv_ell);
    }
};

// Fixed-point fraction of a displacement dr = (x1 + i1*L1) - (x0 + i0*L0),
// rebuilt exactly in FP64 (Fs); the pair's box lengths are held by value
// (an aggregate, so that the functor lives in registers).
struct DispFrac {
    const float *X, *Y, *Z;
    const int16_t *IX, *IY, *IZ;
    const int *list;
    size_t o0, o1;
    double inv_ell;
    float l0x, l0y, l0z, l1x, l1y, l1z;
    struct Raw {
        float x0, x1;
        uint32_t i01;  // i0 in the low 16 bits, i1 in the high
    };
    template <int C>
    __device__ __forceinline__ Raw load(int a) const
    {
        const float *P = C == 0 ? X : (C == 1 ? Y : Z);
        const int16_t *I = C == 0 ? IX : (C == 1 ? IY : IZ);
        const int j = list ? list[a] : a;
        const size_t p0 = o0 + j, p1 = o1 + j;
        return Raw{P[p0], P[p1], (uint32_t)(uint16_t)I[p0] | ((uint32_t)(uint16_t)I[p1] << 16)};
    }
    template <int C>
    __device__ __forceinline__ uint32_t convert(const Raw &r) const
    {
        const float l0 = C == 0 ? l0x : (C == 1 ? l0y : l0z), l1 = C == 0 ? l1x : (C == 1 ? l1y : l1z);
        const int i0 = (int16_t)(r.i01 & 0xFFFFu), i1 = (int16_t)(r.i01 >> 16);
        const double d = ((double)r.x1 - (double)r.x0) + ((double)i1 * (double)l1 - (double)i0 * (double)l0);
        return turns_fixed(d * inv_ell);
    }
};

// rho over one block's atoms: FP64 sums (re, im) per slot.
template <int R, int MM, bool GROUPED = false>
__device__ __forceinline__ void rho_block(const float *__restrict__ X, const float *__restrict__ Y,
                                          const float *__restrict__ Z, const int *__restrict__ list, size_t foff,
                                          int a_begin, int a_end, double inv_ell, const uint32_t *__restrict__ kp,
                                          int4 rng, float2 *tab, float2 *anc, uint32_t *sU,
                                          double (&re)[R], double (&im)[R])
{
    const CoordFrac uf{X, Y, Z, list, foff, inv_ell};
    phase_tiles<R, MM, false, GROUPED>(kp, rng, a_begin, a_end, tab, anc, sU, uf, re, im);
}

template <int R>
__device__ __forceinline__ void rho_write(const int *__restrict__ kd, const double (&re)[R], const double (&im)[R],
                                          double2 *__restrict__ dst)
{
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int d = kd[r * kPhaseBlock + threadIdx.x];
        if (d >= 0) dst[d & ~(1 << 30)] = make_double2(re[r], (d & (1 << 30)) ? -im[r] : im[r]);
    }
}

// Fs over one block's atoms for the pair (o0, o1): FP64 sum of cos(k.dr).
template <int R, int MM, bool GROUPED = false>
__device__ __forceinline__ void fs_block(const float *__restrict__ X, const float *__restrict__ Y,
                                         const float *__restrict__ Z, const int16_t *__restrict__ IX,
                                         const int16_t *__restrict__ IY, const int16_t *__restrict__ IZ,
                                         const float *__restrict__ L0, const float *__restrict__ L1,
                                         const int *__restrict__ list, size_t o0, size_t o1, int a_begin, int a_end,
                                         double inv_ell, const uint32_t *__restrict__ kp, int4 rng, float2 *tab,
                                         float2 *anc, uint32_t *sU, double (&s)[R])
{
    const DispFrac uf{X, Y, Z, IX, IY, IZ, list, o0, o1, inv_ell, L0[0], L0[1], L0[2], L1[0], L1[1], L1[2]};
    double unused[R];
    phase_tiles<R, MM, true, GROUPED>(kp, rng, a_begin, a_end, tab, anc, sU, uf, s, unused);
}

template <int R>
__device__ __forceinline__ void fs_write(const int *__restrict__ kd, const double (&s)[R], double *__restrict__ dst)
{
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int d = kd[r * kPhaseBlock + threadIdx.x];
        if (d >= 0) dst[d & ~(1 << 30)] = s[r];
    }
}

// Blocks of a launch are dealt to the 8 XCDs round robin by linear block
// id, and each XCD has its own L2.  The phase kernels number their work
// items (chunk, row), row = frame or pair + n_rows * split, so that all
// chunks of one row land on one XCD one after another: the row's
// coordinates come from HBM once and from that XCD's L2 for its other
// chunks.  Grid: kXcds * n_chunks * ceil(n_rows_total / kXcds) blocks;
// false for the ids past the last row (the whole block returns).
constexpr int kXcds = 8;
__device__ __forceinline__ bool xcd_item(int n_chunks, int n_rows_total, int &chunk, int &row)
{
    const int L = blockIdx.x;
    const int x = L % kXcds, k = L / kXcds;
    chunk = k % n_chunks;
    row = (k / n_chunks) * kXcds + x;
    return row < n_rows_total;
}

// Pair kernels (MSD, displacement statistics): atom block bx of pair p
// runs on XCD bx % kXcds, each XCD taking its atom blocks pair after pair,
// so the frame slices that nearby pairs share come from its L2.  Grid:
// kXcds * ceil(n_blocks / kXcds) * n_pairs blocks; false for the padding ids.
__device__ __forceinline__ bool xcd_atom_block(int n_blocks, int n_pairs, int &bx, int &p)
{
    const int L = blockIdx.x;
    const int x = L % kXcds, k = L / kXcds;
    const int nb8 = (n_blocks + kXcds - 1) / kXcds;
    bx = (k % nb8) * kXcds + x;
    p = k / nb8;
    return bx < n_blocks && p < n_pairs;
}

#define AMDAT_PHASE_SHARED(MM)                         \
    extern __shared__ float2 tab[];                    \
    __shared__ float2 anc[3 * Geom<MM>::A * 8];          \
    __shared__ uint32_t sU[3 * Geom<MM>::A]


// A work item read by every lane of a wave (the same address), kept in
// scalar registers.
__device__ __forceinline__ int4 uniform_item(const int4 *__restrict__ items, int q)
{
    const int4 it = items[q];
    return make_int4(__builtin_amdgcn_readfirstlane(it.x), __builtin_amdgcn_readfirstlane(it.y),
                     __builtin_amdgcn_readfirstlane(it.z), __builtin_amdgcn_readfirstlane(it.w));
}

// Next work item of a persistent launch: thread 0 draws from the queue
// counter and the block reads the value from LDS.
__device__ __forceinline__ int next_item(int *counter, int *slot)
{
    __syncthreads();  // every thread has read the previous value, and the tables are free
    if (threadIdx.x == 0) *slot = atomicAdd(counter, 1);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*slot);  // same value in every lane: keep it scalar
}

// Sum over a block of kBlock threads (tree order, deterministic).
__device__ __forceinline__ double block_sum(double v, double *red)
{
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

}  // namespace
}  // namespace amdat
