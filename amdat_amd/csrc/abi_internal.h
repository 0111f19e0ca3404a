// abi_internal.h -- shared internals of the C ABI implementation.
#pragma once

#include <stdint.h>

#include <exception>
#include <new>
#include <string>
#include <vector>

namespace amdat {

// Record a message for amdat_last_error() on this thread and return code.
int set_error(int code, const std::string &msg);
void clear_error();

// Body of an entry point of the C ABI: an exception thrown below it (a
// failed host allocation, a thread that cannot start) becomes a status
// instead of unwinding through the C caller's frames.
#define AMDAT_GUARD(NAME, CALL)                                                          \
    try {                                                                                \
        return CALL;                                                                     \
    } catch (const std::bad_alloc &) {                                                   \
        return ::amdat::set_error(AMDAT_ENOMEM, NAME ": host allocation failed");        \
    } catch (const std::exception &amdat_guard_e) {                                      \
        return ::amdat::set_error(AMDAT_EINVAL, std::string(NAME ": ") + amdat_guard_e.what()); \
    } catch (...) {                                                                      \
        return ::amdat::set_error(AMDAT_EINVAL, NAME ": unexpected error");              \
    }

// Block layout of a k-vector plan (host side of the phase kernels).
// Vectors are canonicalised (a vector with nz < 0 is flipped and flagged
// with bit 30 of its destination), sorted by (nx, ny, nz), and dealt into
// thread units of R slots that all share one nx (a thread then reads
// e_x(nx) once per atom for its R vectors).  Units of one nx take the
// group's sorted vectors with stride (units in the group), so the lanes of
// a wave read neighbouring table entries; each unit's slots are then
// permuted so the 32 lanes of a half-wave read y and z entries in distinct
// LDS bank pairs where possible.  Chunks are 512 consecutive units; unused
// slots carry (nx of their unit, 0, 0) and destination -1.
// Table geometry shared with the device code (Geom<MM> in phase_device.h):
// the half-width MM serving a largest |component| M, and the entry offsets
// of the y and z sections in a table row.
constexpr int table_half_width(int M)
{
    return M <= 24 ? 24 : (M <= 50 ? 50 : (M <= 100 ? 100 : (M <= 127 ? 127 : -1)));
}
// Sections hold H = 8 floor(MM/8) + 7 >= MM entries on each side.
constexpr int table_h(int MM) { return 8 * (MM / 8) + 7; }
constexpr int table_oy(int MM) { return 3 * table_h(MM) + 1; }
constexpr int table_oz(int MM) { return 4 * table_h(MM) + 2; }

// kpack bit 24: the slot's roles of y and z are swapped (a paired slot
// whose pair shares nz rather than ny).
constexpr uint32_t kPackSwap = 1u << 24;

struct PlanLayout {
    int R = 0, KB = 0, n_chunks = 0;
    std::vector<uint32_t> pack;  // [n_chunks][R][512]: (nx+128) | (ny+128)<<8 | nz<<16
    std::vector<int> dst;        // same layout: destination index | bit 30 (conjugate), or -1
    std::vector<int> range;      // [n_chunks][4]: nx lo, nx hi, max |ny| | max nz << 16,
                                 // units with vectors | group codes << 16 (2 bits per wave:
                                 // group size 1 << code)
    long conflicts = 0;          // extra LDS passes summed over (chunk, slot, half-wave, axis)
    bool pair = false;           // grouped waves: a group's slots share ny, or nz (kPackSwap)
    bool quad = false;           // mode 3
    long gs8_slots = 0, quad_slots = 0, paired_slots = 0, unpaired_slots = 0;
    int active_waves = 0;        // waves that run the sums, summed over chunks
    double simd_cost = 0.0;      // per atom, summed over chunks: 4 SIMDs x 64 lanes x the busiest
                                 // SIMD's VALU ops (its two waves' slots x (c - 4 + 4/GS))
                                 // slots of waves that run the sums, by group size 8, 4, 2, 1,
                                 // each chunk's rounded up to whole SIMD rounds (4 waves)
};
// mode 0: unpaired; 1: paired, vectors without a partner padded; 2: paired,
// vectors without a partner in unpaired waves; 3: quads sharing ny where a
// ny class holds four, then pairs as mode 1, in waves of one kind each;
// 5: grouped as 4, with waves of single slots holding R/2 vectors per
// thread and waves of pairs 3R/4 (R = 8; dead slots past them, range bit
// 15 set), so that waves carry about equal work per atom;
// 4: grouped, classes sharing nz or ny cut into groups of 8, 4, 2 and 1
// slots (cheapest at c VALU ops per unpaired slot: one accumulate of c - 4
// per slot and one product of 4 per group), waves of one group size each,
// slots, groups and units placed by a local search against LDS bank-pair
// clashes (place = false skips that search: the cost figures are the
// same).  MM: table half-width of the kernels that read the plan.
PlanLayout plan_layout(const int *nvec, int K, int R, int mode, int MM, double c = 9.0, bool place = true);

// ---- trajectory ingestion (host_io.cpp) ----
struct MappedFile {
    const char *data = nullptr;
    size_t size = 0;
    int fd = -1;
    int open(const char *path);
    ~MappedFile();
};
int io_threads(int threads);
// frame start offsets of a custom dump / xyz file (the first `want` frames)
int index_custom(const MappedFile &f, int want, int threads, std::vector<size_t> &off);
int index_xyz(const MappedFile &f, int N, int want, int threads, std::vector<size_t> &off);
// frames [f0, f1) into rows 0 .. f1-f0 of the outputs
int parse_custom(const MappedFile &f, const std::vector<size_t> &off, int f0, int f1, int N, int threads,
                 float *x, float *y, float *z, int16_t *ix, int16_t *iy, int16_t *iz, float *box, float *lo,
                 int *types);
int parse_xyz(const MappedFile &f, const std::vector<size_t> &off, int f0, int f1, int N, int threads, float *x,
              float *y, float *z, int *types);
// image counts from consecutive-frame jumps; prev_xyz / prev_img ([3][N])
// carry the frame before row 0 when a trajectory is read in pieces
int unwrap_images(int N, int T, const float *L, const float *x, const float *y, const float *z, int16_t *ix,
                  int16_t *iy, int16_t *iz, const float *prev_xyz, const int16_t *prev_img, int threads);

}  // namespace amdat

extern "C" {
// Test hook: the layout of plan_layout for K vectors nvec[K][3] (pair = the
// layout mode; paired slots 2p and 2p+1 share ny, or nz where kPackSwap is
// set).
// With
// pack == NULL only *n_chunks is set.  Returns 0 or a negative status.
int amdat_internal_plan_layout(const int *nvec, int K, int R, int pair, uint32_t *pack, int *dst, int *range,
                               int *n_chunks, long *conflicts);
}
