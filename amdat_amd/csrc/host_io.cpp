// host_io.cpp -- trajectory ingestion: LAMMPS custom dumps and xyz files
// parsed in parallel into frame-major SoA (FP32 wrapped coordinates, int16
// image counts, per-frame box), as the device path stores them.
//
// The file is memory-mapped; frame starts are found by a parallel scan,
// then whole frames are parsed by worker threads.  Numbers are read with
// std::from_chars, which rounds correctly, so every float equals what the
// reference's stream extraction (operator>> into float) gives.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <charconv>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/amdat_hip.h"
#include "abi_internal.h"

namespace amdat {

MappedFile::~MappedFile()
{
    if (data && size) munmap((void *)data, size);
    if (fd >= 0) close(fd);
}

int MappedFile::open(const char *path)
{
    fd = ::open(path, O_RDONLY);
    if (fd < 0) return set_error(AMDAT_EIO, std::string("cannot open trajectory file ") + path);
    struct stat st;
    if (fstat(fd, &st) != 0) return set_error(AMDAT_EIO, std::string("cannot stat ") + path);
    size = (size_t)st.st_size;
    if (size == 0) return set_error(AMDAT_EIO, std::string("empty trajectory file ") + path);
    void *p = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (p == MAP_FAILED) {
        size = 0;
        return set_error(AMDAT_EIO, std::string("cannot map ") + path);
    }
    madvise(p, size, MADV_SEQUENTIAL);
    data = (const char *)p;
    return AMDAT_OK;
}

int io_threads(int threads)
{
    if (threads > 0) return threads;
    const unsigned h = std::thread::hardware_concurrency();
    return h ? (int)std::min(h, 64u) : 4;
}

namespace {

// Run f(i) for i in [0, n) on `threads` workers drawing indices from a counter.
template <typename F>
void parallel_for(int n, int threads, F f)
{
    threads = std::max(1, std::min(threads, n));
    if (threads == 1) {
        for (int i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<int> next{0};
    auto worker = [&] {
        for (int i = next.fetch_add(1); i < n; i = next.fetch_add(1)) f(i);
    };
    // the calling thread works too; a thread that cannot be started leaves
    // its share to the others (the indices are drawn from one counter)
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; t++) {
        try {
            pool.emplace_back(worker);
        } catch (...) {
            break;
        }
    }
    worker();
    for (auto &th : pool) th.join();
}

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r'; }

inline const char *line_end(const char *s, const char *e)
{
    const char *nl = (const char *)std::memchr(s, '\n', (size_t)(e - s));
    return nl ? nl : e;
}

// Split [s, e) into whitespace-separated tokens (at most cap).
inline int tokens(const char *s, const char *e, const char **tb, const char **te, int cap)
{
    int n = 0;
    while (s < e && n < cap) {
        while (s < e && is_space(*s)) ++s;
        if (s >= e) break;
        tb[n] = s;
        while (s < e && !is_space(*s)) ++s;
        te[n++] = s;
    }
    return n;
}

template <typename T>
inline bool num(const char *b, const char *e, T &v)
{
    if (b < e && *b == '+') ++b;
    auto r = std::from_chars(b, e, v);
    return r.ec == std::errc() && r.ptr == e;
}

inline bool starts(const char *s, const char *e, const char *lit)
{
    const size_t n = std::strlen(lit);
    return (size_t)(e - s) >= n && std::memcmp(s, lit, n) == 0;
}

}  // namespace

// ---- LAMMPS custom ---------------------------------------------------------

int index_custom(const MappedFile &f, int want, int threads, std::vector<size_t> &off)
{
    // frame starts: "ITEM: TIMESTEP" at the start of a line
    static const char key[] = "ITEM: TIMESTEP";
    const size_t klen = sizeof(key) - 1;
    const int P = io_threads(threads);
    const size_t step = f.size / P + 1;
    std::vector<std::vector<size_t>> found(P);
    parallel_for(P, P, [&](int p) {
        const size_t b = std::min(f.size, (size_t)p * step), e = std::min(f.size, b + step);
        const char *s = f.data + b, *end = f.data + e, *fend = f.data + f.size;
        while (s < end) {
            const char *hit = (const char *)std::memchr(s, 'I', (size_t)(end - s));
            if (!hit) break;
            if ((hit == f.data || hit[-1] == '\n') && (size_t)(fend - hit) >= klen && std::memcmp(hit, key, klen) == 0)
                found[p].push_back((size_t)(hit - f.data));
            s = hit + 1;
        }
    });
    off.clear();
    for (auto &v : found) off.insert(off.end(), v.begin(), v.end());
    if ((int)off.size() < want)
        return set_error(AMDAT_EIO, "trajectory file holds " + std::to_string(off.size()) + " frames, " +
                                        std::to_string(want) + " needed");
    off.resize(want);
    return AMDAT_OK;
}

// Parse frames [f0, f1) of a custom dump; outputs are indexed from frame f0
// (row t - f0).  Columns: x y z with ix iy iz, or xu yu zu (then wrapped
// coordinates and images follow from the box: k = floor((u - lo)/L)), or
// x y z alone (images 0).  types (frame f0, optional): atom types in file order.
int parse_custom(const MappedFile &f, const std::vector<size_t> &off, int f0, int f1, int N, int threads,
                 float *x, float *y, float *z, int16_t *ix, int16_t *iy, int16_t *iz, float *box, float *lo_out,
                 int *types)
{
    std::atomic<int> bad{0};
    std::vector<std::string> msg(f1 - f0);
    parallel_for(f1 - f0, io_threads(threads), [&](int r) {
        const int t = f0 + r;
        const char *s = f.data + off[t];
        const char *fend = f.data + f.size;
        auto fail = [&](const std::string &m) {
            msg[r] = "frame " + std::to_string(t) + ": " + m;
            bad.store(1);
        };
        int natoms = -1;
        float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        int cx = -1, cy = -1, cz = -1, cix = -1, ciy = -1, ciz = -1, cty = -1, ncols = 0;
        bool unwrapped = false;
        // header
        for (;;) {
            if (s >= fend) return fail("truncated header");
            const char *le = line_end(s, fend);
            if (starts(s, le, "ITEM: NUMBER OF ATOMS")) {
                s = le + 1;
                const char *l2 = line_end(s, fend);
                const char *tb[2], *te[2];
                if (tokens(s, l2, tb, te, 2) < 1 || !num(tb[0], te[0], natoms)) return fail("bad atom count");
                s = l2 + 1;
            } else if (starts(s, le, "ITEM: BOX BOUNDS")) {
                s = le + 1;
                for (int c = 0; c < 3; c++) {
                    const char *l2 = line_end(s, fend);
                    const char *tb[3], *te[3];
                    if (tokens(s, l2, tb, te, 3) < 2 || !num(tb[0], te[0], lo[c]) || !num(tb[1], te[1], hi[c]))
                        return fail("bad box bounds");
                    s = l2 + 1;
                }
            } else if (starts(s, le, "ITEM: ATOMS")) {
                const char *tb[64], *te[64];
                const int nt = tokens(s, le, tb, te, 64);
                for (int k = 2; k < nt; k++) {
                    const std::string name(tb[k], te[k]);
                    const int c = k - 2;
                    if (name == "x") cx = c;
                    else if (name == "y") cy = c;
                    else if (name == "z") cz = c;
                    else if (name == "xu") { cx = c; unwrapped = true; }
                    else if (name == "yu") cy = c;
                    else if (name == "zu") cz = c;
                    else if (name == "ix") cix = c;
                    else if (name == "iy") ciy = c;
                    else if (name == "iz") ciz = c;
                    else if (name == "type") cty = c;
                }
                ncols = nt - 2;
                s = le + 1;
                break;
            } else {
                s = le + 1;
            }
        }
        if (natoms != N) return fail("atom count " + std::to_string(natoms) + " != " + std::to_string(N));
        if (cx < 0 || cy < 0 || cz < 0) return fail("needs x y z (or xu yu zu) columns");
        const bool images = cix >= 0 && ciy >= 0 && ciz >= 0 && !unwrapped;
        float L[3];
        for (int c = 0; c < 3; c++) {
            L[c] = hi[c] - lo[c];
            box[3 * r + c] = L[c];
            if (lo_out) lo_out[3 * r + c] = lo[c];
        }
        const size_t row = (size_t)r * N;
        const int need = std::max({cx, cy, cz, cix, ciy, ciz, cty}) + 1;
        for (int j = 0; j < N; j++) {
            if (s >= fend) return fail("truncated at atom " + std::to_string(j));
            const char *le = line_end(s, fend);
            const char *tb[64], *te[64];
            const int nt = tokens(s, le, tb, te, std::min(64, std::max(need, ncols)));
            if (nt < need) return fail("short atom line " + std::to_string(j));
            float w[3];
            if (!num(tb[cx], te[cx], w[0]) || !num(tb[cy], te[cy], w[1]) || !num(tb[cz], te[cz], w[2]))
                return fail("bad coordinate at atom " + std::to_string(j));
            int im[3] = {0, 0, 0};
            if (images) {
                if (!num(tb[cix], te[cix], im[0]) || !num(tb[ciy], te[ciy], im[1]) || !num(tb[ciz], te[ciz], im[2]))
                    return fail("bad image count at atom " + std::to_string(j));
            } else if (unwrapped) {
                for (int c = 0; c < 3; c++) {
                    const double u = w[c];
                    const double k = std::floor((u - lo[c]) / L[c]);
                    w[c] = (float)(u - k * L[c]);
                    im[c] = (int)k;
                }
            }
            for (int c = 0; c < 3; c++)
                if (im[c] < -32768 || im[c] > 32767) return fail("image count beyond int16");
            x[row + j] = w[0]; y[row + j] = w[1]; z[row + j] = w[2];
            ix[row + j] = (int16_t)im[0]; iy[row + j] = (int16_t)im[1]; iz[row + j] = (int16_t)im[2];
            if (types && r == 0 && cty >= 0) {
                int ty = 0;
                if (!num(tb[cty], te[cty], ty)) return fail("bad type at atom " + std::to_string(j));
                types[j] = ty;
            }
            s = le + 1;
        }
    });
    if (bad.load()) {
        for (auto &m : msg)
            if (!m.empty()) return set_error(AMDAT_EIO, m);
    }
    return AMDAT_OK;
}

// ---- xyz -------------------------------------------------------------------

// Frame t starts at line t*(N+2): a count line, a comment line, N atoms.
int index_xyz(const MappedFile &f, int N, int want, int threads, std::vector<size_t> &off)
{
    const int P = io_threads(threads);
    const size_t step = f.size / P + 1;
    std::vector<size_t> nl(P, 0);
    parallel_for(P, P, [&](int p) {
        const size_t b = std::min(f.size, (size_t)p * step), e = std::min(f.size, b + step);
        size_t c = 0;
        for (const char *s = f.data + b, *end = f.data + e; s < end;) {
            const char *h = (const char *)std::memchr(s, '\n', (size_t)(end - s));
            if (!h) break;
            ++c;
            s = h + 1;
        }
        nl[p] = c;
    });
    // line L starts after the L-th newline; the lines wanted are t*(N+2)
    const size_t per = (size_t)N + 2;
    off.assign(want, (size_t)-1);
    if (want > 0) off[0] = 0;
    std::vector<size_t> first(P, 0);
    for (int p = 1; p < P; p++) first[p] = first[p - 1] + nl[p - 1];
    parallel_for(P, P, [&](int p) {
        const size_t b = std::min(f.size, (size_t)p * step), e = std::min(f.size, b + step);
        size_t line = first[p];  // newlines before this range
        for (const char *s = f.data + b, *end = f.data + e; s < end;) {
            const char *h = (const char *)std::memchr(s, '\n', (size_t)(end - s));
            if (!h) break;
            ++line;  // a new line starts at h + 1
            if (line % per == 0) {
                const size_t t = line / per;
                if (t < (size_t)want && (size_t)(h + 1 - f.data) < f.size) off[t] = (size_t)(h + 1 - f.data);
            }
            s = h + 1;
        }
    });
    for (int t = 0; t < want; t++)
        if (off[t] == (size_t)-1)
            return set_error(AMDAT_EIO, "xyz file holds fewer than " + std::to_string(want) + " frames");
    return AMDAT_OK;
}

int parse_xyz(const MappedFile &f, const std::vector<size_t> &off, int f0, int f1, int N, int threads, float *x,
              float *y, float *z, int *types)
{
    std::atomic<int> bad{0};
    std::vector<std::string> msg(f1 - f0);
    parallel_for(f1 - f0, io_threads(threads), [&](int r) {
        const int t = f0 + r;
        const char *s = f.data + off[t], *fend = f.data + f.size;
        auto fail = [&](const std::string &m) {
            msg[r] = "xyz frame " + std::to_string(t) + ": " + m;
            bad.store(1);
        };
        const char *le = line_end(s, fend);
        const char *tb[8], *te[8];
        int n = -1;
        if (tokens(s, le, tb, te, 1) < 1 || !num(tb[0], te[0], n)) return fail("bad atom count line");
        if (n != N) return fail("The number of atoms listed in the xyz file is inconsistent with user input: " +
                                std::to_string(n) + " != " + std::to_string(N));
        s = line_end(le + 1, fend) + 1;  // skip the comment line
        const size_t row = (size_t)r * N;
        for (int j = 0; j < N; j++) {
            if (s >= fend) return fail("truncated at atom " + std::to_string(j));
            le = line_end(s, fend);
            if (tokens(s, le, tb, te, 4) < 4) return fail("short atom line " + std::to_string(j));
            float w[3];
            if (!num(tb[1], te[1], w[0]) || !num(tb[2], te[2], w[1]) || !num(tb[3], te[3], w[2]))
                return fail("bad coordinate at atom " + std::to_string(j));
            x[row + j] = w[0]; y[row + j] = w[1]; z[row + j] = w[2];
            if (types && r == 0) types[j] = std::atoi(std::string(tb[0], te[0]).c_str());
            s = le + 1;
        }
    });
    if (bad.load()) {
        for (auto &m : msg)
            if (!m.empty()) return set_error(AMDAT_EIO, m);
    }
    return AMDAT_OK;
}

// Image counts of an xyz trajectory by the reference's unwrap rule
// (src/trajectory.cpp:233-270): between consecutive frames the FP32 jump
// d = r(t) - r(t-1) moves the particle by -L * int(d / (0.55 L)), so
// image(t) = image(t-1) - int(d / (0.55 L)).  Rows [0, T) of the arrays;
// prev (optional) holds the frame before row 0 and its images.
int unwrap_images(int N, int T, const float *L, const float *x, const float *y, const float *z, int16_t *ix,
                  int16_t *iy, int16_t *iz, const float *prev_xyz, const int16_t *prev_img, int threads)
{
    const float s[3] = {L[0] * 0.55f, L[1] * 0.55f, L[2] * 0.55f};
    std::atomic<int> bad{0};
    const int P = io_threads(threads);
    const int per = (N + P - 1) / P;
    parallel_for(P, P, [&](int p) {
        const int j0 = p * per, j1 = std::min(N, j0 + per);
        const float *src[3] = {x, y, z};
        int16_t *dst[3] = {ix, iy, iz};
        for (int j = j0; j < j1; j++)
            for (int c = 0; c < 3; c++) {
                float last = prev_xyz ? prev_xyz[(size_t)c * N + j] : src[c][j];
                int img = prev_img ? prev_img[(size_t)c * N + j] : 0;
                for (int t = 0; t < T; t++) {
                    const float cur = src[c][(size_t)t * N + j];
                    if (t > 0 || prev_xyz) img -= (int)((cur - last) / s[c]);
                    if (img < -32768 || img > 32767) { bad.store(1); return; }
                    dst[c][(size_t)t * N + j] = (int16_t)img;
                    last = cur;
                }
            }
    });
    if (bad.load()) return set_error(AMDAT_ERANGE, "unwrapped image count beyond int16");
    return AMDAT_OK;
}

}  // namespace amdat

using namespace amdat;

extern "C" {

static int amdat_read_custom_body(const char *path, int n_atoms, int n_frames, int threads, float *x, float *y, float *z,
                      int16_t *ix, int16_t *iy, int16_t *iz, float *box, float *box_lo, int *types)
{
    if (!path || n_atoms < 1 || n_frames < 1 || !x || !y || !z || !ix || !iy || !iz || !box)
        return set_error(AMDAT_EINVAL, "amdat_read_custom: bad arguments");
    MappedFile f;
    int rc = f.open(path);
    if (rc) return rc;
    std::vector<size_t> off;
    rc = index_custom(f, n_frames, threads, off);
    if (rc) return rc;
    return parse_custom(f, off, 0, n_frames, n_atoms, threads, x, y, z, ix, iy, iz, box, box_lo, types);
}

int amdat_read_custom(const char *path, int n_atoms, int n_frames, int threads, float *x, float *y, float *z,
                      int16_t *ix, int16_t *iy, int16_t *iz, float *box, float *box_lo, int *types)
{
    AMDAT_GUARD("amdat_read_custom", amdat_read_custom_body(path, n_atoms, n_frames, threads, x, y, z, ix, iy, iz, box, box_lo, types))
}

static int amdat_read_xyz_body(const char *path, int n_atoms, int n_frames, const float *box3, int threads, float *x, float *y,
                   float *z, int16_t *ix, int16_t *iy, int16_t *iz, int *types)
{
    if (!path || n_atoms < 1 || n_frames < 1 || !box3 || !x || !y || !z || !ix || !iy || !iz)
        return set_error(AMDAT_EINVAL, "amdat_read_xyz: bad arguments");
    MappedFile f;
    int rc = f.open(path);
    if (rc) return rc;
    std::vector<size_t> off;
    rc = index_xyz(f, n_atoms, n_frames, threads, off);
    if (rc) return rc;
    rc = parse_xyz(f, off, 0, n_frames, n_atoms, threads, x, y, z, types);
    if (rc) return rc;
    return unwrap_images(n_atoms, n_frames, box3, x, y, z, ix, iy, iz, nullptr, nullptr, threads);
}

int amdat_read_xyz(const char *path, int n_atoms, int n_frames, const float *box3, int threads, float *x, float *y,
                   float *z, int16_t *ix, int16_t *iy, int16_t *iz, int *types)
{
    AMDAT_GUARD("amdat_read_xyz", amdat_read_xyz_body(path, n_atoms, n_frames, box3, threads, x, y, z, ix, iy, iz, types))
}

}  // extern "C"
