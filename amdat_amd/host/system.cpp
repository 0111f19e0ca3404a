// system.cpp -- System, Trajectory_List, Coordinate and the device
// context of the C++ drop-in layer.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>

#include "amdat_host.h"

namespace amdat_hip {

const char *kSemver = "1.0.1+mi355x";

void warn(const std::string &msg) { std::cout << "\nWarning: " << msg << std::endl; }

static void check(int rc, const char *what)
{
    if (rc < 0) throw Error(std::string(what) + ": " + amdat_last_error(), 0);
}

amdat_ctx *device_context()
{
    static amdat_ctx *ctx = nullptr;
    if (!ctx) {
        int dev = 0;
        if (const char *e = std::getenv("AMDAT_DEVICE")) dev = std::atoi(e);
        check(amdat_ctx_create(dev, &ctx), "device context");
    }
    return ctx;
}

float Coordinate::length() const { return std::sqrt(x_ * x_ + y_ * y_ + z_ * z_); }

std::vector<std::string> tokenize(const std::string &line)
{
    std::vector<std::string> out;
    std::istringstream is(line);
    std::string t;
    while (is >> t) out.push_back(t);
    return out;
}

// ---------------------------------------------------------------------------
// Trajectory_List

Trajectory_List::Trajectory_List(System *sys, std::vector<int> atoms) : system_(sys), atoms_(std::move(atoms))
{
    check(amdat_list_create(device_context(), sys->traj(), (int)atoms_.size(),
                            atoms_.empty() ? nullptr : atoms_.data(), &list_),
          "trajectory list");
}

Trajectory_List::Trajectory_List(System *sys, amdat_list *timed) : system_(sys), list_(timed)
{
    check(amdat_list_shape(timed, &n_times_, &n_entries_), "trajectory list");
    if (n_times_ > 0) {
        conv_.resize(sys->show_n_timesteps());
        off_.resize(n_times_ + 1);
        check(amdat_list_download(device_context(), timed, conv_.data(), off_.data(), nullptr), "trajectory list");
        atoms_loaded_ = false;  // the atoms stay on the device until asked for
    } else {
        atoms_.resize(n_entries_);
        check(amdat_list_download(device_context(), timed, nullptr, nullptr, atoms_.empty() ? nullptr : atoms_.data()),
              "trajectory list");
    }
}

void Trajectory_List::load_atoms() const
{
    if (atoms_loaded_) return;
    atoms_.resize(n_entries_);
    check(amdat_list_download(device_context(), list_, nullptr, nullptr, atoms_.empty() ? nullptr : atoms_.data()),
          "trajectory list");
    atoms_loaded_ = true;
}

Trajectory_List::Trajectory_List(System *sys, amdat_list *paired, PairKeyed)
    : system_(sys), pair_keyed_(true), list_(paired)
{
}

Trajectory_List::~Trajectory_List() { amdat_list_free(device_context(), list_); }

int Trajectory_List::show_n_trajectories(int timeii) const
{
    if (pair_keyed_) throw Error("a pair-keyed list has no atoms per frame", 0);
    if (n_times_ == 0) return (int)atoms_.size();
    if (timeii < 0 || timeii >= (int)conv_.size()) return 0;
    const int c = conv_[timeii];
    return (c >= 0 && c < n_times_) ? off_[c + 1] - off_[c] : 0;
}

std::vector<int> Trajectory_List::atoms_at(int timeii) const
{
    if (pair_keyed_) throw Error("a pair-keyed list has no atoms per frame", 0);
    if (n_times_ == 0) return atoms_;
    if (timeii < 0 || timeii >= (int)conv_.size()) return {};
    const int c = conv_[timeii];
    if (c < 0 || c >= n_times_) return {};
    load_atoms();
    return std::vector<int>(atoms_.begin() + off_[c], atoms_.begin() + off_[c + 1]);
}

// ---------------------------------------------------------------------------
// Trajectory_List_Bins

Trajectory_List_Bins::Trajectory_List_Bins(System *sys, int xbins, int ybins, int zbins) : system_(sys), n_{xbins, ybins, zbins}
{
    const int T = sys->show_n_timesteps();
    std::vector<float> lo = sys->box_lo(), hi(3 * (size_t)T);
    for (int t = 0; t < T; t++) {
        const Coordinate L = sys->size(t);
        hi[3 * t] = lo[3 * t] + L.show_x();
        hi[3 * t + 1] = lo[3 * t + 1] + L.show_y();
        hi[3 * t + 2] = lo[3 * t + 2] + L.show_z();
    }
    assign(lo, hi);
}

Trajectory_List_Bins::Trajectory_List_Bins(System *sys, int xbins, int ybins, int zbins, float x_lo, float x_hi,
                                           float y_lo, float y_hi, float z_lo, float z_hi)
    : system_(sys), n_{xbins, ybins, zbins}
{
    // src/trajectory_list_bins.cpp:309-356: a bound of 0, or beyond the
    // box, is the box boundary of the frame
    const int T = sys->show_n_timesteps();
    const std::vector<float> blo = sys->box_lo();
    std::vector<float> lo(3 * (size_t)T), hi(3 * (size_t)T);
    const float ulo[3] = {x_lo, y_lo, z_lo}, uhi[3] = {x_hi, y_hi, z_hi};
    for (int t = 0; t < T; t++) {
        const Coordinate L = sys->size(t);
        const float len[3] = {L.show_x(), L.show_y(), L.show_z()};
        for (int c = 0; c < 3; c++) {
            const float bl = blo[3 * t + c], bh = blo[3 * t + c] + len[c];
            lo[3 * t + c] = (ulo[c] == 0 || ulo[c] < bl) ? bl : ulo[c];
            hi[3 * t + c] = (uhi[c] == 0 || uhi[c] > bh) ? bh : uhi[c];
        }
    }
    assign(lo, hi);
}

void Trajectory_List_Bins::assign(const std::vector<float> &lo, const std::vector<float> &hi)
{
    if (n_[0] < 1 || n_[1] < 1 || n_[2] < 1) throw Error("bin counts must be at least 1", 0);
    check(amdat_cells_assign(device_context(), system_->traj(), n_[0], n_[1], n_[2], lo.data(), hi.data(), &cells_),
          "bins");
}

Trajectory_List_Bins::~Trajectory_List_Bins() { amdat_cells_free(device_context(), cells_); }

// The bin cap of the point and cluster forms (src/trajectory_list_bins.cpp:
// 487-517, 792-822): the largest half box diagonal over the frames, formed
// with pow in double and kept as float, gives int(length / thickness) + 1.
static int distance_bin_cap(const System *sys, float thickness, int n_bins)
{
    float length = 0;
    for (int t = 0; t < sys->show_n_timesteps(); t++) {
        const Coordinate L = sys->size(t);
        const float xl = L.show_x() / 2, yl = L.show_y() / 2, zl = L.show_z() / 2;
        const double d = std::pow(std::pow(xl, 2) + std::pow(yl, 2) + std::pow(zl, 2), .5);
        if (d > length) length = (float)d;
    }
    const int max_bins = int(length / thickness) + 1;
    if (n_bins < max_bins) return n_bins;
    std::cout << std::endl << "system too small for " << n_bins << " bins. using " << max_bins << " instead." << std::endl;
    return max_bins;
}

Trajectory_List_Bins::Trajectory_List_Bins(System *sys, float thickness, int n_bins, const Trajectory_List *binned,
                                           const Trajectory_List *cluster)
    : system_(sys), n_{1, 1, 1}
{
    n_[2] = distance_bin_cap(sys, thickness, n_bins);
    assign_distance(AMDAT_DIST_LIST, nullptr, 0, 0, binned, cluster, thickness, n_[2]);
}

Trajectory_List_Bins::Trajectory_List_Bins(System *sys, float thickness, int n_bins, const Trajectory_List *binned,
                                           const Coordinate &point)
    : system_(sys), n_{1, 1, 1}
{
    n_[2] = distance_bin_cap(sys, thickness, n_bins);
    const float p[3] = {point.show_x(), point.show_y(), point.show_z()};
    assign_distance(AMDAT_DIST_POINT, p, 0, 0, binned, nullptr, thickness, n_[2]);
}

Trajectory_List_Bins::Trajectory_List_Bins(System *sys, float thickness, int n_bins, const Trajectory_List *binned,
                                           const std::string &plane, float position, const std::string &direction)
    : system_(sys), n_{1, 1, 1}
{
    // src/trajectory_list_bins.cpp:585-700: the reach is the last frame's
    // half box (average) or whole box (above, below) along the plane normal
    int dir;
    if (direction == "average") dir = 0;
    else if (direction == "above") dir = 1;
    else if (direction == "below") dir = 2;
    else throw Error("could not recognize direction " + direction + ". valid choices are above, below or average.", 0);
    int axis;
    if (plane == "x") axis = 0;
    else if (plane == "y") axis = 1;
    else if (plane == "z") axis = 2;
    else throw Error("plane " + plane + " not recognized, plane must be x, y, or z.", 0);
    const Coordinate L = sys->size(sys->show_n_timesteps() - 1);
    const float Lk = axis == 0 ? L.show_x() : (axis == 1 ? L.show_y() : L.show_z());
    const float reach = dir == 0 ? Lk / 2 : Lk;
    const int max_bins = int(reach / thickness) + 1;
    int nb = n_bins;
    if (n_bins >= max_bins) {
        std::cout << std::endl << "system too small for " << n_bins << " bins. Using " << max_bins << " instead."
                  << std::endl;
        nb = max_bins;
    }
    n_[axis] = nb;
    const float p[3] = {position, 0.0f, 0.0f};
    assign_distance(AMDAT_DIST_PLANE, p, axis, dir, binned, nullptr, thickness, nb);
}

void Trajectory_List_Bins::assign_distance(int kind, const float *point, int axis, int direction,
                                           const Trajectory_List *binned, const Trajectory_List *cluster,
                                           float thickness, int n_bins)
{
    check(amdat_cells_distance(device_context(), system_->traj(), binned ? binned->handle() : nullptr, kind, point,
                               axis, direction, cluster ? cluster->handle() : nullptr, thickness, n_bins, &cells_),
          "distance bins");
}

std::unique_ptr<Trajectory_List> Trajectory_List_Bins::intersect(const Trajectory_List *l, int xii, int yii,
                                                                 int zii) const
{
    const int bin = (xii * n_[1] + yii) * n_[2] + zii;
    amdat_list *h = nullptr;
    check(amdat_cells_bin_lists(device_context(), cells_, l ? l->handle() : nullptr, bin, 1, &h), "bin list");
    return std::make_unique<Trajectory_List>(system_, h);
}

std::vector<std::unique_ptr<Trajectory_List>> Trajectory_List_Bins::intersect_all(const Trajectory_List *l) const
{
    const int nb = n_[0] * n_[1] * n_[2];
    std::vector<amdat_list *> h(nb, nullptr);
    check(amdat_cells_bin_lists(device_context(), cells_, l ? l->handle() : nullptr, 0, nb, h.data()), "bin list");
    std::vector<std::unique_ptr<Trajectory_List>> out;
    for (int c = 0; c < nb; c++) out.push_back(std::make_unique<Trajectory_List>(system_, h[c]));
    return out;
}

std::vector<std::unique_ptr<Trajectory_List>> Trajectory_List_Bins::intersect_pairs(const Trajectory_List *l) const
{
    const int nb = n_[0] * n_[1] * n_[2];
    std::vector<int> tg, t0, t1;
    system_->displacement_pair_list(false, tg, t0, t1);
    // One key per distinct frame pair; membership depends on the frames only.
    std::vector<std::pair<int, int>> keys;
    for (size_t i = 0; i < t0.size(); i++) keys.emplace_back(t0[i], t1[i]);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const int nk = (int)keys.size();
    std::vector<int> k0(nk), k1(nk);
    for (int k = 0; k < nk; k++) k0[k] = keys[k].first, k1[k] = keys[k].second;
    std::vector<amdat_list *> h(nb, nullptr);
    check(amdat_cells_pair_lists(device_context(), cells_, l ? l->handle() : nullptr, nk, k0.data(), k1.data(), 0,
                                 nb, h.data()),
          "bin list");
    std::vector<std::unique_ptr<Trajectory_List>> out;
    for (int c = 0; c < nb; c++)
        out.push_back(std::make_unique<Trajectory_List>(system_, h[c], Trajectory_List::PairKeyed{}));
    return out;
}

void Trajectory_List::write_count(const std::string &filename) const
{
    const int nt = n_times_ > 0 ? n_times_ : 1;
    std::vector<int> counts(nt);
    for (int c = 0; c < nt; c++) counts[c] = n_times_ > 0 ? off_[c + 1] - off_[c] : (int)atoms_.size();
    double avg = 0.0;
    for (int c = 0; c < nt; c++) avg += double(counts[c]) / double(nt);
    std::ofstream out(filename);
    out << "Average_trajectories:\t" << avg << "\n";
    out << "Count_List\n";
    for (int c = 0; c < nt; c++) {
        if (n_times_ > 0) {
            const int t = c * system_->show_n_exponential_steps() + system_->show_frt();
            out << system_->show_time(t) << "\t" << counts[c] << "\n";
        } else {
            out << "\t" << counts[c] << "\n";
        }
    }
}

// ---------------------------------------------------------------------------
// System

void System::init_timescheme(const std::vector<std::string> &a)
{
    if (a.empty()) throw Error("missing time scheme line", 0);
    const std::string &type = a[0];
    // src/system.cpp:88-141
    if (type == "linear") {
        if (a.size() != 3) throw Error("Incorrect number of arguments for linear time style.", 0);
        n_timesteps_ = std::atoi(a[1].c_str());
        n_exponentials_ = n_timesteps_ - 1;
        exp_base_ = 1.0f;
        n_exponential_steps_ = 1;
        first_exponent_ = 0;
        frt_ = 0;
        time_unit_ = (float)std::atof(a[2].c_str());
    } else if (type == "exponential") {
        if (a.size() != 7) throw Error("Incorrect number of arguments for exponential time style.", 0);
        n_exponentials_ = std::atoi(a[1].c_str());
        n_exponential_steps_ = std::atoi(a[2].c_str());
        exp_base_ = (float)std::atof(a[3].c_str());
        frt_ = std::atoi(a[4].c_str());
        first_exponent_ = std::atoi(a[5].c_str());
        n_timesteps_ = n_exponential_steps_ * n_exponentials_ + 1;
        time_unit_ = (float)std::atof(a[6].c_str());
    } else if (type == "snapshot") {
        if (a.size() != 1) throw Error("Incorrect number of arguments for snapshot time style.", 0);
        n_timesteps_ = 1;
        n_exponentials_ = 1;
        exp_base_ = 1.0f;
        n_exponential_steps_ = 0;
        first_exponent_ = 0;
        frt_ = 0;
        time_unit_ = 1.0f;
    } else {
        throw Error("Time scheme " + type + " not recognized!", 0);
    }
    if (n_exponentials_ < 1 || n_exponential_steps_ < 0) throw Error("bad time scheme", 0);
}

void System::finish()
{
    n_timegaps_ = n_exponential_steps_ + n_exponentials_ - 1 + (frt_ ? 0 : 1);
    timelist_.assign(n_timesteps_, 0.0f);
    check(amdat_ts_timelist(n_exponentials_, n_exponential_steps_, exp_base_, first_exponent_, time_unit_,
                            timelist_.data(), n_timesteps_),
          "time list");
    if (!traj_)
        check(amdat_traj_upload(device_context(), n_atoms_, n_timesteps_, x_.data(), y_.data(), z_.data(),
                                ix_.data(), iy_.data(), iz_.data(), box_.data(), &traj_),
              "trajectory upload");
}

System::System(int nexp, int nes, float base, int frt, int first_exp, float dt, int n_atoms, const float *x,
               const float *y, const float *z, const int16_t *ix, const int16_t *iy, const int16_t *iz,
               const float *box, const int *types)
{
    n_exponentials_ = nexp;
    n_exponential_steps_ = nes;
    exp_base_ = base;
    frt_ = frt;
    first_exponent_ = first_exp;
    time_unit_ = dt;
    n_timesteps_ = nes * nexp + 1;
    if (nes == 0 && nexp == 1) n_timesteps_ = 1;
    n_atoms_ = n_atoms;
    const size_t n = (size_t)n_atoms * n_timesteps_;
    x_.assign(x, x + n); y_.assign(y, y + n); z_.assign(z, z + n);
    ix_.assign(ix, ix + n); iy_.assign(iy, iy + n); iz_.assign(iz, iz + n);
    box_.assign(box, box + 3 * (size_t)n_timesteps_);
    type_.assign(n_atoms, 1);
    if (types) type_.assign(types, types + n_atoms);
    // One species of single-atom molecules: list order is file order.
    list_order_.resize(n_atoms);
    for (int i = 0; i < n_atoms; i++) list_order_[i] = i;
    atom_species_.assign(n_atoms, 0);
    species_names_ = {"all"};
    finish();
}

System::System(const std::vector<std::string> &block)
{
    // block: system type, trajectory type, file line, time scheme, species
    // line, types line, one count row per species (src/system.cpp:49-226;
    // the custom layout of Manual/custom.md).
    if (block.size() < 6) throw Error("system block too short", 0);
    const std::string systype = tokenize(block[0]).at(0);
    if (systype != "system" && systype != "system_nv" && systype != "system_np")
        throw Error("unknown system type " + systype, 0);
    const std::string trajtype = tokenize(block[1]).at(0);
    if (trajtype != "custom" && trajtype != "xyz")
        throw Error(std::string("Trajectory file type ") + trajtype + " not recognized.", 0);
    auto files = tokenize(block[2]);
    if (files.empty()) throw Error("missing trajectory file name", 0);
    init_timescheme(tokenize(block[3]));
    auto sp = tokenize(block[4]);
    if (sp.size() % 2) throw Error("species line must be name/count pairs", 0);
    for (size_t i = 0; i < sp.size(); i += 2) {
        species_names_.push_back(sp[i]);
        species_mols_.push_back(std::atoi(sp[i + 1].c_str()));
    }
    for (auto &t : tokenize(block[5])) type_ids_.push_back(std::atoi(t.c_str()));
    if (block.size() < 6 + species_names_.size()) throw Error("missing per-species type counts", 0);
    for (size_t s = 0; s < species_names_.size(); s++) {
        std::vector<int> row;
        for (auto &t : tokenize(block[6 + s])) row.push_back(std::atoi(t.c_str()));
        if (row.size() != type_ids_.size()) throw Error("type count row does not match the types line", 0);
        species_type_counts_.push_back(row);
    }
    if (trajtype == "xyz") {
        // box lengths, then bounds (src/system.cpp:385-397)
        const size_t at = 6 + species_names_.size();
        if (block.size() < at + 2) throw Error("xyz system block needs the box and bounds lines", 0);
        auto L = tokenize(block[at]);
        if (L.size() < 3) throw Error("xyz box line needs L_x L_y L_z", 0);
        for (int c = 0; c < 3; c++) xyz_box_[c] = (float)std::atof(L[c].c_str());
    }
    read_trajectory(trajtype, files[0]);
    finish();
}

// Trajectory file straight into device memory (amdat_traj_read): a LAMMPS
// custom dump (x y z with ix iy iz, or xu yu zu; src/system.cpp:1026-1367)
// or an xyz file (src/system.cpp:610-705, images by the reference's
// unwrapping rule), atoms in the same order in every frame.
void System::read_trajectory(const std::string &trajtype, const std::string &file)
{
    n_atoms_ = 0;
    for (size_t s = 0; s < species_names_.size(); s++) {
        int per = 0;
        for (int c : species_type_counts_[s]) per += c;
        n_atoms_ += per * species_mols_[s];
    }
    const int N = n_atoms_, T = n_timesteps_;
    type_.assign(N, 0);
    check(amdat_traj_read(device_context(), trajtype.c_str(), file.c_str(), N, T,
                          trajtype == "xyz" ? xyz_box_ : nullptr, 0, type_.data(), &traj_),
          "reading the trajectory");
    box_.assign(3 * (size_t)T, 0.0f);
    check(amdat_traj_shape(traj_, nullptr, nullptr, box_.data()), "trajectory shape");
    // Reference list order (src/system.cpp:4783-4791): species -> molecule
    // -> type (types-line order) -> index within the molecule.
    list_order_.clear();
    atom_species_.assign(N, 0);
    int pos = 0;
    for (size_t s = 0; s < species_names_.size(); s++) {
        int per = 0;
        for (int c : species_type_counts_[s]) per += c;
        for (int m = 0; m < species_mols_[s]; m++) {
            for (int a = 0; a < per; a++) atom_species_[pos + a] = (int)s;
            for (size_t ti = 0; ti < type_ids_.size(); ti++)
                for (int a = 0; a < per; a++)
                    if (type_[pos + a] == type_ids_[ti]) list_order_.push_back(pos + a);
            pos += per;
        }
    }
    if ((int)list_order_.size() != N) throw Error("atom types in the file do not match the types line", 0);
}

System::~System() { amdat_traj_free(device_context(), traj_); }

Coordinate System::min_box_dimensions() const
{
    float m[3] = {box_[0], box_[1], box_[2]};
    for (int t = 1; t < n_timesteps_; t++)
        for (int c = 0; c < 3; c++) m[c] = std::min(m[c], box_[3 * t + c]);
    return Coordinate(m[0], m[1], m[2]);
}

Coordinate System::size(int t) const { return Coordinate(box_[3 * t], box_[3 * t + 1], box_[3 * t + 2]); }

std::vector<float> System::box_lo() const
{
    std::vector<float> lo(3 * (size_t)n_timesteps_, 0.0f);
    check(amdat_traj_bounds(traj_, lo.data()), "box bounds");
    return lo;
}

std::vector<int> System::timegap_weighting(bool fullblock) const
{
    std::vector<int> w(n_timegaps_);
    check(amdat_ts_weighting(n_exponentials_, n_exponential_steps_, frt_, fullblock, displacement_limit_, w.data(),
                             n_timegaps_),
          "timegap weighting");
    return w;
}

std::vector<float> System::displacement_times() const
{
    std::vector<float> out(n_timegaps_);
    check(amdat_ts_displacement_times(n_exponentials_, n_exponential_steps_, frt_, timelist_.data(), out.data(),
                                      n_timegaps_),
          "displacement times");
    return out;
}

amdat_pairs *System::gap_pairs(int time_index) const
{
    int n = 0;
    int rc = amdat_ts_gap_pairs(n_exponentials_, n_exponential_steps_, frt_, displacement_limit_, time_index, nullptr,
                                nullptr, nullptr, 0, &n);
    if (rc < 0) throw Error(amdat_last_error(), 0);
    std::vector<int> tg(n), t0(n), t1(n);
    check(amdat_ts_gap_pairs(n_exponentials_, n_exponential_steps_, frt_, displacement_limit_, time_index, tg.data(),
                             t0.data(), t1.data(), n, &n),
          "pairs");
    amdat_pairs *p = nullptr;
    check(amdat_pairs_upload(device_context(), n, tg.data(), t0.data(), t1.data(), n_timegaps_, &p), "pair upload");
    return p;
}

void System::displacement_pair_list(bool fullblock, std::vector<int> &tg, std::vector<int> &t0,
                                    std::vector<int> &t1) const
{
    int n = 0;
    check(amdat_ts_pairs(n_exponentials_, n_exponential_steps_, frt_, fullblock, displacement_limit_, nullptr, nullptr,
                         nullptr, 0, &n),
          "pairs");
    tg.resize(n);
    t0.resize(n);
    t1.resize(n);
    check(amdat_ts_pairs(n_exponentials_, n_exponential_steps_, frt_, fullblock, displacement_limit_, tg.data(),
                         t0.data(), t1.data(), n, &n),
          "pairs");
}

amdat_pairs *System::timegap_loop_pairs(int timegapii, bool fullblock) const
{
    // src/system.cpp:3972-4023, the break test as there: count == limit
    std::vector<int> tg, t0, t1;
    const int f = frt_ ? 1 : 0, nes = n_exponential_steps_, nexp = n_exponentials_;
    int count = 0;
    auto add = [&](int g, int a, int b) {
        tg.push_back(g);
        t0.push_back(a);
        t1.push_back(b);
        return ++count == displacement_limit_;
    };
    if (timegapii < nes) {
        for (int b = 0; b < nexp; b++)
            if (add(timegapii, nes * b + f, nes * b + f + timegapii)) break;
    } else if (timegapii < n_timegaps_ - 1 + f) {
        const int bg = timegapii - nes + 1;
        bool stop = false;
        for (int e = 0; e < (fullblock ? nes - 1 : 0) + 1 && !stop; e++)
            for (int b = 0; b < nexp - bg; b++) {
                const int a = nes * b + e + f;
                if (add(timegapii, a, a + nes * bg)) {
                    stop = true;
                    break;
                }
            }
    } else if (timegapii == n_timegaps_ - 1 + f) {
        add(n_timegaps_ - 1, 0, n_timesteps_ - 1);
    } else {
        throw Error("Requested timegap out of range.", 0);
    }
    amdat_pairs *p = nullptr;
    check(amdat_pairs_upload(device_context(), (int)tg.size(), tg.data(), t0.data(), t1.data(), n_timegaps_, &p),
          "pair upload");
    return p;
}

amdat_pairs *System::displacement_pairs(bool fullblock, std::vector<int> *tg_out, std::vector<int> *t0_out) const
{
    int n = 0;
    check(amdat_ts_pairs(n_exponentials_, n_exponential_steps_, frt_, fullblock, displacement_limit_, nullptr, nullptr,
                         nullptr, 0, &n),
          "pairs");
    std::vector<int> tg(n), t0(n), t1(n);
    check(amdat_ts_pairs(n_exponentials_, n_exponential_steps_, frt_, fullblock, displacement_limit_, tg.data(),
                         t0.data(), t1.data(), n, &n),
          "pairs");
    amdat_pairs *p = nullptr;
    check(amdat_pairs_upload(device_context(), n, tg.data(), t0.data(), t1.data(), n_timegaps_, &p), "pair upload");
    if (tg_out) *tg_out = tg;
    if (t0_out) *t0_out = t0;
    return p;
}

std::vector<int> System::select(const std::vector<std::string> &args) const
{
    // create_list keywords (Manual/create_list.md): all, type_system <t>,
    // species <name>, type_species <name> <t>; several keywords on one
    // line select the union.
    std::vector<char> keep(n_atoms_, 0);
    size_t i = 0;
    auto species_index = [&](const std::string &nm) {
        for (size_t s = 0; s < species_names_.size(); s++)
            if (species_names_[s] == nm) return (int)s;
        throw Error("species " + nm + " not found", 1);
    };
    while (i < args.size()) {
        const std::string &k = args[i];
        if (k == "all") {
            std::fill(keep.begin(), keep.end(), 1);
            i += 1;
        } else if (k == "type_system" && i + 1 < args.size()) {
            const int t = std::atoi(args[i + 1].c_str());
            for (int a = 0; a < n_atoms_; a++) if (type_[a] == t) keep[a] = 1;
            i += 2;
        } else if (k == "species" && i + 1 < args.size()) {
            const int s = species_index(args[i + 1]);
            for (int a = 0; a < n_atoms_; a++) if (atom_species_[a] == s) keep[a] = 1;
            i += 2;
        } else if (k == "type_species" && i + 2 < args.size()) {
            const int s = species_index(args[i + 1]);
            const int t = std::atoi(args[i + 2].c_str());
            for (int a = 0; a < n_atoms_; a++) if (atom_species_[a] == s && type_[a] == t) keep[a] = 1;
            i += 3;
        } else {
            throw Error("create_list keyword " + k + " is not supported by this build", 1);
        }
    }
    std::vector<int> out;
    for (int a : list_order_) if (keep[a]) out.push_back(a);
    return out;
}

}  // namespace amdat_hip
