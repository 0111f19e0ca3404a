// analyses.cpp -- Wave_Vectors and the Analysis classes of the drop-in
// layer.  Each analyze() runs the device path through the C ABI; each
// write() reproduces the reference writer's text format.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>

#include "amdat_host.h"

namespace amdat_hip {

namespace {
constexpr double kPi = 3.14159265358979323846;  // src/wave_vectors.cpp:16

void check(int rc, const char *what)
{
    if (rc < 0) throw Error(std::string(what) + ": " + amdat_last_error(), 0);
}

// q-vector file stem of a plane: AMDAT_WV3D, AMDAT_WV2D or AMDAT_WV1D (the
// reference's compile-time WV3D / WV2D / WV1D paths).
std::string wv_stem(const std::string &plane)
{
    const char *var = plane == "xyz" ? "AMDAT_WV3D" : (plane.size() == 2 ? "AMDAT_WV2D" : "AMDAT_WV1D");
    const char *dir = plane == "xyz" ? "qvectors3d" : (plane.size() == 2 ? "qvectors2d" : "qvectors1d");
    const char *e = std::getenv(var);
    if (!e || !*e)
        throw Error(std::string(var) + " is not set: point it at the q-vector file stem (the reference's " + dir +
                        "/qvector)", 0);
    return e;
}
}  // namespace

// ---------------------------------------------------------------------------
// Wave_Vectors (src/wave_vectors.cpp:77-181, 430-484, 527-566)

Wave_Vectors::Wave_Vectors(System *sys, const std::string &plane, float max_length_scale, int n_wavenumbers)
    : system_(sys)
{
    if (n_wavenumbers < 1 || n_wavenumbers > 299) throw Error("n_wavenumbers must be in 1..299", 0);
    float min_size;
    const Coordinate mb = sys->min_box_dimensions();
    if (max_length_scale <= 0.0f) {
        std::cout << "Maximum lengthscale for wavevector decomposition <= 0; defaulting to box size.\n";
        if (plane == "xy") min_size = mb.show_x() < mb.show_y() ? mb.show_x() : mb.show_y();
        else if (plane == "xz") min_size = mb.show_x() < mb.show_z() ? mb.show_x() : mb.show_z();
        else if (plane == "yz") min_size = mb.show_z() < mb.show_y() ? mb.show_z() : mb.show_y();
        else if (plane == "xyz") min_size = mb.min();
        else if (plane == "x") min_size = mb.show_x();
        else if (plane == "y") min_size = mb.show_y();
        else if (plane == "z") min_size = mb.show_z();
        else throw Error("plane command not recognized. valid inputs are: x, y, z, xy, xz, yz, xyz", 1);
    } else {
        if (max_length_scale > mb.min())
            std::cout << "\nWARNING: Maximum length scale larger than smallest box dimension. Consider choosing a "
                         "different maximum for wavevector decomposition.";
        min_size = max_length_scale;
    }
    wavegrid_spacing_ = (float)(2.0 * kPi / min_size);
    delta_wavenumber_ = (float)(wavegrid_spacing_ / 2.0);
    for (int wn = 0; wn < n_wavenumbers; wn++) approx_.push_back(wavegrid_spacing_ + delta_wavenumber_ * wn);
    off_.assign(n_wavenumbers + 1, 0);
    const int cap = 300 * n_wavenumbers + 1024;
    nvec_.assign(3 * (size_t)cap, 0);
    int n = 0;
    check(amdat_shells_read_plane(wv_stem(plane).c_str(), plane.c_str(), 2, n_wavenumbers, off_.data(), nvec_.data(),
                                  cap, &n, &denominator_),
          "reading wave vectors");
    nvec_.resize(3 * (size_t)n);
    // the kernels take k = 2 pi n / l with integer n: 1D shells (half-integer
    // multiples of the spacing) are held as 2n over 2 l
    length_scale_ = min_size * (float)denominator_;
    check(amdat_shells_upload(device_context(), n_wavenumbers, off_.data(), nvec_.data(), &shells_),
          "wave vector upload");
}

Wave_Vectors::~Wave_Vectors() { amdat_shells_free(device_context(), shells_); }

std::vector<Coordinate> Wave_Vectors::vectorlist(int index) const
{
    std::vector<Coordinate> out;
    const float inv_d = 1.0f / (float)denominator_;  // exact: 1 or 1/2
    for (int v = off_[index]; v < off_[index + 1]; v++)
        out.emplace_back((float)nvec_[3 * v] * inv_d * wavegrid_spacing_,
                         (float)nvec_[3 * v + 1] * inv_d * wavegrid_spacing_,
                         (float)nvec_[3 * v + 2] * inv_d * wavegrid_spacing_);
    return out;
}

float Wave_Vectors::show_mean_wavenumber(int index) const
{
    float mean = 0;
    auto vl = vectorlist(index);
    for (auto &c : vl) mean += c.length();
    mean /= vl.size();
    return mean;
}

float Wave_Vectors::show_stdev_wavenumber(int index) const
{
    float variance = 0;
    const float mean = show_mean_wavenumber(index);
    auto vl = vectorlist(index);
    for (auto &c : vl) variance += std::pow(c.length() - mean, 2.0);
    variance /= vl.size();
    return (float)std::pow(variance, 0.5);
}

Coordinate Wave_Vectors::show_mean_wavevector(int index) const
{
    float mx = 0, my = 0, mz = 0;
    auto vl = vectorlist(index);
    for (auto &c : vl) { mx += c.show_x(); my += c.show_y(); mz += c.show_z(); }
    const float n = (float)vl.size();
    return Coordinate(mx / n, my / n, mz / n);
}

// ---------------------------------------------------------------------------
// Analysis (src/analysis.h:83-106): the reference's messages for methods a
// class does not implement

void Analysis::analyze(Trajectory_List *, Trajectory_List *)
{
    std::cout << "Error: Trajectory list targets with two lists not implemented for this analysis method.\n";
}

void Analysis::analyze(Trajectory_List *)
{
    std::cout << "Error: Trajectory list targets with one list not implemented for this analysis method.\n";
}

void Analysis::bin_hook(Trajectory_List *, int, int, int)
{
    std::cout << "Error: Trajectory list bins not implemented for this analysis method.\n";
}

void Analysis::write(std::string) const
{
    std::cout << "Error: No standard write method implemented for this analysis method.\n";
}

// ---------------------------------------------------------------------------
// Wave_Density (src/wave_density.cpp:46-365)

static std::shared_ptr<amdat_rho> own_rho(amdat_rho *r)
{
    return std::shared_ptr<amdat_rho>(r, [](amdat_rho *p) { amdat_rho_free(device_context(), p); });
}

Wave_Density::Wave_Density(System *sys, const Wave_Vectors *wv, bool stf)
    : Analysis(sys), system_(sys), wv_(wv), first_(0), last_(wv->show_n_wavenumbers() - 1), sf_(stf)
{
}

Wave_Density::Wave_Density(System *sys, const Wave_Vectors *wv, int inner, int outer)
    : Analysis(sys), system_(sys), wv_(wv), first_(inner), last_(outer)
{
    if (inner < 0 || outer < inner || outer >= wv->show_n_wavenumbers())
        throw Error("wavenumber range outside the wave vector table", 0);
}

void Wave_Density::set(System *sys, const Wave_Vectors *wv)
{
    set(sys, wv, 0, wv->show_n_wavenumbers() - 1);
}

void Wave_Density::set(System *sys, const Wave_Vectors *wv, int inner, int outer)
{
    if (inner < 0 || outer < inner || outer >= wv->show_n_wavenumbers())
        throw Error("wavenumber range outside the wave vector table", 0);
    system_ = sys;
    system_base_ = sys;
    wv_ = wv;
    first_ = inner;
    last_ = outer;
    n_atoms_.clear();
    rho_.reset();
    host_.reset();
}

void Wave_Density::analyze(Trajectory_List *t_list)
{
    const int T = system_->show_n_timesteps();
    if ((int)n_atoms_.size() != T) n_atoms_.assign(T, 0);
    host_.reset();
    amdat_rho *r = nullptr;
    if (!sf_) {
        for (int t = 0; t < T; t++) n_atoms_[t] += t_list->show_n_trajectories(t);
        check(amdat_rho_compute(device_context(), system_->traj(), t_list->handle(), wv_->handle(), first_, last_,
                                wv_->show_length_scale(), nullptr, 0, AMDAT_PREC_FP32C, &r),
              "wave density");
    } else {
        // the list's atoms at the block starts only: a time-varying list
        // whose other frames hold no atoms
        std::vector<int> conv(T, -1), off(1, 0), idx;
        int c = 0;
        for (int b = 0; b < system_->show_n_exponentials(); b++) {
            const int t = system_->show_n_exponential_steps() * b;
            if (t >= T) break;
            const std::vector<int> a = t_list->atoms_at(t);
            n_atoms_[t] += (int)a.size();
            idx.insert(idx.end(), a.begin(), a.end());
            off.push_back((int)idx.size());
            conv[t] = c++;
        }
        amdat_list *h = nullptr;
        check(amdat_list_create_timed(device_context(), system_->traj(), c, conv.data(), off.data(),
                                      idx.empty() ? nullptr : idx.data(), &h),
              "wave density list");
        Trajectory_List starts(system_, h);
        check(amdat_rho_compute(device_context(), system_->traj(), starts.handle(), wv_->handle(), first_, last_,
                                wv_->show_length_scale(), nullptr, 0, AMDAT_PREC_FP32C, &r),
              "wave density");
    }
    if (!rho_) {
        rho_ = own_rho(r);
        return;
    }
    // a later call adds to the density: a new device object, so copies made
    // in between keep theirs
    std::shared_ptr<amdat_rho> add = own_rho(r);
    amdat_rho *sum = nullptr;
    check(amdat_rho_sum(device_context(), rho_.get(), add.get(), &sum), "wave density sum");
    rho_ = own_rho(sum);
}

int Wave_Density::show_n_atoms_looped() const
{
    int n = 0;
    for (int v : n_atoms_) n += v;
    return n;
}

void Wave_Density::fetch() const
{
    if (host_) return;
    const amdat_rho *r = device();
    if (!r) throw Error("wave density not computed", 0);
    int nf = 0, nv = 0;
    amdat_rho_shape(r, &nf, &nv);
    auto h = std::make_shared<std::vector<double>>(2 * (size_t)nf * nv, 0.0);
    check(amdat_rho_download(device_context(), r, h->data(), h->size()), "wave density download");
    host_ = h;
}

std::complex<double> Wave_Density::show_density(int t, int wn, int v) const
{
    fetch();
    const int K = wv_->offsets()[last_ + 1] - wv_->offsets()[first_];
    const int idx = wv_->offsets()[first_ + wn] - wv_->offsets()[first_] + v;
    const size_t o = 2 * ((size_t)t * K + idx);
    return {(*host_)[o], (*host_)[o + 1]};
}

double Wave_Density::show_density_re(int t, int wn, int v) const { return show_density(t, wn, v).real(); }

double Wave_Density::show_density_im(int t, int wn, int v) const { return show_density(t, wn, v).imag(); }

// ---------------------------------------------------------------------------
// Structure_Factor (src/structure_factor.cpp:55-344, src/analysis_onetime.cpp:22-85)

Structure_Factor::Structure_Factor(System *sys, const Wave_Vectors *wv, int timescheme)
    : Analysis(sys), system_(sys), wv_(wv), time_scheme_(timescheme), structure_factor_(wv->show_n_wavenumbers(), 0.0f)
{
}

std::vector<int> Structure_Factor::frames() const
{
    std::vector<int> f;
    if (time_scheme_ == -1) {
        for (int t = 0; t < system_->show_n_timesteps(); t++) f.push_back(t);
    } else if (time_scheme_ < -1) {
        f.push_back(-time_scheme_ - 2);
    } else {
        // The reference bounds this loop by the block count, not the frame
        // count (src/analysis_onetime.cpp:43); kept for parity.
        for (int t = time_scheme_; t < system_->show_n_exponentials(); t += system_->show_n_exponential_steps()) {
            f.push_back(t);
            if (system_->show_n_exponential_steps() == 0) break;
        }
    }
    return f;
}

static void structure_factor_run(System *sys, const Wave_Vectors *wv, const std::vector<int> &fr,
                                 Trajectory_List *l1, Trajectory_List *l2, std::vector<float> &S)
{
    const int nsh = wv->show_n_wavenumbers();
    std::fill(S.begin(), S.end(), 0.0f);
    if (fr.empty()) return;
    for (int f : fr)
        if (f < 0 || f >= sys->show_n_timesteps()) throw Error("structure factor frame out of range", 0);
    amdat_rho *r1 = nullptr, *r2 = nullptr;
    check(amdat_rho_compute(device_context(), sys->traj(), l1->handle(), wv->handle(), 0, nsh - 1,
                            wv->show_length_scale(), fr.data(), (int)fr.size(), AMDAT_PREC_FP32C, &r1),
          "structure factor density");
    if (l2) {
        int rc = amdat_rho_compute(device_context(), sys->traj(), l2->handle(), wv->handle(), 0, nsh - 1,
                                   wv->show_length_scale(), fr.data(), (int)fr.size(), AMDAT_PREC_FP32C, &r2);
        if (rc < 0) { amdat_rho_free(device_context(), r1); check(rc, "structure factor density"); }
    }
    std::vector<double> s(nsh);
    int rc = amdat_sk(device_context(), r1, r2 ? r2 : r1, s.data());
    amdat_rho_free(device_context(), r1);
    amdat_rho_free(device_context(), r2);
    check(rc, "structure factor");
    for (int i = 0; i < nsh; i++) S[i] = (float)s[i];
}

void Structure_Factor::analyze(Trajectory_List *t_list)
{
    structure_factor_run(system_, wv_, frames(), t_list, nullptr, structure_factor_);
}

void Structure_Factor::analyze(Trajectory_List *t_list, Trajectory_List *t_list2)
{
    structure_factor_run(system_, wv_, frames(), t_list, t_list2, structure_factor_);
}

void Structure_Factor::write(std::string filename) const
{
    std::ofstream out(filename);
    std::cout << "\nWriting to file " << filename << ".";
    out << "Structure factor data created by AMDAT v." << kSemver << "\n";
    out << "Nominal_q\tMean_q\tStddev_q\tMean_qx\tMean_qy\tMean_qz\tS(q)\n";
    for (int b = 0; b < wv_->show_n_wavenumbers(); b++) {
        Coordinate m = wv_->show_mean_wavevector(b);
        out << wv_->show_approx_wavenumber(b) << "\t" << wv_->show_mean_wavenumber(b) << "\t"
            << wv_->show_stdev_wavenumber(b) << "\t" << m.show_x() << "\t" << m.show_y() << "\t" << m.show_z() << "\t"
            << structure_factor_[b] << std::endl;
    }
    out << "\n";
}

// ---------------------------------------------------------------------------
// Correlation_2D-style writer (src/correlation_2d.cpp:475-504)

static void write_correlation(const std::string &filename, const Wave_Vectors *wv, int first, int last,
                              const std::vector<float> &timetable, const std::vector<std::vector<float>> &corr)
{
    std::ofstream out(filename);
    std::cout << "\nWriting to file " << filename << "." << std::endl;
    out << "Correlation data created by AMDAT v." << kSemver << "\n";
    out << "\t";
    for (int b = first; b <= last; b++) out << wv->show_approx_wavenumber(b) << "\t";
    out << "\n";
    for (size_t t = 0; t < corr.size(); t++) {
        out << timetable[t] << "\t";
        for (float v : corr[t]) out << v << "\t";
        out << "\n";
    }
}

// ---------------------------------------------------------------------------
// Intermediate_Scattering_Function (src/intermediate_scattering_function.cpp:237-322)

static void check_densities(const Wave_Density *wd1, const Wave_Density *wd2)
{
    if (wd1->show_wavevectors() != wd2->show_wavevectors())
        throw Error("Wave vectors for density distributions do not match!", 0);
    if (wd1->show_system() != wd2->show_system()) throw Error("Systems for density distributions do not match!", 0);
    if (!wd1->device() || !wd2->device()) throw Error("wave density not computed", 0);
}

static void check_same_range(const Wave_Density *wd1, const Wave_Density *wd2)
{
    if (wd1->show_first() != wd2->show_first() || wd1->show_last() != wd2->show_last())
        throw Error("Wave vector bins for density distributions do not match!", 0);
}

Intermediate_Scattering_Function::Intermediate_Scattering_Function(System *sys, Wave_Density *wd1,
                                                                   Wave_Density *wd2, bool parallel)
    : Analysis(sys), system_(sys), first_bin_(wd1->show_first()), last_bin_(wd1->show_last())
{
    check_densities(wd1, wd2);
    check_same_range(wd1, wd2);
    amdat_pairs *p = sys->displacement_pairs(parallel);
    compute(wd1, wd2, p, -1);
}

Intermediate_Scattering_Function::Intermediate_Scattering_Function(System *sys, Wave_Density *wd1,
                                                                   Wave_Density *wd2, int wn, bool parallel,
                                                                   bool physics_correct)
    : Intermediate_Scattering_Function(sys, wd1, wd2, wn, wn, parallel, physics_correct)
{
}

Intermediate_Scattering_Function::Intermediate_Scattering_Function(System *sys, Wave_Density *wd1,
                                                                   Wave_Density *wd2, int wn1, int wn2,
                                                                   bool parallel, bool physics_correct)
    : Analysis(sys), system_(sys), first_bin_(wn1), last_bin_(wn2)
{
    check_densities(wd1, wd2);
    if (wn1 < wd1->show_first() || wn1 < wd2->show_first())
        throw Error("requested wavenumber below range of that contained in fourier density object.", 0);
    if (wn1 > wd1->show_last() || wn1 > wd2->show_last() || wn2 > wd1->show_last() || wn2 > wd2->show_last() ||
        wn2 < wn1)
        throw Error("requested wavenumber above range of that contained in fourier density object.", 0);
    if (wd1->show_first() != wd2->show_first())
        throw Error("Wave vector bins for density distributions do not match!", 0);
    amdat_pairs *p = sys->displacement_pairs(parallel);
    compute(wd1, wd2, p, -1, !physics_correct && wn1 != wd1->show_first());
}

Intermediate_Scattering_Function::Intermediate_Scattering_Function(System *sys, int timegapii, Wave_Density *wd1,
                                                                   Wave_Density *wd2, bool parallel)
    : Analysis(sys), system_(sys), first_bin_(wd1->show_first()), last_bin_(wd1->show_last())
{
    check_densities(wd1, wd2);
    check_same_range(wd1, wd2);
    if (timegapii < 0 || timegapii >= sys->show_n_timegaps()) throw Error("Requested timegap out of range.", 0);
    amdat_pairs *p = sys->timegap_loop_pairs(timegapii, parallel);
    compute(wd1, wd2, p, timegapii);
}

// F over the given pairs (amdat_fkt; FP64 sums normalised by the atoms at
// t0, src/intermediate_scattering_function.cpp:551-566 and 310-314); takes
// ownership of pairs.  only_timegap >= 0 keeps that row and zeroes the rest.
// slot_quirk: column i sums the density's slot i (shell first + i) over the
// vector count of shell first_bin_ + i (amdat_fkt_vectors), as the
// reference's displacementkernel indexes the densities.
void Intermediate_Scattering_Function::compute(Wave_Density *wd1, Wave_Density *wd2, amdat_pairs *p,
                                               int only_timegap, bool slot_quirk)
{
    System *sys = system_;
    wv_ = wd1->show_wavevectors();
    timetable_ = sys->displacement_times();
    const int ntg = sys->show_n_timegaps(), nsh = last_bin_ - first_bin_ + 1;
    std::vector<double> F((size_t)ntg * nsh), cnt(ntg);
    int rc;
    if (slot_quirk) {
        const std::vector<int> &off = wv_->offsets();
        const int first = wd1->show_first();
        std::vector<int> vf(nsh), vc(nsh);
        for (int i = 0; i < nsh; i++) {
            const int slot = first + i, wn = first_bin_ + i;
            vf[i] = off[slot] - off[first];
            vc[i] = off[wn + 1] - off[wn];
            if (vc[i] > off[slot + 1] - off[slot]) {
                amdat_pairs_free(device_context(), p);
                throw Error("Intermediate_Scattering_Function: shell " + std::to_string(wn) +
                                " is read from the density's slot " + std::to_string(i) +
                                ", which holds fewer vectors (the reference reads past it)",
                            10);
            }
        }
        rc = amdat_fkt_vectors(device_context(), wd1->device(), wd2->device(), p, nsh, vf.data(), vc.data(),
                               F.data(), cnt.data());
    } else {
        rc = amdat_fkt(device_context(), wd1->device(), wd2->device(), p, first_bin_ - wd1->show_first(),
                       last_bin_ - wd1->show_first(), F.data(), cnt.data());
    }
    amdat_pairs_free(device_context(), p);
    check(rc, "intermediate scattering function");
    correlation_.assign(ntg, std::vector<float>(nsh, 0.0f));
    for (int t = 0; t < ntg; t++) {
        if (only_timegap >= 0 && t != only_timegap) continue;
        for (int s = 0; s < nsh; s++) correlation_[t][s] = (float)F[(size_t)t * nsh + s];
    }
}

void Intermediate_Scattering_Function::write(std::string filename) const
{
    write_correlation(filename, wv_, first_bin_, last_bin_, timetable_, correlation_);
}

// ---------------------------------------------------------------------------
// Incoherent_Scattering_Function (src/incoherent_scattering_function.cpp:176-299)

Incoherent_Scattering_Function::Incoherent_Scattering_Function(System *sys, const Wave_Vectors *wv, int inner,
                                                               int outer, bool fblock)
    : Analysis(sys), system_(sys), wv_(wv), first_bin_(inner), last_bin_(outer), fullblock_(fblock)
{
    if (inner < 0 || outer < inner || outer >= wv->show_n_wavenumbers())
        throw Error("wavenumber range outside the wave vector table", 0);
    timetable_ = sys->displacement_times();
    correlation_.assign(sys->show_n_timegaps(), std::vector<float>(outer - inner + 1, 0.0f));
}

void Incoherent_Scattering_Function::analyze(Trajectory_List *t_list)
{
    amdat_pairs *p = system_->displacement_pairs(fullblock_);
    const int ntg = system_->show_n_timegaps(), nsh = last_bin_ - first_bin_ + 1;
    std::vector<double> F((size_t)ntg * nsh), cnt(ntg);
    int rc = amdat_fs(device_context(), system_->traj(), t_list->handle(), wv_->handle(), first_bin_, last_bin_,
                      wv_->show_length_scale(), p, AMDAT_PREC_FP32C, F.data(), cnt.data());
    amdat_pairs_free(device_context(), p);
    check(rc, "incoherent scattering function");
    for (int t = 0; t < ntg; t++)
        for (int s = 0; s < nsh; s++) correlation_[t][s] = (float)F[(size_t)t * nsh + s];
}

void Incoherent_Scattering_Function::write(std::string filename) const
{
    write_correlation(filename, wv_, first_bin_, last_bin_, timetable_, correlation_);
}

// ---------------------------------------------------------------------------
// Mean_Square_Displacement (src/mean_square_displacement.cpp:54-200)

Mean_Square_Displacement::Mean_Square_Displacement(System *sys) : Analysis(sys), system_(sys)
{
    timetable_ = sys->displacement_times();
    msd_.assign(sys->show_n_timegaps(), 0.0f);
}

void Mean_Square_Displacement::analyze(Trajectory_List *t_list)
{
    // MSD always uses block-start time origins (fullblock = false).
    amdat_pairs *p = system_->displacement_pairs(false);
    const int ntg = system_->show_n_timegaps();
    std::vector<double> s2(ntg);
    s4_.assign(ntg, 0.0);
    count_.assign(ntg, 0.0);
    int rc = amdat_msd_ngp(device_context(), system_->traj(), t_list->handle(), p, s2.data(), s4_.data(),
                           count_.data());
    amdat_pairs_free(device_context(), p);
    check(rc, "mean square displacement");
    msd64_.assign(ntg, 0.0);
    for (int t = 0; t < ntg; t++) {
        msd64_[t] = s2[t] / count_[t];
        msd_[t] = (float)msd64_[t];
    }
}

void Mean_Square_Displacement::write(std::string filename) const
{
    std::cout << "\nWriting msd to file " << filename << ".";
    std::ofstream out(filename);
    out << "Mean square displacement data created by AMDAT v." << kSemver << "\n";
    for (size_t t = 0; t < msd_.size(); t++) out << timetable_[t] << "\t" << msd_[t] << "\n";
}

// ---------------------------------------------------------------------------
// Non_Gaussian_Parameter (src/non_gaussian_parameter.cpp:96-190)

Non_Gaussian_Parameter::Non_Gaussian_Parameter(System *sys, const Mean_Square_Displacement *msd)
    : Analysis(sys), system_(sys), msd_(msd)
{
    timetable_ = sys->displacement_times();
    ngp_.assign(sys->show_n_timegaps(), 0.0f);
}

void Non_Gaussian_Parameter::analyze(Trajectory_List *t_list)
{
    // Fourth moment from the same pass as the MSD's second moment.
    Mean_Square_Displacement m(system_);
    m.analyze(t_list);
    for (size_t t = 0; t < ngp_.size(); t++) {
        const double msd = msd_ && (int)msd_->msd64_.size() == (int)ngp_.size() ? msd_->msd64_[t] : m.msd64_[t];
        // alpha2 = 3 <r^4> / (5 <r^2>^2) - 1; NaN at t = 0 as in the reference.
        ngp_[t] = (float)(m.s4_[t] * (3.0 / m.count_[t]) / (5.0 * msd * msd) - 1.0);
    }
}

void Non_Gaussian_Parameter::write(std::string filename) const
{
    std::cout << "\nWriting non-Gaussian parameter to file " << filename << ".";
    std::ofstream out(filename);
    out << "Non-Gaussian parameter data created by AMDAT v." << kSemver << "\n";
    for (size_t t = 0; t < ngp_.size(); t++) out << timetable_[t] << "\t" << ngp_[t] << "\n";
}

int Non_Gaussian_Parameter::max() const
{
    int maxtime = -1;
    float maxvalue = 0;
    for (size_t t = 1; t < ngp_.size(); t++)
        if (ngp_[t] > maxvalue) { maxvalue = ngp_[t]; maxtime = (int)t; }
    return maxtime;
}


// ---------------------------------------------------------------------------
// Displacement statistics (src/mean_square_displacement_2d.cpp,
// src/mean_displacement.cpp, src/van_hove_self.cpp): pairs with
// fullblock = false (System::displacement_list's default)

void Displacement_Stats::run(System *sys, Trajectory_List *t_list, int n_bins, float max_range)
{
    amdat_pairs *p = sys->displacement_pairs(false);
    const int ntg = sys->show_n_timegaps();
    sum_d.assign(3 * (size_t)ntg, 0.0);
    sum_d2.assign(3 * (size_t)ntg, 0.0);
    count.assign(ntg, 0.0);
    hist.assign((size_t)ntg * std::max(n_bins, 1), 0ull);
    int rc = amdat_displacements(device_context(), sys->traj(), t_list->handle(), p, n_bins, max_range,
                                 sum_d.data(), sum_d2.data(), hist.data(), count.data());
    amdat_pairs_free(device_context(), p);
    check(rc, "displacement statistics");
}

Mean_Square_Displacement_2D::Mean_Square_Displacement_2D(System *sys, const std::string &plane)
    : Analysis(sys), system_(sys), plane_(plane)
{
    if (plane == "xy") { a0_ = 0; a1_ = 1; }
    else if (plane == "xz") { a0_ = 0; a1_ = 2; }
    else if (plane == "yz") { a0_ = 1; a1_ = 2; }
    else throw Error("Error: plane command " + plane + " not understood.", 1);
    timetable_ = sys->displacement_times();
    msd_.assign(sys->show_n_timegaps(), 0.0f);
}

void Mean_Square_Displacement_2D::analyze(Trajectory_List *t_list)
{
    Displacement_Stats d;
    d.run(system_, t_list, 0, 0.0f);
    for (size_t t = 0; t < msd_.size(); t++)
        msd_[t] = (float)((d.sum_d2[3 * t + a0_] + d.sum_d2[3 * t + a1_]) / d.count[t]);
}

void Mean_Square_Displacement_2D::write(std::string filename) const
{
    std::cout << "\nWriting msd to file.";
    std::ofstream out(filename);
    out << "2-D mean square displacement data for " << plane_ << " plane created by AMDAT v." << kSemver << "\n";
    for (size_t t = 0; t < msd_.size(); t++) out << timetable_[t] << "\t" << msd_[t] << "\n";
}

Mean_Displacement::Mean_Displacement(System *sys) : Analysis(sys), system_(sys)
{
    timetable_ = sys->displacement_times();
    md_.assign(sys->show_n_timegaps(), Coordinate());
}

void Mean_Displacement::analyze(Trajectory_List *t_list)
{
    Displacement_Stats d;
    d.run(system_, t_list, 0, 0.0f);
    for (size_t t = 0; t < md_.size(); t++)
        md_[t] = Coordinate((float)(d.sum_d[3 * t] / d.count[t]), (float)(d.sum_d[3 * t + 1] / d.count[t]),
                            (float)(d.sum_d[3 * t + 2] / d.count[t]));
}

void Mean_Displacement::write(std::string filename) const
{
    std::cout << "\nWriting mean displacement to file " << filename << ".";
    std::ofstream out(filename);
    out << "Mean displacement data created by AMDAT v." << kSemver << "\n";
    for (size_t t = 0; t < md_.size(); t++)
        out << timetable_[t] << "\t" << md_[t].show_x() << "\t" << md_[t].show_y() << "\t" << md_[t].show_z()
            << "\n";
}

// ---------------------------------------------------------------------------
// Gaussian_Comparison (src/gaussian_comparison.cpp:100-332)

namespace {
constexpr double kGcPi = 3.14159265;  // src/gaussian_comparison.cpp:14
}

Gaussian_Comparison::Gaussian_Comparison(System *sys, const Non_Gaussian_Parameter *ngp, const Van_Hove_Self *vhs,
                                         const Mean_Square_Displacement *msd)
    : Analysis(sys), system_(sys)
{
    time_index_ = ngp->max();
    init(vhs, msd);
}

Gaussian_Comparison::Gaussian_Comparison(System *sys, int time_in, const Van_Hove_Self *vhs,
                                         const Mean_Square_Displacement *msd)
    : Analysis(sys), system_(sys)
{
    time_index_ = time_in;
    init(vhs, msd);
}

void Gaussian_Comparison::init(const Van_Hove_Self *vhs, const Mean_Square_Displacement *msd)
{
    if (time_index_ < 0 || time_index_ >= msd->show_n_times())
        throw Error("gaussian comparison: time index outside the time gaps", 0);
    n_bins_ = vhs->show_n_bins();
    bin_size_ = vhs->show_bin_size();
    mean_square_displacement_ = msd->show(time_index_);
    gaussian_approx_.assign(n_bins_, 0.0f);
    van_hove_.assign(n_bins_, 0.0f);
    // scale_van_hoves (src/gaussian_comparison.cpp:190-206)
    const float alpha = 3.0 / (2.0 * mean_square_displacement_);
    for (int b = 0; b < n_bins_; b++) {
        const float shellvolume = 4.0 / 3.0 * kGcPi *
                                  (std::pow((float(b) + 1.0) * bin_size_, 3.0) - std::pow(float(b) * bin_size_, 3.0));
        van_hove_[b] = vhs->show(time_index_, b) * shellvolume;
        gaussian_approx_[b] = std::pow(alpha / kGcPi, 3.0 / 2.0) *
                              std::exp(-alpha * std::pow((float(b) + .5) * bin_size_, 2.0)) * shellvolume;
    }
    // find_crossovers (src/gaussian_comparison.cpp:211-249)
    for (int b = 1; b < n_bins_; b++)
        if (gaussian_approx_[b] >= van_hove_[b]) {
            slowboundary_index_ = b - 1;
            break;
        }
    for (int b = slowboundary_index_ + 1; b < n_bins_; b++)
        if (gaussian_approx_[b] <= van_hove_[b]) {
            fastboundary_index_ = b;
            break;
        }
    if (fastboundary_index_ == 0) {
        std::cout << "\nWarning: program cannot identify two crossing points between self Van Hove and Gaussian "
                     "approximation self Van Hove.  Fractions of slow and fast particles will be incorrect.\n";
        errorstate_ = true;
    }
    slowboundary_ = std::pow((float(slowboundary_index_) + .5) * bin_size_, 2);
    fastboundary_ = std::pow((float(fastboundary_index_) - .5) * bin_size_, 2);
    // determine_fractions (src/gaussian_comparison.cpp:256-272)
    for (int b = 1; b <= slowboundary_index_; b++) {
        slowfraction_ += van_hove_[b];
        gaussian_slowfraction_ += gaussian_approx_[b];
    }
    fastfraction_ = 1.0 - slowfraction_;
    gaussian_fastfraction_ = 1.0 - gaussian_slowfraction_;
    for (int b = slowboundary_index_ + 1; b < fastboundary_index_; b++) {
        fastfraction_ -= van_hove_[b];
        gaussian_fastfraction_ -= gaussian_approx_[b];
    }
}

void Gaussian_Comparison::write(std::string filename) const
{
    std::cout << "\nWriting fractions of slow and fast particles and comparison between gaussian and actual self van "
                 "hove to file "
              << filename << ".";
    std::ofstream out(filename);
    if (errorstate_)
        out << "Warning: program cannot identify two crossing points between self Van Hove and Gaussian "
               "approximation self Van Hove.  Fractions of slow and fast particles will be incorrect.\n";
    out << "Gaussian comparison data created by AMDAT v." << kSemver << "\n";
    out << "t*\n" << system_->displacement_times()[time_index_] << "\n";
    out << "Slow_Cutoff: " << std::pow(slowboundary_, 0.5) << "\n";
    out << "Fast_Cutoff: " << std::pow(fastboundary_, 0.5) << "\n";
    out << "\t" << "Slow_Fraction" << "\t" << "Fast_Fraction\n";
    out << "Gaussian" << "\t" << gaussian_slowfraction_ << "\t" << gaussian_fastfraction_ << "\n";
    out << "Actual" << "\t" << slowfraction_ << "\t" << fastfraction_ << "\n";
    out << "Excess" << "\t" << slowfraction_ - gaussian_slowfraction_ << "\t" << fastfraction_ - gaussian_fastfraction_
        << "\n\n";
    out << "Self_Van_Hove*shellvolume\n";
    out << "radius\tgaussian\tactual\n";
    for (int b = 0; b < n_bins_; b++)
        out << (b + .5) * bin_size_ << "\t" << gaussian_approx_[b] << "\t" << van_hove_[b] << "\n";
}

// ---------------------------------------------------------------------------
// Displacement_Distribution (src/displacement_distribution.cpp:19-141),
// DebyeWaller_Dist (src/debyewaller_dist.cpp:14-141)

Displacement_Distribution::Displacement_Distribution(System *sys, float power, int bins, float maxvalue, float t)
    : Analysis(sys), system_(sys), power_(power), n_bins_(bins + 1), time_index_((int)t), max_value_(maxvalue)
{
    if (bins < 1 || bins > 16383) throw Error("distribution: number of bins must be 1..16383", 0);
    binsize_ = maxvalue / bins;
    const std::vector<float> times = sys->displacement_times();
    if (time_index_ < 0 || time_index_ >= (int)times.size()) throw Error("Requested timegap out of range.", 0);
    time_ = times[time_index_];
    distribution_.assign(n_bins_, 0.0f);
}

void Displacement_Distribution::analyze(Trajectory_List *t_list)
{
    amdat_pairs *p = system_->gap_pairs(time_index_);
    std::vector<unsigned long long> counts(n_bins_);
    double s2 = 0, s4 = 0, c = 0;
    int rc = amdat_displacement_dist(device_context(), system_->traj(), t_list->handle(), p, n_bins_ - 1,
                                     max_value_, counts.data(), &s2, &s4, &c);
    amdat_pairs_free(device_context(), p);
    check(rc, "displacement distribution");
    // postprocess_list: mean and distribution per atom (the 1/weighting
    // factors of the reference cancel); an empty selection gives 0/0 as there
    const double mean = s2 / c;
    mean_ = (float)mean;
    variance_ = (float)(s4 / c - mean * mean);
    for (int b = 0; b < n_bins_; b++) distribution_[b] = (float)((double)counts[b] / c);
}

void Displacement_Distribution::write(std::string filename) const
{
    std::ofstream out(filename);
    std::cout << "\nWriting distribution of particle displacement magnitudes to the " << power_ << " power to file "
              << filename << ".";
    out << "Particle displacement distribution to the " << power_ << " power created by AMDAT v." << kSemver << "\n";
    out << "Last bin includes overflow.";
    out << "time = " << time_ << "\n";
    out << "Mean\t" << mean_ << "\n";
    out << "Variance\t" << variance_ << "\n\n";
    for (int b = 0; b < n_bins_; b++) out << (float(b) + 0.5) * binsize_ << "\t" << distribution_[b] << "\n";
}

DebyeWaller_Dist::DebyeWaller_Dist(System *sys, int bins, float maxvalue, float t)
    : Displacement_Distribution(sys, 2.0f, bins, maxvalue, t)
{
}

void DebyeWaller_Dist::write(std::string filename) const
{
    std::ofstream out(filename);
    std::cout << "\nWriting distribution of mean-square dispalcements to file " << filename << ".";
    out << "MSD distribution created by AMDAT v." << kSemver << "\n";
    out << "MSD time = " << time_ << "\n";
    out << "Mean_stiffness\t" << mean_ << "\n";
    out << "Variance\t" << variance_ << "\n\n";
    for (int b = 0; b < n_bins_; b++) out << (float(b) + 0.5) * binsize_ << "\t" << distribution_[b] << "\n";
    out << "Last bin includes overflow.";
}

Van_Hove_Self::Van_Hove_Self(System *sys, int bin_count, float value_max)
    : Analysis(sys), system_(sys), n_bins_(bin_count), max_value_(value_max)
{
    if (bin_count < 1 || bin_count > 16384) throw Error("vhs: number of bins must be 1..16384", 0);
    bin_size_ = max_value_ / (float)n_bins_;
    timetable_ = sys->displacement_times();
    corr_.assign((size_t)sys->show_n_timegaps() * n_bins_, 0.0f);
}

void Van_Hove_Self::analyze(Trajectory_List *t_list)
{
    Displacement_Stats d;
    d.run(system_, t_list, n_bins_, max_value_);
    const int ntg = (int)timetable_.size();
    for (int b = 0; b < n_bins_; b++) {
        // src/space_time_correlation_function.cpp:345-360
        const float rshell = b * bin_size_;
        const float shellvolume = (float)((4.0 / 3.0) * kPi * (std::pow(rshell + bin_size_, 3.0) - std::pow(rshell, 3.0)));
        for (int t = 0; t < ntg; t++)
            corr_[(size_t)t * n_bins_ + b] =
                (float)((double)d.hist[(size_t)t * n_bins_ + b] / (d.count[t] * (double)shellvolume));
    }
}

void Van_Hove_Self::write(std::string filename) const
{
    std::ofstream out(filename);
    out << "Correlation data created by AMDAT v." << kSemver << "\n";
    out << n_bins_ << " bins\n";
    out << timetable_.size() << " times\n\n";
    std::cout << "\nWriting correlation function to file " << filename << ".";
    out << "\t";
    for (int b = 0; b < n_bins_; b++) out << bin_size_ / 2 + float(b) * bin_size_ << "\t";
    out << "\n";
    for (size_t t = 0; t < timetable_.size(); t++) {
        out << timetable_[t] << "\t";
        for (int b = 0; b < n_bins_ - 1; b++) out << corr_[t * n_bins_ + b] << "\t";
        out << corr_[t * n_bins_ + n_bins_ - 1] << "\n";
    }
}

}  // namespace amdat_hip
