// amdat_host.h -- C++ drop-in for AMDAT's reciprocal-space and
// time-correlation analyses, running on the C ABI of libamdat_hip.
//
// Class names, constructor signatures, analyze()/write() and the show_*
// accessors follow the reference's Analysis hierarchy so that code written
// against AMDAT's classes, and AMDAT input scripts using the msd, ngp, isfs,
// isf and structure_factor commands, run unchanged:
//   System                          src/system.h
//   Trajectory_List                 src/trajectory_list.h (static lists)
//   Wave_Vectors                    src/wave_vectors.h
//   Wave_Density                    src/wave_density.h
//   Structure_Factor                src/structure_factor.h
//   Intermediate_Scattering_Function src/intermediate_scattering_function.h
//   Incoherent_Scattering_Function  src/incoherent_scattering_function.h
//   Mean_Square_Displacement        src/mean_square_displacement.h
//   Non_Gaussian_Parameter          src/non_gaussian_parameter.h
//   Fast_Particles lists            src/fast_particles.h (time-varying)
//   Control                         src/control.h (the commands above)
// The arithmetic runs on the device; results are FP64 there and are
// stored as float here, as the reference stores them, before writing.
#pragma once

#include <cstdint>
#include <complex>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/amdat_hip.h"

namespace amdat_hip {

// Version string written into output headers ("... created by AMDAT v.<X>").
extern const char *kSemver;

// Error severities of the reference's Error(msg, code) (src/error.cpp:
// 51-146): code < 1 is severe (the program exits with that code), code <
// 10000 is moderate (the current command is skipped), otherwise minor (a
// warning).  Severe and moderate errors are thrown as Error.
class Error : public std::runtime_error {
  public:
    Error(const std::string &msg, int code) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }
    bool severe() const { return code_ < 1; }

  private:
    int code_;
};
void warn(const std::string &msg);

// Process-wide device context (one per host thread in the C ABI's sense).
amdat_ctx *device_context();

class Coordinate {
  public:
    Coordinate(float x = 0, float y = 0, float z = 0) : x_(x), y_(y), z_(z) {}
    float show_x() const { return x_; }
    float show_y() const { return y_; }
    float show_z() const { return z_; }
    float min() const { return x_ < y_ ? (x_ < z_ ? x_ : z_) : (y_ < z_ ? y_ : z_); }
    float length() const;

  private:
    float x_, y_, z_;
};

class System;
class Trajectory_List;

// Kinds an analysis reports through what_are_you() (src/analysis.h:27-56).
enum Analysis_Type {
    analysis,
    composition,
    correlation_2d,
    debyewaller_dist,
    displacement_dist,
    displacement_map,
    exptime_trajectory_list,
    fast_particles,
    gaussian_comparison,
    incoherent_scattering_function,
    intermediate_scattering_function,
    mean_displacement,
    mean_square_displacement,
    mean_square_displacement_2d,
    n_fold_order_parameter,
    non_gaussian_parameter,
    radial_debye_waller,
    rgtensor_stats,
    space_time_correlation_function,
    stiffness_dist,
    strings,
    edge_detector_timedependent,
    radial_distribution_function,
    displacement_list,
    distance_neighbor_list,
    neighbor_decorrelation_function,
    find_between
};

// Parent of the analyses (src/analysis.h:62-134): the virtual interface
// through which AMDAT's Control drives an analysis it holds as Analysis*
// (run_analysis(Analysis*, string), src/control.cpp:1316-1363).  A method a
// class does not implement prints the reference's message and does
// nothing.  The per-atom kernels of the reference's interface
// (listkernel, list_displacementkernel, atomkernel) have no counterpart:
// the loops they served run on the device inside analyze().
class Analysis {
  public:
    virtual ~Analysis() = default;
    virtual Analysis_Type what_are_you() { return analysis; }
    virtual void analyze(Trajectory_List *, Trajectory_List *);
    virtual void analyze(Trajectory_List *);
    // Trajectory_List_Bins targets (Bin_Static_Analysis / Bin_Dynamics_
    // Analysis call these per bin): in this implementation the Control
    // forms per-bin lists on the device and calls analyze() with them.
    virtual void bin_hook(Trajectory_List *, int, int, int);
    virtual void postprocess_bins() {}
    virtual void write(std::string) const;
    System *show_system() { return system_base_; }
    virtual bool isThreadSafe() { return false; }

  protected:
    Analysis() = default;
    explicit Analysis(System *sys) : system_base_(sys) {}
    Analysis(const Analysis &) = default;
    Analysis &operator=(const Analysis &) = default;
    virtual void preprocess() {}
    virtual void postprocess() {}
    virtual void postprocess_list() {}
    System *system_base_ = nullptr;
};

// Selection of atoms.  Static (Static_Trajectory_List: the same atoms at
// every frame, in the reference's species -> molecule -> type -> index
// order) or time-varying (Exptime_Trajectory_List and the Fast_Particles
// lists built on it: one list per exponential block, the list of frame t
// at list time int((t - frt) / n_exponential_steps);
// src/exptime_trajectory_list.cpp:29-56).
class Trajectory_List {
  public:
    Trajectory_List(System *sys, std::vector<int> atoms);
    // Takes ownership of a time-varying device list.
    Trajectory_List(System *sys, amdat_list *timed);
    // Takes ownership of a pair-keyed device list (amdat_list_create_paired):
    // the analyses over frame pairs take it; it has no atoms per frame.
    struct PairKeyed {};
    Trajectory_List(System *sys, amdat_list *paired, PairKeyed);
    ~Trajectory_List();
    Trajectory_List(const Trajectory_List &) = delete;
    Trajectory_List &operator=(const Trajectory_List &) = delete;
    // Atoms at frame timeii (src/trajectory_list.h:58).
    int show_n_trajectories(int timeii) const;
    bool time_varying() const { return n_times_ > 0; }
    bool pair_keyed() const { return pair_keyed_; }
    // Every list time's atoms, concatenated in list-time order (a device
    // list's atoms are copied to the host on first use).
    const std::vector<int> &atoms() const
    {
        load_atoms();
        return atoms_;
    }
    // The atoms at frame timeii, in list order.
    std::vector<int> atoms_at(int timeii) const;
    amdat_list *handle() const { return list_; }
    // Counts per list time: Trajectory_List::write_count
    // (src/trajectory_list.cpp:231-249) for a static list,
    // Exptime_Trajectory_List::write_count (src/exptime_trajectory_list.
    // cpp:93-114: block start time and count) for a time-varying one.
    void write_count(const std::string &filename) const;

  private:
    void load_atoms() const;
    System *system_;
    mutable std::vector<int> atoms_;
    mutable bool atoms_loaded_ = true;
    int n_entries_ = 0;
    int n_times_ = 0;
    bool pair_keyed_ = false;
    std::vector<int> conv_, off_;
    amdat_list *list_ = nullptr;
};

// Rectangular spatial bins of all atoms at every frame
// (Trajectory_List_Bins, src/trajectory_list_bins.cpp: "all" 207-279,
// "region" 280-394, assign_bins 949-1096), computed on the device.
class Trajectory_List_Bins {
  public:
    // all: the whole box of every frame
    Trajectory_List_Bins(System *sys, int xbins, int ybins, int zbins);
    // region: bounds 0 or outside the box fall back to the box
    Trajectory_List_Bins(System *sys, int xbins, int ybins, int zbins, float x_lo, float x_hi, float y_lo,
                         float y_hi, float z_lo, float z_hi);
    // Distance bins of the atoms of `binned`: from the nearest atom of
    // `cluster` (src/trajectory_list_bins.cpp:461-559), from a plane x, y or
    // z at `position` with direction average, above or below (:560-765), from
    // a point (:766-889).  n_bins is capped at int(reach / thickness) + 1.
    Trajectory_List_Bins(System *sys, float thickness, int n_bins, const Trajectory_List *binned,
                         const Trajectory_List *cluster);
    Trajectory_List_Bins(System *sys, float thickness, int n_bins, const Trajectory_List *binned,
                         const std::string &plane, float position, const std::string &direction);
    Trajectory_List_Bins(System *sys, float thickness, int n_bins, const Trajectory_List *binned,
                         const Coordinate &point);
    int show_n_xbins() const { return n_[0]; }
    int show_n_ybins() const { return n_[1]; }
    int show_n_zbins() const { return n_[2]; }
    ~Trajectory_List_Bins();
    Trajectory_List_Bins(const Trajectory_List_Bins &) = delete;
    Trajectory_List_Bins &operator=(const Trajectory_List_Bins &) = delete;
    // The atoms of list l that lie in bin (x, y, z), frame by frame, in atom
    // order (Trajectory_List::operator&&, src/trajectory_list.cpp:518-):
    // a time-varying list with one list time per frame, formed on the device.
    std::unique_ptr<Trajectory_List> intersect(const Trajectory_List *l, int xii, int yii, int zii) const;
    // intersect for every bin, in bin order ((x * ny + y) * nz + z).
    std::vector<std::unique_ptr<Trajectory_List>> intersect_all(const Trajectory_List *l) const;
    // Per bin, in bin order ((x * ny + y) * nz + z), the pair-keyed list of
    // Bin_Dynamics_Analysis (src/bin_dynamics_analysis.h:120-141): for
    // each frame pair of System::displacement_list(analysis) (fullblock 0,
    // with limit), the atoms of l at t0 that lie in the bin at t0 and at t1,
    // in atom order.
    std::vector<std::unique_ptr<Trajectory_List>> intersect_pairs(const Trajectory_List *l) const;

  private:
    void assign(const std::vector<float> &lo, const std::vector<float> &hi);
    void assign_distance(int kind, const float *point, int axis, int direction, const Trajectory_List *binned,
                         const Trajectory_List *cluster, float thickness, int n_bins);
    System *system_;
    int n_[3];
    amdat_cells *cells_ = nullptr;  // device cell table [T][N], -1 outside
};

class System {
  public:
    // Parse the system block (src/system.cpp:49-226): system type,
    // trajectory type (custom), file, time scheme, species/type table.
    explicit System(const std::vector<std::string> &block);
    // Build from arrays already in memory (frame-major SoA, file order).
    System(int n_exponentials, int n_exponential_steps, float exp_base, int frt, int first_exponent,
           float time_unit, int n_atoms, const float *x, const float *y, const float *z, const int16_t *ix,
           const int16_t *iy, const int16_t *iz, const float *box, const int *types);
    ~System();
    System(const System &) = delete;
    System &operator=(const System &) = delete;

    int show_n_timesteps() const { return n_timesteps_; }
    int show_n_timegaps() const { return n_timegaps_; }
    int show_n_exponentials() const { return n_exponentials_; }
    int show_n_exponential_steps() const { return n_exponential_steps_; }
    int show_frt() const { return frt_; }
    int show_n_atoms() const { return n_atoms_; }
    float show_exp_base() const { return exp_base_; }
    const float *show_timelist() const { return timelist_.data(); }
    float show_time(int timeii) const { return timelist_.at(timeii); }
    Coordinate min_box_dimensions() const;
    Coordinate size(int timeii) const;
    // Box lower bounds [T][3] (System::boundaries(t)[0], src/system.cpp);
    // the upper bounds are these plus size(t).
    std::vector<float> box_lo() const;
    void set_limit(int limit) { displacement_limit_ = limit; }
    int show_limit() const { return displacement_limit_; }

    // src/system.cpp:4599-4628
    std::vector<int> timegap_weighting(bool fullblock) const;
    // src/system.cpp:3763-3784
    std::vector<float> displacement_times() const;
    // (tg, t0, t1) in the order of src/system.cpp:4090-4179, as a device table.
    // Pairs of one time gap, first frame of each block only
    // (src/system.cpp:4349-4398), as a device table.
    amdat_pairs *gap_pairs(int time_index) const;
    amdat_pairs *displacement_pairs(bool fullblock, std::vector<int> *tg = nullptr,
                                    std::vector<int> *t0 = nullptr) const;
    // The same pairs on the host only.
    void displacement_pair_list(bool fullblock, std::vector<int> &tg, std::vector<int> &t0,
                                std::vector<int> &t1) const;
    // Pairs of one time gap as System::displacement_loop(analysis, traj,
    // timegap, fullblock) visits them (src/system.cpp:3972-4023), as a
    // device table.
    amdat_pairs *timegap_loop_pairs(int timegapii, bool fullblock) const;

    // Atom selections (create_list keywords all / type_system / species /
    // type_species); indices in the reference's list order.
    std::vector<int> select(const std::vector<std::string> &args) const;

    amdat_traj *traj() const { return traj_; }

  private:
    void init_timescheme(const std::vector<std::string> &args);
    void read_trajectory(const std::string &trajtype, const std::string &file);
    void finish();

    int n_exponentials_ = 1, n_exponential_steps_ = 0, frt_ = 0, first_exponent_ = 0;
    float exp_base_ = 1.0f, time_unit_ = 1.0f;
    int n_timesteps_ = 1, n_timegaps_ = 1, n_atoms_ = 0;
    int displacement_limit_ = 0;
    std::vector<float> timelist_;
    // host copies in file order
    std::vector<float> x_, y_, z_, box_;
    std::vector<int16_t> ix_, iy_, iz_;
    std::vector<int> type_;                     // atom type per atom
    std::vector<std::string> species_names_;
    std::vector<int> species_mols_;             // molecules per species
    std::vector<int> type_ids_;                 // types line
    std::vector<std::vector<int>> species_type_counts_;
    std::vector<int> list_order_;               // reference list order over all atoms
    std::vector<int> atom_species_;             // species index per atom (file order)
    amdat_traj *traj_ = nullptr;
    float xyz_box_[3] = {0, 0, 0};  // xyz: constant box lengths of the system block
};

class Wave_Vectors {
  public:
    // src/wave_vectors.cpp:77-181: plane "xyz", "xy", "xz", "yz", "x", "y"
    // or "z"; max_length_scale <= 0 means the smallest box dimension in the
    // plane.  Reads n_wavenumbers shells (99 in the reference: files
    // stem.002 .. stem.100) from the q-vector files named by AMDAT_WV3D,
    // AMDAT_WV2D or AMDAT_WV1D (the reference's WV3D / WV2D / WV1D paths);
    // n_wavenumbers > 99 opts into later files.
    Wave_Vectors(System *sys, const std::string &plane, float max_length_scale = 0, int n_wavenumbers = 99);
    ~Wave_Vectors();
    Wave_Vectors(const Wave_Vectors &) = delete;
    Wave_Vectors &operator=(const Wave_Vectors &) = delete;

    int show_n_wavenumbers() const { return (int)off_.size() - 1; }
    int vectorcount(int index) const { return off_[index + 1] - off_[index]; }
    std::vector<Coordinate> vectorlist(int index) const;
    float show_delta_wavenumber() const { return delta_wavenumber_; }
    float show_approx_wavenumber(int index) const { return approx_[index]; }
    float show_mean_wavenumber(int index) const;
    float show_stdev_wavenumber(int index) const;
    Coordinate show_mean_wavevector(int index) const;
    // length scale l of the integer vectors held: k = 2 pi n / l
    float show_length_scale() const { return length_scale_; }
    int show_denominator() const { return denominator_; }
    amdat_shells *handle() const { return shells_; }
    const std::vector<int> &offsets() const { return off_; }
    const std::vector<int> &integer_vectors() const { return nvec_; }

  private:
    System *system_;
    float length_scale_, wavegrid_spacing_, delta_wavenumber_;
    int denominator_ = 1;
    std::vector<float> approx_;
    std::vector<int> off_, nvec_;
    amdat_shells *shells_ = nullptr;
};

class Wave_Density : public Analysis {
  public:
    // src/wave_density.cpp:46-80: every shell of wv; stf: only the first
    // frame of each exponential block, t = n_exponential_steps * b (the
    // static S(q) form, :316-323), other frames hold zero.
    Wave_Density(System *sys, const Wave_Vectors *wv, bool stf = false);
    // src/wave_density.cpp:83-124: shells inner .. outer.
    Wave_Density(System *sys, const Wave_Vectors *wv, int inner, int outer);
    // Copies share the computed device density, which is never modified
    // after analyze() (a later analyze() or set() gives the object a new
    // one), so a copy behaves as the reference's deep copy.
    Wave_Density(const Wave_Density &other) = default;
    Wave_Density &operator=(const Wave_Density &other) = default;
    // src/wave_density.cpp:226-299: new range, density cleared.
    void set(System *sys, const Wave_Vectors *wv);
    void set(System *sys, const Wave_Vectors *wv, int inner, int outer);
    // src/wave_density.cpp:313-336: rho_k(t) for every frame (or the block
    // starts), added to the density and atom counts of earlier calls since
    // construction or set(), as the reference adds them.
    void analyze(Trajectory_List *t_list) override;
    using Analysis::analyze;
    // The reference's wd2 = wd1.
    void copy_from(const Wave_Density &other) { *this = other; }
    std::complex<double> show_density(int time, int wavenumberindex, int vectorindex) const;
    double show_density_re(int time, int wavenumberindex, int vectorindex) const;
    double show_density_im(int time, int wavenumberindex, int vectorindex) const;
    int show_n_atoms(int timeii) const { return n_atoms_.at(timeii); }
    // Atom-frames summed into the density (src/wave_density.cpp:339-365).
    int show_n_atoms_looped() const;
    int show_first() const { return first_; }
    int show_last() const { return last_; }
    const Wave_Vectors *show_wavevectors() const { return wv_; }
    System *show_system() const { return system_; }
    const amdat_rho *device() const { return rho_.get(); }

  private:
    void fetch() const;
    System *system_;
    const Wave_Vectors *wv_;
    int first_, last_;
    bool sf_ = false;
    std::vector<int> n_atoms_;  // atoms of the list at each frame
    std::shared_ptr<amdat_rho> rho_;
    mutable std::shared_ptr<std::vector<double>> host_;
};

class Structure_Factor : public Analysis {
  public:
    Structure_Factor(System *sys, const Wave_Vectors *wv, int timescheme = -1);
    void analyze(Trajectory_List *t_list) override;
    void analyze(Trajectory_List *t_list, Trajectory_List *t_list2) override;
    void write(std::string filename) const override;
    float show(int wavenumberii) const { return structure_factor_[wavenumberii]; }
    // Frames the time scheme selects (src/analysis_onetime.cpp:22-52).
    std::vector<int> frames() const;

  private:
    System *system_;
    const Wave_Vectors *wv_;
    int time_scheme_;
    std::vector<float> structure_factor_;
};

class Intermediate_Scattering_Function : public Analysis {
  public:
    // src/intermediate_scattering_function.cpp:52-147: every shell of the
    // densities (which must share wave vectors, system and shell range)
    Intermediate_Scattering_Function(System *sys, Wave_Density *wd1, Wave_Density *wd2, bool parallel = true);
    // :149-235: one shell, read as the range constructor reads wn1 = wn2 =
    // wn (the density's first slot with shell wn's vector count);
    // physics_correct = true reads shell wn
    Intermediate_Scattering_Function(System *sys, Wave_Density *wd1, Wave_Density *wd2, int wn,
                                     bool parallel = true, bool physics_correct = false);
    // :237-322: shells wn1 .. wn2, as the reference reads them: column
    // wn - wn1 sums the density's slot wn - wn1 (counted from the density's
    // first shell) over the vector count of shell wn, :556-562.  That is
    // shell wn when wn1 is the density's first shell (the isf command's
    // case); otherwise it is the slot's shell, over the first vectors of it,
    // and a slot holding fewer vectors than shell wn is refused (the
    // reference reads past it).  physics_correct = true reads shell wn.
    Intermediate_Scattering_Function(System *sys, Wave_Density *wd1, Wave_Density *wd2, int wn1, int wn2,
                                     bool parallel = true, bool physics_correct = false);
    // :324-412: every shell at one time gap only (pairs of System::
    // displacement_loop(analysis, traj, timegap, fullblock), src/system.cpp:
    // 3972-4023); time gap 0 gives S(k) over the block starts.  Other rows
    // stay 0.
    Intermediate_Scattering_Function(System *sys, int timegapii, Wave_Density *wd1, Wave_Density *wd2,
                                     bool parallel = true);
    Analysis_Type what_are_you() override { return intermediate_scattering_function; }
    void write(std::string filename) const override;
    float show(int timeii, int binii) const { return correlation_[timeii][binii]; }

  private:
    void compute(Wave_Density *wd1, Wave_Density *wd2, amdat_pairs *pairs, int only_timegap,
                 bool slot_quirk = false);
    System *system_;
    const Wave_Vectors *wv_;
    int first_bin_, last_bin_;
    std::vector<float> timetable_;
    std::vector<std::vector<float>> correlation_;
};

class Incoherent_Scattering_Function : public Analysis {
  public:
    // src/incoherent_scattering_function.cpp:176-212
    Incoherent_Scattering_Function(System *sys, const Wave_Vectors *wv, int inner, int outer, bool fblock);
    Analysis_Type what_are_you() override { return incoherent_scattering_function; }
    void analyze(Trajectory_List *t_list) override;
    using Analysis::analyze;
    void write(std::string filename) const override;
    float show(int timeii, int binii) const { return correlation_[timeii][binii]; }

  private:
    System *system_;
    const Wave_Vectors *wv_;
    int first_bin_, last_bin_;
    bool fullblock_;
    std::vector<float> timetable_;
    std::vector<std::vector<float>> correlation_;
};

class Mean_Square_Displacement : public Analysis {
  public:
    explicit Mean_Square_Displacement(System *sys);
    Analysis_Type what_are_you() override { return mean_square_displacement; }
    void analyze(Trajectory_List *t_list) override;
    using Analysis::analyze;
    void write(std::string filename) const override;
    float show(int t) const { return msd_[t]; }
    double show_f64(int t) const { return msd64_[t]; }
    int show_n_times() const { return (int)msd_.size(); }

  private:
    friend class Non_Gaussian_Parameter;
    System *system_;
    std::vector<float> timetable_, msd_;
    std::vector<double> msd64_, s4_, count_;
};

// Displacement statistics shared by the three analyses below: one device
// pass (amdat_displacements) over the fullblock = false pairs.
struct Displacement_Stats {
    std::vector<double> sum_d, sum_d2, count;  // [gap][3], [gap][3], [gap]
    std::vector<unsigned long long> hist;      // [gap][bins] when bins > 0
    void run(System *sys, Trajectory_List *t_list, int n_bins, float max_range);
};

// src/mean_square_displacement_2d.cpp: MSD in the xy, xz or yz plane.
class Mean_Square_Displacement_2D : public Analysis {
  public:
    Mean_Square_Displacement_2D(System *sys, const std::string &plane);
    Analysis_Type what_are_you() override { return mean_square_displacement_2d; }
    void analyze(Trajectory_List *t_list) override;
    using Analysis::analyze;
    void write(std::string filename) const override;
    float show(int t) const { return msd_[t]; }

  private:
    System *system_;
    std::string plane_;
    int a0_ = 0, a1_ = 1;
    std::vector<float> timetable_, msd_;
};

// src/mean_displacement.cpp: mean displacement vector per time gap.
class Mean_Displacement : public Analysis {
  public:
    explicit Mean_Displacement(System *sys);
    Analysis_Type what_are_you() override { return mean_displacement; }
    void analyze(Trajectory_List *t_list) override;
    using Analysis::analyze;
    void write(std::string filename) const override;
    Coordinate show(int t) const { return md_[t]; }

  private:
    System *system_;
    std::vector<float> timetable_;
    std::vector<Coordinate> md_;
};

// src/displacement_distribution.cpp: distribution of the squared
// displacement at one time gap (the reference ignores `power` there too),
// the last bin taking the overflow; mean and variance of the squared
// displacement.  DebyeWaller_Dist (src/debyewaller_dist.cpp) computes the
// same and writes another header.
class Displacement_Distribution : public Analysis {
  public:
    Displacement_Distribution(System *sys, float power, int bins, float maxvalue, float t);
    Analysis_Type what_are_you() override { return displacement_dist; }
    void analyze(Trajectory_List *t_list) override;
    using Analysis::analyze;
    void write(std::string filename) const override;
    float show_mean() const { return mean_; }
    float show_variance() const { return variance_; }
    float show(int bin) const { return distribution_.at(bin); }

  protected:
    System *system_;
    float power_;
    int n_bins_, time_index_;
    float max_value_, binsize_, time_;
    float mean_ = 0, variance_ = 0;
    std::vector<float> distribution_;
};

class DebyeWaller_Dist : public Displacement_Distribution {
  public:
    DebyeWaller_Dist(System *sys, int bins, float maxvalue, float t);
    Analysis_Type what_are_you() override { return debyewaller_dist; }
    void write(std::string filename) const override;
};

// src/van_hove_self.cpp + src/space_time_correlation_function.cpp: self
// part of the van Hove function, counts of |dr| per bin over atom pairs /
// (weighting * shell volume).
class Van_Hove_Self : public Analysis {
  public:
    Van_Hove_Self(System *sys, int bin_count, float value_max);
    Analysis_Type what_are_you() override { return space_time_correlation_function; }
    void analyze(Trajectory_List *t_list) override;
    using Analysis::analyze;
    void write(std::string filename) const override;
    float show(int t, int b) const { return corr_[(size_t)t * n_bins_ + b]; }
    int show_n_bins() const { return n_bins_; }
    float show_bin_size() const { return bin_size_; }

  private:
    System *system_;
    int n_bins_;
    float max_value_, bin_size_;
    std::vector<float> timetable_, corr_;
};

class Non_Gaussian_Parameter;

// src/gaussian_comparison.cpp: the self van Hove at one time gap against
// its Gaussian approximation of the same MSD (both times the shell
// volume); the crossings give the slow and fast boundaries (squared
// displacements) and the fractions.  Host arithmetic on the device-computed
// van Hove and MSD, in the reference's types and order.
class Gaussian_Comparison : public Analysis {
  public:
    // time index: the non-Gaussian parameter's maximum (src/gaussian_comparison.cpp:126-149)
    Gaussian_Comparison(System *sys, const Non_Gaussian_Parameter *ngp, const Van_Hove_Self *vhs,
                        const Mean_Square_Displacement *msd);
    // given time index (src/gaussian_comparison.cpp:160-185)
    Gaussian_Comparison(System *sys, int time_in, const Van_Hove_Self *vhs, const Mean_Square_Displacement *msd);
    Analysis_Type what_are_you() override { return gaussian_comparison; }
    void write(std::string filename) const override;
    int show_time_index() const { return time_index_; }
    float show_fastboundary() const { return fastboundary_; }
    float show_slowboundary() const { return slowboundary_; }
    float show_slowfraction() const { return slowfraction_; }
    float show_fastfraction() const { return fastfraction_; }

  private:
    void init(const Van_Hove_Self *vhs, const Mean_Square_Displacement *msd);
    System *system_;
    int n_bins_ = 0, time_index_ = 0;
    float bin_size_ = 0, mean_square_displacement_ = 0;
    std::vector<float> gaussian_approx_, van_hove_;
    int slowboundary_index_ = 0, fastboundary_index_ = 0;
    float slowboundary_ = 0, fastboundary_ = 0;
    float slowfraction_ = 0, fastfraction_ = 0, gaussian_slowfraction_ = 0, gaussian_fastfraction_ = 0;
    bool errorstate_ = false;
};

class Non_Gaussian_Parameter : public Analysis {
  public:
    Non_Gaussian_Parameter(System *sys, const Mean_Square_Displacement *msd);
    Analysis_Type what_are_you() override { return non_gaussian_parameter; }
    void analyze(Trajectory_List *t_list) override;
    using Analysis::analyze;
    void write(std::string filename) const override;
    int max() const;
    float show(int t) const { return ngp_[t]; }

  private:
    System *system_;
    const Mean_Square_Displacement *msd_;
    std::vector<float> timetable_, ngp_;
};

// Input-script interpreter for the commands above (src/control.cpp:108-531
// for the script structure).  Runs a script file; returns the exit code.
class Control {
  public:
    explicit Control(const std::string &input_file);
    int run();

  private:
    std::string read_line();
    void create_list(const std::vector<std::string> &args);
    void msd(const std::vector<std::string> &args);
    void ngp(const std::vector<std::string> &args);
    void msd_2d(const std::vector<std::string> &args);
    void mean_displacement(const std::vector<std::string> &args);
    void vhs(const std::vector<std::string> &args);
    void u2dist(const std::vector<std::string> &args);
    void displacement_dist(const std::vector<std::string> &args);
    void isfs(const std::vector<std::string> &args);
    void isf(const std::vector<std::string> &args);
    void structure_factor(const std::vector<std::string> &args);
    void find_fast_fixedthreshold(const std::vector<std::string> &args);
    void gaussian_comparison(const std::vector<std::string> &args);
    void find_fast(const std::vector<std::string> &args);
    void add_list(const std::string &name, std::unique_ptr<Trajectory_List> l);
    Trajectory_List *target(const std::string &line);
    // src/control.cpp:1316-1363
    void run_analysis(Analysis *analyzer, const std::string &setline);
    // Run fn(file suffix, list) for a "list <name>" target (suffix "") or,
    // for a "bin_list <bins> <list> [persistence]" target, once per bin
    // with the suffix Bin_Static_Analysis::write forms (src/bin_static_
    // analysis.h:114-140).  For a dynamic analysis with persistence 1
    // (src/control.h:353-366) each bin's list is pair-keyed, as
    // Bin_Dynamics_Analysis forms it, and the suffix is its write's
    // (src/bin_dynamics_analysis.h:145-175).
    void for_each_target(const std::string &line,
                         const std::function<void(const std::string &, Trajectory_List *)> &fn,
                         bool dynamic = false);
    void create_bin_list(const std::vector<std::string> &args);

    std::vector<std::string> lines_;
    size_t cursor_ = 0;
    std::unique_ptr<System> system_;
    std::vector<std::pair<std::string, std::unique_ptr<Trajectory_List>>> lists_;
    std::unique_ptr<Van_Hove_Self> vhs_;               // the last vhs command's result
    std::unique_ptr<Gaussian_Comparison> gaussian_;    // the last gaussian_comparison
    std::vector<std::pair<std::string, std::unique_ptr<Trajectory_List_Bins>>> bin_lists_;
};

std::vector<std::string> tokenize(const std::string &line);

}  // namespace amdat_hip
