// control.cpp -- input-script interpreter and command-line entry of the
// drop-in layer, for the commands of the reciprocal-space and
// time-correlation path.  Script structure: src/control.cpp:125-531
// (comments with #, a system block, then command lines each followed by
// their target lines); command syntax: src/control.cpp:1860-1916 (limit,
// create_list), 2163-2210 (msd), 2655-2758 (structure_factor), 3107-3192
// (isf), 3198-3238 (isfs), 3345-3370 (ngp), 2219-2271 (msd_2d,
// mean_displacement), 2517-2550 (vhs), 3466-3496 (find_fast_fixedthreshold),
// 3247-3273 (u2dist), 3312-3339 (displacement_dist), 3377-3420
// (gaussian_comparison), 3430-3458 (find_fast), 3841-4003 (create_bin_list),
// and bin_list targets (src/control.h:269-391, Bin_Static_Analysis).
#include <cstdlib>
#include <functional>
#include <ctime>
#include <fstream>
#include <iostream>

#include "amdat_host.h"

namespace amdat_hip {

Control::Control(const std::string &input_file)
{
    std::ifstream in(input_file);
    if (!in) throw Error("Error opening input file " + input_file, 0);
    std::string line;
    bool skipping = false;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        const size_t h = line.find('#');
        if (h != std::string::npos) line = line.substr(0, h);
        auto t = tokenize(line);
        if (!t.empty() && t[0] == "skip") { skipping = true; continue; }
        if (!t.empty() && t[0] == "endskip") { skipping = false; continue; }
        if (skipping) continue;
        lines_.push_back(line);
    }
}

std::string Control::read_line()
{
    while (cursor_ < lines_.size()) {
        const std::string &l = lines_[cursor_++];
        if (!tokenize(l).empty()) return l;
    }
    throw Error("unexpected end of input file", 0);
}

Trajectory_List *Control::target(const std::string &line)
{
    auto t = tokenize(line);
    if (t.size() != 2 || t[0] != "list")
        throw Error("target '" + line + "' is not supported by this build (list <name> only)", 1);
    for (auto &p : lists_)
        if (p.first == t[1]) return p.second.get();
    throw Error("Trajectory list '" + t[1] + "' not found.", 0);
}

// Bin_Static_Analysis::write (src/bin_static_analysis.h:123-137) forms
// each bin's index strings with one stringstream per axis that is rewound
// (seekp(0)) but never cleared, so a shorter index overwrites the first
// characters of a longer earlier one ("02" after "12").  Kept for identical
// file names.
namespace {
struct RewoundIndex {
    std::string buf;
    const std::string &put(int v)
    {
        const std::string s = std::to_string(v);
        if (s.size() > buf.size()) buf.resize(s.size());
        buf.replace(0, s.size(), s);
        return buf;
    }
};
}  // namespace

void Control::for_each_target(const std::string &line,
                              const std::function<void(const std::string &, Trajectory_List *)> &fn,
                              bool dynamic)
{
    auto t = tokenize(line);
    if (!t.empty() && t[0] == "bin_list") {
        // src/control.h:340-380: bin_list <bin list> <list> [persistence];
        // persistence counts for the dynamic analyses only (src/control.h:362)
        if (t.size() < 3) throw Error("Insufficient number of arguments in bin_list target line.", 0);
        const bool persist = dynamic && t.size() == 4 && std::atoi(t[3].c_str()) != 0;
        const Trajectory_List_Bins *bins = nullptr;
        for (auto &p : bin_lists_)
            if (p.first == t[1]) bins = p.second.get();
        if (!bins) throw Error("Bin list '" + t[1] + "' not found.", 0);
        Trajectory_List *l = target("list " + t[2]);
        // every bin's list at once, formed on the device
        std::vector<std::unique_ptr<Trajectory_List>> bin_lists =
            persist ? bins->intersect_pairs(l) : bins->intersect_all(l);
        const char *sep = persist ? "." : "_";
        RewoundIndex sx, sy, sz;
        for (int x = 0; x < bins->show_n_xbins(); x++)
            for (int y = 0; y < bins->show_n_ybins(); y++)
                for (int z = 0; z < bins->show_n_zbins(); z++) {
                    std::cout << "Bin (" << x << "," << y << "," << z << ")" << "\r";
                    std::cout.flush();
                    const std::string suffix =
                        "_bin_" + sx.put(x) + sep + sy.put(y) + sep + sz.put(z) + ".bindata";
                    fn(suffix, bin_lists[(x * bins->show_n_ybins() + y) * bins->show_n_zbins() + z].get());
                }
        std::cout << std::endl;
        return;
    }
    fn("", target(line));
}

// create_bin_list <name>, then "all <nx> <ny> <nz>" or "region <nx> <ny>
// <nz> <xlo> <xhi> <ylo> <yhi> <zlo> <zhi>" (src/control.cpp:3841-4003).
void Control::create_bin_list(const std::vector<std::string> &args)
{
    if (args.size() != 2) throw Error("create_bin_list expects one list name", 0);
    std::cout << "\nCreating Binned Trajectory List";
    auto a = tokenize(read_line());
    if (a.empty()) throw Error("No set type found.", 0);
    std::unique_ptr<Trajectory_List_Bins> b;
    auto I = [&](int k) { return std::atoi(a[k].c_str()); };
    auto F = [&](int k) { return (float)std::atof(a[k].c_str()); };
    if (a[0] == "all") {
        if (a.size() != 4) throw Error("Incorrect number of arguments for create_bin_list all.", 0);
        b = std::make_unique<Trajectory_List_Bins>(system_.get(), I(1), I(2), I(3));
    } else if (a[0] == "region") {
        if (a.size() != 10) throw Error("Incorrect number of arguments for create_bin_list region.", 0);
        b = std::make_unique<Trajectory_List_Bins>(system_.get(), I(1), I(2), I(3), F(4), F(5), F(6), F(7), F(8),
                                                   F(9));
    } else if (a[0] == "distance" && a.size() > 1) {
        // src/control.cpp:3915-3985
        auto L = [&](int k) { return target("list " + a[k]); };
        if (a[1] == "trajectory") {
            // distance trajectory <list to bin> <list to measure from> <thickness> <n_bins>
            if (a.size() != 6) throw Error("Incorrect number of arguments for create_bin_list distance trajectory.", 0);
            b = std::make_unique<Trajectory_List_Bins>(system_.get(), F(4), I(5), L(2), L(3));
        } else if (a[1] == "point") {
            // distance point <list to bin> <x> <y> <z> <thickness> <n_bins>
            if (a.size() != 8) throw Error("Incorrect number of arguments for create_bin_list distance point.", 0);
            b = std::make_unique<Trajectory_List_Bins>(system_.get(), F(6), I(7), L(2), Coordinate(F(3), F(4), F(5)));
        } else if (a[1] == "plane") {
            // distance plane <direction> <list to bin> <plane> <position> <thickness> <n_bins>
            if (a.size() != 8) throw Error("Incorrect number of arguments for create_bin_list distance plane.", 0);
            b = std::make_unique<Trajectory_List_Bins>(system_.get(), F(6), I(7), L(3), a[4], F(5), a[2]);
        } else {
            throw Error("distance binning type " + a[1] + " not recognized, use trajectory, point or plane.", 0);
        }
    } else {
        throw Error("bin_list of type " + a[0] + " invalid", 0);
    }
    for (auto &p : bin_lists_)
        if (p.first == args[1]) {
            p.second = std::move(b);
            return;
        }
    bin_lists_.emplace_back(args[1], std::move(b));
}

static std::vector<std::string> strip_store_flag(std::vector<std::string> a)
{
    std::vector<std::string> out;
    for (auto &s : a)
        if (s.rfind("\\s", 0) != 0) out.push_back(s);
    return out;
}

void Control::create_list(const std::vector<std::string> &args)
{
    if (args.size() != 2) throw Error("create_list expects one list name", 0);
    auto sel = tokenize(read_line());
    auto atoms = system_->select(sel);
    lists_.emplace_back(args[1], std::make_unique<Trajectory_List>(system_.get(), atoms));
    std::cout << "\nTrajectory list " << args[1] << " created with " << atoms.size() << " trajectories." << std::endl;
}

void Control::msd(const std::vector<std::string> &args)
{
    if (args.size() != 2) throw Error("msd expects an output file", 0);
    const std::string runline = read_line();
    std::cout << "\nCalculating mean square displacement.\n";
    const time_t start = time(nullptr);
    for_each_target(
        runline,
        [&](const std::string &suffix, Trajectory_List *l) {
            Mean_Square_Displacement m(system_.get());
            m.analyze(l);
            m.write(args[1] + suffix);
        },
        true);
    std::cout << "\nCalculated mean square displacement in " << time(nullptr) - start << " seconds." << std::endl;
}

// msd_2d <file> <plane> (src/control.cpp:2219-2243)
void Control::msd_2d(const std::vector<std::string> &args)
{
    if (args.size() != 3) throw Error("msd_2d expects an output file and a plane", 0);
    const std::string runline = read_line();
    std::cout << "\nCalculating mean square displacement.";
    const time_t start = time(nullptr);
    for_each_target(
        runline,
        [&](const std::string &suffix, Trajectory_List *l) {
            Mean_Square_Displacement_2D m(system_.get(), args[2]);
            m.analyze(l);
            m.write(args[1] + suffix);
        },
        true);
    std::cout << "\nCalculated mean square displacement in " << time(nullptr) - start << " seconds." << std::endl;
}

// mean_displacement <file> (src/control.cpp:2249-2271)
void Control::mean_displacement(const std::vector<std::string> &args)
{
    if (args.size() != 2) throw Error("mean_displacement expects an output file", 0);
    const std::string runline = read_line();
    std::cout << "\nCalculating mean displacement.\n";
    const time_t start = time(nullptr);
    for_each_target(
        runline,
        [&](const std::string &suffix, Trajectory_List *l) {
            Mean_Displacement m(system_.get());
            m.analyze(l);
            m.write(args[1] + suffix);
        },
        true);
    std::cout << "\nCalculated mean displacement in " << time(nullptr) - start << " seconds.";
}

// vhs <file> <max_range> <n_bins> (src/control.cpp:2517-2550)
void Control::vhs(const std::vector<std::string> &args)
{
    if (args.size() != 4) throw Error("vhs expects an output file, a maximum range and a bin count", 0);
    Trajectory_List *l = target(read_line());
    auto v = std::make_unique<Van_Hove_Self>(system_.get(), std::atoi(args[3].c_str()),
                                             (float)std::atof(args[2].c_str()));
    std::cout << "\nCalculating self part of Van Hove correlation function.";
    const time_t start = time(nullptr);
    v->analyze(l);
    std::cout << "\nCalculated self Van Hove in " << time(nullptr) - start << " seconds.";
    v->write(args[1]);
    vhs_ = std::move(v);  // kept for gaussian_comparison (src/control.cpp:2541-2546)
}

// gaussian_comparison <file> [time index], then the target line
// (src/control.cpp:3377-3420): MSD and NGP of the target, then the last
// vhs against its Gaussian approximation at the NGP's maximum or the given
// time index.
void Control::gaussian_comparison(const std::vector<std::string> &args)
{
    if (!vhs_)
        throw Error("self van hove not yet defined.  Self van hove must be calculated prior to compare_gaussian "
                    "calculation.",
                    0);
    if (args.size() != 2 && args.size() != 3)
        throw Error("Incorrect number of arguments for gaussian comparison method.", 0);
    const std::string runline = read_line();
    std::cout << "\n" << runline;
    Trajectory_List *l = target(runline);
    Mean_Square_Displacement msd(system_.get());
    msd.analyze(l);
    std::cout << "\nCalculating non-Gaussian parameter";
    Non_Gaussian_Parameter ngpar(system_.get(), &msd);
    ngpar.analyze(l);
    std::cout << "\nComparing self Van Hove with Gaussian approximation.";
    if (args.size() == 2)
        gaussian_ = std::make_unique<Gaussian_Comparison>(system_.get(), &ngpar, vhs_.get(), &msd);
    else
        gaussian_ = std::make_unique<Gaussian_Comparison>(system_.get(), std::atoi(args[2].c_str()), vhs_.get(), &msd);
    gaussian_->write(args[1]);
}

// find_fast <list> <count file>, then the target line (src/control.cpp:
// 3430-3458): Fast_Particles with the last Gaussian comparison's time index
// and fast boundary (src/fast_particles.cpp:38-56).
void Control::find_fast(const std::vector<std::string> &args)
{
    if (args.size() < 3) throw Error("find_fast expects a list name and a count file", 0);
    const std::string runline = read_line();
    std::cout << "\n" << runline;
    if (!gaussian_)
        throw Error("A guassian comparison analysis myst be defined prior to calculating fast particle list.", 0);
    std::cout << "\nIdentifying 'fast' particles.";
    Trajectory_List *in = target(runline);
    System *s = system_.get();
    amdat_list *out = nullptr;
    int rc = amdat_list_select_displacement(device_context(), s->traj(), in->handle(), s->show_n_exponentials(),
                                            s->show_n_exponential_steps(), s->show_frt(), s->show_limit(),
                                            gaussian_->show_time_index(), AMDAT_SELECT_FAST,
                                            gaussian_->show_fastboundary(), 0.0f, &out);
    if (rc < 0) throw Error(std::string("fast particles: ") + amdat_last_error(), 0);
    auto l = std::make_unique<Trajectory_List>(s, out);
    l->write_count(args[2]);
    add_list(args[1], std::move(l));
}

// Control::add_trajectorylist (src/control.cpp:1417-1429): an existing name
// is kept.
void Control::add_list(const std::string &name, std::unique_ptr<Trajectory_List> l)
{
    for (auto &p : lists_)
        if (p.first == name) {
            std::cout << "\nWarning: trajectory_list " << name
                      << " not created because a trajectory_list with this name already exists. Replacement of a "
                         "trajectory_list requires that you first delete the existing list with the same name.\n";
            return;
        }
    lists_.emplace_back(name, std::move(l));
}

// u2dist <file> <n_bins> <maxvalue> <time index> (src/control.cpp:3247-3273)
void Control::u2dist(const std::vector<std::string> &args)
{
    if (args.size() != 5) throw Error("Incorrect number of arguments for u2dist.", 0);
    const std::string runline = read_line();
    std::cout << "\n" << runline;
    DebyeWaller_Dist d(system_.get(), std::atoi(args[2].c_str()), (float)std::atof(args[3].c_str()),
                       (float)std::atoi(args[4].c_str()));
    d.analyze(target(runline));
    d.write(args[1]);
}

// displacement_dist <file> <power> <n_bins> <maxvalue> <time index>
// (src/control.cpp:3312-3339)
void Control::displacement_dist(const std::vector<std::string> &args)
{
    if (args.size() != 6) throw Error("Incorrect number of arguments for displacement_dist.", 0);
    const std::string runline = read_line();
    std::cout << "\n" << runline;
    Displacement_Distribution d(system_.get(), (float)std::atof(args[2].c_str()), std::atoi(args[3].c_str()),
                                (float)std::atof(args[4].c_str()), (float)std::atoi(args[5].c_str()));
    d.analyze(target(runline));
    d.write(args[1]);
}

void Control::ngp(const std::vector<std::string> &args)
{
    if (args.size() != 2) throw Error("ngp expects an output file", 0);
    Trajectory_List *l = target(read_line());
    Mean_Square_Displacement m(system_.get());
    m.analyze(l);
    Non_Gaussian_Parameter g(system_.get(), &m);
    g.analyze(l);
    g.write(args[1]);
    std::cout << "\n Peak time index of non-Gaussian parameter is " << g.max() << ".";
}

void Control::isfs(const std::vector<std::string> &args)
{
    // isfs <file> <inner> <outer> <plane> <max_length_scale> [fullblock]
    if (args.size() != 6 && args.size() != 7) throw Error("Incorrect number of arguments for isfs.", 0);
    const int inner = std::atoi(args[2].c_str()), outer = std::atoi(args[3].c_str());
    const float ell = (float)std::atof(args[5].c_str());
    const bool fb = args.size() == 7 ? std::atoi(args[6].c_str()) != 0 : false;
    const std::string runline = read_line();
    const time_t start = time(nullptr);
    Wave_Vectors wv(system_.get(), args[4], ell);
    for_each_target(
        runline,
        [&](const std::string &suffix, Trajectory_List *l) {
            // Bin_Dynamics_Analysis::analyze loops the pairs of
            // displacement_list(this), fullblock 0 (src/bin_dynamics_analysis.h:99)
            Incoherent_Scattering_Function f(system_.get(), &wv, inner, outer, l->pair_keyed() ? false : fb);
            f.analyze(l);
            f.write(args[1] + suffix);
        },
        true);
    std::cout << "\nSelf Intermediate Scattering Function calculated in " << time(nullptr) - start
              << " seconds.\n";
}

// src/control.cpp:1316-1363: run an analysis held as Analysis* on the
// target of a "list <name>" line (the reference's older set-based targets
// are not built).
void Control::run_analysis(Analysis *analyzer, const std::string &setline)
{
    auto t = tokenize(setline);
    if (t.empty()) throw Error("No atom set command found.", 5);
    if (t[0] != "list")
        throw Error("target '" + setline + "' is not supported by this build (list <name> only)", 1);
    if (t.size() != 2) throw Error("Incorrect number of arguments for list.", 0);
    for (auto &p : lists_)
        if (p.first == t[1]) {
            analyzer->analyze(p.second.get());
            return;
        }
    std::cout << "\nTrajectory list '" << t[1] << "' not found.";
}

void Control::isf(const std::vector<std::string> &args)
{
    // isf <file> auto <symmetric|asymmetric> <plane> <fullblock> <inner> <outer>
    // isf <file> manual <symmetric|asymmetric> <plane> <fullblock> <-> <S(k) file>
    // (src/control.cpp:3107-3192)
    if (args.size() < 6) throw Error("Incorrect number of arguments for isf.", 0);
    const std::string method = args[2], symmetry = args[3], plane = args[4];
    const bool fb = std::atoi(args[5].c_str()) != 0;
    std::cout << "\nDetermining wave vectors.";
    std::cout.flush();
    Wave_Vectors wv(system_.get(), plane);  // isf always uses the box size
    Wave_Density wd1(system_.get(), &wv), wd2(system_.get(), &wv);
    int inner = 0, outer = 0;
    auto densities = [&]() {
        std::cout << "\nCalculating fourier transform of density for first set of particles.";
        std::cout.flush();
        run_analysis(&wd1, read_line());
        if (symmetry == "symmetric") {
            wd2.copy_from(wd1);
        } else if (symmetry == "asymmetric") {
            std::cout << "\nCalculating fourier transform of density for second set of particles.";
            std::cout.flush();
            const std::string line = read_line();
            std::cout << "\n" << line;
            run_analysis(&wd2, line);
        } else {
            throw Error("command for structure factor calculation atom set types not understood.", 0);
        }
    };
    if (method == "auto") {
        if (args.size() != 8) throw Error("Incorrect number of arguments for isf auto.", 0);
        inner = std::atoi(args[6].c_str());
        outer = std::atoi(args[7].c_str());
        wd1.set(system_.get(), &wv, inner, outer);
        wd2.set(system_.get(), &wv, inner, outer);
        densities();
    } else if (method == "manual") {
        if (args.size() != 8) throw Error("Incorrect number of arguments for isf manual.", 0);
        densities();
        std::cout << "\nCalculating structure factor.";
        std::cout.flush();
        Intermediate_Scattering_Function struct_fac(system_.get(), 0, &wd1, &wd2);
        struct_fac.write(args[7]);
        std::cout << "Please input the lower and upper wavenumber index bounds, separated by spaces.\n";
        if (!(std::cin >> inner >> outer)) throw Error("isf manual: no wavenumber bounds on the input.", 0);
    } else {
        throw Error("isf method " + method + " not understood.", 0);
    }
    std::cout << "\nCalculating intermediate scattering function.";
    std::cout.flush();
    if (method == "manual" && inner != wd1.show_first()) {
        // The reference reads the density's slot wn - inner with shell wn's
        // vector count (src/intermediate_scattering_function.cpp:556-562);
        // over densities of every shell that runs past the slots, so the
        // shells asked for are computed instead.
        try {
            Intermediate_Scattering_Function f(system_.get(), &wd1, &wd2, inner, outer, fb);
            f.write(args[1]);
            return;
        } catch (const Error &e) {
            if (e.severe()) throw;
            warn(std::string(e.what()) + "; computing shells " + std::to_string(inner) + " to " +
                 std::to_string(outer) + " as requested");
        }
        Intermediate_Scattering_Function f(system_.get(), &wd1, &wd2, inner, outer, fb, true);
        f.write(args[1]);
        return;
    }
    Intermediate_Scattering_Function f(system_.get(), &wd1, &wd2, inner, outer, fb);
    f.write(args[1]);
}

void Control::structure_factor(const std::vector<std::string> &args)
{
    // structure_factor <file> <symmetry> <plane> <max_length_scale> [timescheme]
    if (args.size() != 5 && args.size() != 6) throw Error("Incorrect number of arguments for structure_factor.", 0);
    float ell = (float)std::atof(args[4].c_str());
    // A 5-token line has no timescheme: the reference reads atoi("") = 0
    // (src/control.cpp:2664-2673).
    const int ts = args.size() == 6 ? std::atoi(args[5].c_str()) : 0;
    if (ell < 0 && ts < -1) ell = system_->size(-ts - 2).min();
    Wave_Vectors wv(system_.get(), args[3], ell);
    const std::string runline = read_line();
    const time_t start = time(nullptr);
    const auto rt = tokenize(runline);
    if (!rt.empty() && rt[0] == "bin_list") {
        // src/control.cpp:2730-2744: binning works with symmetric S(k) only
        if (args[2] != "symmetric") throw Error("Binning currently only functions with symmetric S(k).", 0);
        std::cout << "\nCalculating structure factor.";
        for_each_target(runline, [&](const std::string &suffix, Trajectory_List *l) {
            Structure_Factor sf(system_.get(), &wv, ts);
            sf.analyze(l);
            sf.write(args[1] + suffix);
        });
        std::cout << "\nCalculated structure factor in " << time(nullptr) - start << " seconds.";
        return;
    }
    Structure_Factor sf(system_.get(), &wv, ts);
    Trajectory_List *l1 = target(runline);
    if (args[2] == "symmetric") {
        sf.analyze(l1);
    } else if (args[2] == "asymmetric") {
        sf.analyze(l1, target(read_line()));
    } else {
        throw Error("command for structure factor calculation atom set types not understood.", 0);
    }
    std::cout << "\nCalculated structure factor in " << time(nullptr) - start << " seconds.\n";
    sf.write(args[1]);
}

// find_fast_fixedthreshold <list> <count file> <time index> <threshold>,
// then the target line (src/control.cpp:3466-3496): the atoms whose squared
// displacement over the time index exceeds the threshold, per exponential
// block (Fast_Particles, src/fast_particles.cpp:117-139); the list is
// stored under <list> and its counts are written.
void Control::find_fast_fixedthreshold(const std::vector<std::string> &args)
{
    if (args.size() != 5) throw Error("Incorrect number of arguments for find_fast_fixedthreshold.", 0);
    const std::string listname = args[1], filename = args[2];
    const int time_index = std::atoi(args[3].c_str());
    const float threshold = (float)std::atof(args[4].c_str());
    const std::string runline = read_line();
    std::cout << "\n" << runline;
    std::cout << "\nIdentifying 'fast' particles.";
    Trajectory_List *in = target(runline);
    amdat_list *out = nullptr;
    System *s = system_.get();
    int rc = amdat_list_select_displacement(device_context(), s->traj(), in->handle(), s->show_n_exponentials(),
                                            s->show_n_exponential_steps(), s->show_frt(), s->show_limit(), time_index,
                                            AMDAT_SELECT_FAST, threshold, 0.0f, &out);
    if (rc < 0) throw Error(std::string("fast particles: ") + amdat_last_error(), 0);
    auto l = std::make_unique<Trajectory_List>(s, out);
    l->write_count(filename);
    add_list(listname, std::move(l));
}

int Control::run()
{
    // System block: type, trajectory type, file, time scheme, species,
    // types, one count row per species.
    std::vector<std::string> block;
    for (int i = 0; i < 6; i++) block.push_back(read_line());
    const size_t n_species = tokenize(block[4]).size() / 2;
    for (size_t s = 0; s < n_species; s++) block.push_back(read_line());
    // xyz trajectories: box lengths and box bounds follow (src/system.cpp:385-397)
    if (!tokenize(block[1]).empty() && tokenize(block[1])[0] == "xyz")
        for (int i = 0; i < 2; i++) block.push_back(read_line());
    system_ = std::make_unique<System>(block);
    std::cout << "\nSystem created." << std::endl;

    while (cursor_ < lines_.size()) {
        std::string line;
        try {
            line = read_line();
        } catch (const Error &) {
            break;
        }
        auto args = strip_store_flag(tokenize(line));
        const std::string &cmd = args[0];
        std::cout << "\n" << line << std::endl;
        try {
            if (cmd == "exit") break;
            else if (cmd == "create_list") create_list(args);
            else if (cmd == "limit") system_->set_limit(std::atoi(args.at(1).c_str()));
            else if (cmd == "processors") continue;
            else if (cmd == "msd") msd(args);
            else if (cmd == "ngp") ngp(args);
            else if (cmd == "isfs") isfs(args);
            else if (cmd == "isf") isf(args);
            else if (cmd == "structure_factor") structure_factor(args);
            else if (cmd == "msd_2d") msd_2d(args);
            else if (cmd == "mean_displacement") mean_displacement(args);
            else if (cmd == "vhs") vhs(args);
            else if (cmd == "find_fast_fixedthreshold") find_fast_fixedthreshold(args);
            else if (cmd == "u2dist") u2dist(args);
            else if (cmd == "gaussian_comparison") gaussian_comparison(args);
            else if (cmd == "create_bin_list") create_bin_list(args);
            else if (cmd == "find_fast") find_fast(args);
            else if (cmd == "displacement_dist") displacement_dist(args);
            else if (cmd == "print") { std::cout << line.substr(line.find("print") + 5) << std::endl; }
            else throw Error("command " + cmd + " is not supported by this build", 0);
        } catch (const Error &e) {
            if (e.severe()) {
                std::cout << "\nError: " << e.what() << std::endl;
                return e.code() == 0 ? 1 : -e.code();
            }
            std::cout << "\nError: " << e.what() << " Skipping command." << std::endl;
        }
    }
    return 0;
}

}  // namespace amdat_hip

int main(int argc, char **argv)
{
    std::string input;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a == "-i" && i + 1 < argc) input = argv[++i];
        else if (a == "-v") { std::cout << "AMDAT " << amdat_hip::kSemver << std::endl; return 0; }
        else if (a == "-n" && i + 1 < argc) ++i;  // thread count: the device path ignores it
    }
    if (input.empty()) {
        std::cerr << "usage: amdat_hip -i <input file>" << std::endl;
        return 2;
    }
    try {
        amdat_hip::Control c(input);
        return c.run();
    } catch (const amdat_hip::Error &e) {
        std::cout << "\nError: " << e.what() << std::endl;
        return 1;
    }
}
