// Exercises the drop-in class API beyond what the input-script driver
// calls: Wave_Density over every shell, its static S(q) form, set(), copies,
// show_density and accumulation over two analyze() calls; the
// Intermediate_Scattering_Function constructors for every shell, one shell,
// a shell range (the reference's slot indexing, the physics-correct option
// and a refused read past a slot) and one time gap; every class driven
// through an Analysis* (analyze, write, what_are_you).  Reads a raw
// trajectory (x, y, z FP32 [T][N], ix, iy, iz int16 [T][N], box FP32 [T][3])
// and prints the values tests/test_gpu_dropin_api.py checks against the
// oracle.  Test infrastructure: built by amdat_amd/Makefile as bin/api_check.
//
// usage: api_check <raw file> <n_atoms> <n_frames> <n_exponentials> <n_exponential_steps> <base>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <numeric>
#include <vector>

#include "../../amdat_amd/host/amdat_host.h"

using namespace amdat_hip;

static void print_isf(const char *key, const Intermediate_Scattering_Function &f, int ntg, int nsh)
{
    for (int t = 0; t < ntg; t++)
        for (int s = 0; s < nsh; s++) std::printf("%s %d %d %.9e\n", key, t, s, (double)f.show(t, s));
}

int main(int argc, char **argv)
{
    if (argc != 7) {
        std::fprintf(stderr, "usage: api_check <raw> <N> <T> <nexp> <nes> <base>\n");
        return 2;
    }
    const int N = std::atoi(argv[2]), T = std::atoi(argv[3]), nexp = std::atoi(argv[4]), nes = std::atoi(argv[5]);
    const float base = (float)std::atof(argv[6]);
    const size_t NT = (size_t)N * T;
    std::vector<float> x(NT), y(NT), z(NT), box(3 * (size_t)T);
    std::vector<int16_t> ix(NT), iy(NT), iz(NT);
    std::ifstream in(argv[1], std::ios::binary);
    for (auto *v : {&x, &y, &z}) in.read((char *)v->data(), NT * sizeof(float));
    for (auto *v : {&ix, &iy, &iz}) in.read((char *)v->data(), NT * sizeof(int16_t));
    in.read((char *)box.data(), box.size() * sizeof(float));
    if (!in) {
        std::fprintf(stderr, "short raw file\n");
        return 2;
    }
    std::vector<int> types(N, 1);
    try {
        System sys(nexp, nes, base, 0, 0, 1.0f, N, x.data(), y.data(), z.data(), ix.data(), iy.data(), iz.data(),
                   box.data(), types.data());
        Wave_Vectors wv(&sys, "xyz");
        std::vector<int> all(N);
        std::iota(all.begin(), all.end(), 0);
        Trajectory_List list(&sys, all);
        const int ntg = sys.show_n_timegaps(), nsh = wv.show_n_wavenumbers();

        Wave_Density wd(&sys, &wv);
        wd.analyze(&list);
        Wave_Density wd2(wd);
        Wave_Density sf(&sys, &wv, true);
        sf.analyze(&list);
        Wave_Density part(&sys, &wv, 0, 0);
        part.set(&sys, &wv, 2, 5);
        part.analyze(&list);
        std::printf("LOOPED %d %d %d\n", wd.show_n_atoms_looped(), sf.show_n_atoms_looped(), part.show_first());
        for (int t : {0, 1, nes})
            for (int s : {0, 3})
                for (int v = 0; v < 2; v++) {
                    const std::complex<double> a = wd2.show_density(t, s, v), b = sf.show_density(t, s, v);
                    std::printf("RHO %d %d %d %.17e %.17e %.17e %.17e\n", t, s, v, a.real(), a.imag(), b.real(),
                                b.imag());
                }

        Intermediate_Scattering_Function f_all(&sys, &wd, &wd2);
        print_isf("ALL", f_all, ntg, nsh);
        Intermediate_Scattering_Function f_one(&sys, &wd, &wd2, 3, false, true);
        print_isf("ONE", f_one, ntg, 1);
        Intermediate_Scattering_Function f_range(&sys, &part, &part, 3, 5, false, true);
        print_isf("RANGE", f_range, ntg, 3);
        // the reference's reading of a range that starts past the density's
        // first shell: column i sums slot i (shell 20 + i) over the vector
        // count of shell 22 + i (300 vectors each from shell 18 on)
        Wave_Density high(&sys, &wv, 20, 30);
        high.analyze(&list);
        Intermediate_Scattering_Function f_slots(&sys, &high, &high, 22, 28, false);
        print_isf("SLOTS", f_slots, ntg, 7);
        Intermediate_Scattering_Function f_phys(&sys, &high, &high, 22, 28, false, true);
        print_isf("PHYS", f_phys, ntg, 7);
        // one shell of that density: slot 0 (shell 20) over shell 25's count
        Intermediate_Scattering_Function f_one_slot(&sys, &high, &high, 25, false);
        print_isf("ONESLOT", f_one_slot, ntg, 1);
        // slot 0 of part (shell 2, 15 vectors) over shell 3's 12 vectors
        Intermediate_Scattering_Function f_short(&sys, &part, &part, 3, 3, false);
        print_isf("SHORT", f_short, ntg, 1);
        try {
            // slot 0 (shell 2, 15 vectors) would be read over shell 4's 33
            Intermediate_Scattering_Function past(&sys, &part, &part, 4, 5, false);
            std::printf("PASTSLOT accepted\n");
        } catch (const Error &) {
            std::printf("PASTSLOT rejected\n");
        }
        Intermediate_Scattering_Function f_gap0(&sys, 0, &sf, &sf, false);
        print_isf("GAP0", f_gap0, ntg, nsh);
        Intermediate_Scattering_Function f_gap(&sys, nes, &wd, &wd2, true);
        print_isf("GAPN", f_gap, ntg, nsh);
        try {
            Intermediate_Scattering_Function bad(&sys, &wd, &part);
            std::printf("MISMATCH accepted\n");
        } catch (const Error &) {
            std::printf("MISMATCH rejected\n");
        }

        // accumulation: two analyze() calls over disjoint lists add up to
        // the density of their union; a copy taken in between keeps the first
        std::vector<int> even, odd;
        for (int j = 0; j < N; j++) (j % 3 ? odd : even).push_back(j);
        Trajectory_List le(&sys, even), lo(&sys, odd);
        Wave_Density acc(&sys, &wv, 0, 9);
        Analysis *pa = &acc;
        pa->analyze(&le);
        Wave_Density first_only(acc);
        pa->analyze(&lo);
        std::printf("ACCLOOPED %d %d %d\n", acc.show_n_atoms_looped(), first_only.show_n_atoms_looped(),
                    acc.show_n_atoms(1));
        for (int t : {0, 2, T - 1})
            for (int sh : {0, 4, 9})
                for (int v = 0; v < 2; v++) {
                    const std::complex<double> a = acc.show_density(t, sh, v), b = first_only.show_density(t, sh, v);
                    std::printf("ACC %d %d %d %.17e %.17e %.17e %.17e\n", t, sh, v, a.real(), a.imag(), b.real(),
                                b.imag());
                }

        // every analysis through Analysis*: kind, analyze, write
        Mean_Square_Displacement msd(&sys);
        Non_Gaussian_Parameter ngp(&sys, &msd);
        Incoherent_Scattering_Function isfs(&sys, &wv, 5, 7, false);
        Structure_Factor sk(&sys, &wv, -1);
        Mean_Square_Displacement_2D msd2(&sys, "xy");
        Mean_Displacement md(&sys);
        Van_Hove_Self vhs(&sys, 20, 2.0f);
        Analysis *all_a[] = {&msd, &ngp, &isfs, &sk, &msd2, &md, &vhs, &wd, &f_all};
        const char *names[] = {"msd", "ngp", "isfs", "sk", "msd2", "md", "vhs", "wd", "isf"};
        for (int i = 0; i < 9; i++) {
            Analysis *a = all_a[i];
            if (i < 7) a->analyze(&list);
            std::printf("\nKIND %s %d %d\n", names[i], (int)a->what_are_you(), (int)(a->show_system() == &sys));
            if (i < 7 || i == 8) a->write(std::string("api_") + names[i] + ".dat");
        }
        sk.analyze(&le, &list);
        std::printf("\nSKASYM");
        for (int b = 0; b < 5; b++) std::printf(" %.9e", (double)sk.show(b));
        std::printf("\n");
        // methods a class does not implement print the reference's message
        Analysis *w = &wd;
        w->write("never.dat");
        msd.analyze(&le, &lo);
    } catch (const Error &e) {
        std::printf("ERROR %s\n", e.what());
        return 1;
    }
    return 0;
}
