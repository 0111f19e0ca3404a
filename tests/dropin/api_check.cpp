// Exercises the drop-in class API beyond what the input-script driver
// calls: Wave_Density over every shell, its static S(q) form, set(), copies
// and show_density; the Intermediate_Scattering_Function constructors for
// every shell, one shell, a shell range and one time gap.  Reads a raw
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
        Intermediate_Scattering_Function f_one(&sys, &wd, &wd2, 3, false);
        print_isf("ONE", f_one, ntg, 1);
        Intermediate_Scattering_Function f_range(&sys, &part, &part, 3, 5, false);
        print_isf("RANGE", f_range, ntg, 3);
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
    } catch (const Error &e) {
        std::printf("ERROR %s\n", e.what());
        return 1;
    }
    return 0;
}
