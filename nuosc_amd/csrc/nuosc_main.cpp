// nuosc_main.cpp -- the ./nuosc command line (nuosc.cpp:17-261) over the GPU
// NuOsc class.  Same flags, defaults and outputs (stdout analysis lines,
// analysis.dat, wall-time lines, [Summ] line).  Additions: --precision 64|32,
// --seed S (for --ipt 5), --vw voronoi|equal, --snap_p3 1 (P3 .bin/.meta
// snapshot every DUMP_EVERY steps), --graph 0|1, --device N, --bc
// periodic|open (open: z ends open, an ipt-0 pulse flowing in), --zc Z
// (centre of the ipt-0 profile, 0 in the reference) and --srk3 1 (step with
// NuOsc::step_srk3 instead of step_rk4).
#include <sys/resource.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>
#include <stdexcept>

#include "NuOsc.h"

namespace {
double max_rss_mb() {  // utils.h:10-14
    struct rusage r;
    getrusage(RUSAGE_SELF, &r);
    return r.ru_maxrss / 1024.0;
}
}  // namespace

int main(int argc, char* argv[]) {
    // defaults: nuosc.cpp:19-61
    int px = 1, pz = 1;
    double dz = 0.1, dx = 0.1;
    double x0 = -1, x1 = -x0;
    double z0 = -10, z1 = -z0;
    int nv_in = 33, nphi = 32;
    double cfl = 0.4, ko = 0.0;
    double mu = 1.0, pmo = 1.0;
    bool renorm = false;
    double alpha = 0.9, lnue = 0.6, lnueb = 0.53;
    double lnuex = std::numeric_limits<double>::max(), lnuebx = std::numeric_limits<double>::max();
    double ipt = 0, eps0 = 0.1, sigma = 100.0;
    int ANAL_EVERY = 5, END_STEP = 5, DUMP_EVERY = 99999999;
    int gz = 2, gx = 2;
    // additions
    int precision = NUOSC_PREC_FP64, vgrid = NUOSC_VGRID_VORONOI, snap_p3 = 0, use_graph = 1, device = 0;
    int bc = NUOSC_BC_PERIODIC, srk3 = 0;
    double zc = 0.0;
    int ko_order = 3, adv = NUOSC_ADV_CFD4, nflavor = 2;
    bool gz_set = false;
    unsigned long long seed = 12345;

    for (int t = 1; t < argc; t++) {
        auto val = [&](void) -> const char* {
            if (t + 1 >= argc) {
                std::printf("Missing value for %s!\n", argv[t]);
                std::exit(0);
            }
            return argv[++t];
        };
        const char* a = argv[t];
        if (!std::strcmp(a, "--dz")) { dz = std::atof(val()); dx = dz; }
        else if (!std::strcmp(a, "--dx")) dx = std::atof(val());
        else if (!std::strcmp(a, "--zmax")) { z1 = std::atof(val()); z0 = -z1; x1 = z1; x0 = z0; }
        else if (!std::strcmp(a, "--xmax")) { x1 = std::atof(val()); x0 = -x1; }
        else if (!std::strcmp(a, "--gz")) { gz = (int)std::atof(val()); gx = gz; gz_set = true; }
        else if (!std::strcmp(a, "--cfl")) cfl = std::atof(val());
        else if (!std::strcmp(a, "--nv")) nv_in = std::atoi(val());
        else if (!std::strcmp(a, "--nphi")) nphi = std::atoi(val());
        else if (!std::strcmp(a, "--ko")) ko = std::atof(val());
        else if (!std::strcmp(a, "--mu")) mu = std::atof(val());
        else if (!std::strcmp(a, "--pmo")) pmo = std::atof(val());
        else if (!std::strcmp(a, "--renorm")) renorm = bool(std::atoi(val()));
        else if (!std::strcmp(a, "--ANA_EVERY")) { ANAL_EVERY = std::atoi(val()); std::cout << " ** ANAL_EVERY: " << ANAL_EVERY << std::endl; }
        else if (!std::strcmp(a, "--DUMP_EVERY")) { DUMP_EVERY = std::atoi(val()); std::cout << " ** DUMP_EVERY: " << DUMP_EVERY << std::endl; }
        else if (!std::strcmp(a, "--END_STEP")) { END_STEP = std::atoi(val()); std::cout << " ** END_STEP: " << END_STEP << std::endl; }
        // *_T flags use cfl and dz as parsed so far (nuosc.cpp:105-113)
        else if (!std::strcmp(a, "--ANA_EVERY_T")) { ANAL_EVERY = int(std::atof(val()) / (cfl * dz) + 0.5); std::cout << " ** ANAL_EVERY: " << ANAL_EVERY << std::endl; }
        else if (!std::strcmp(a, "--DUMP_EVERY_T")) { DUMP_EVERY = int(std::atof(val()) / (cfl * dz) + 0.5); std::cout << " ** DUMP_EVERY: " << DUMP_EVERY << std::endl; }
        else if (!std::strcmp(a, "--END_STEP_T")) { END_STEP = int(std::atof(val()) / (cfl * dz) + 0.5); std::cout << " ** END_STEP: " << END_STEP << std::endl; }
        else if (!std::strcmp(a, "--lnue")) lnue = std::atof(val());
        else if (!std::strcmp(a, "--lnueb")) lnueb = std::atof(val());
        else if (!std::strcmp(a, "--lnuex")) lnuex = std::atof(val());
        else if (!std::strcmp(a, "--lnuebx")) lnuebx = std::atof(val());
        else if (!std::strcmp(a, "--sigma")) sigma = std::atof(val());
        else if (!std::strcmp(a, "--eps0")) eps0 = std::atof(val());
        else if (!std::strcmp(a, "--alpha")) alpha = std::atof(val());
        else if (!std::strcmp(a, "--ipt")) ipt = std::atoi(val());
        else if (!std::strcmp(a, "--np")) { px = std::atoi(val()); pz = std::atoi(val()); }
        // additions
        else if (!std::strcmp(a, "--precision")) precision = std::atoi(val());
        else if (!std::strcmp(a, "--seed")) seed = std::strtoull(val(), nullptr, 10);
        else if (!std::strcmp(a, "--vw")) {
            const char* m = val();
            if (!std::strcmp(m, "voronoi")) vgrid = NUOSC_VGRID_VORONOI;
            else if (!std::strcmp(m, "equal")) vgrid = NUOSC_VGRID_EQUAL;
            else { std::printf("Unreconganized parameters %s!\n", m); std::exit(0); }
        }
        else if (!std::strcmp(a, "--snap_p3")) snap_p3 = std::atoi(val());
        else if (!std::strcmp(a, "--graph")) use_graph = std::atoi(val());
        else if (!std::strcmp(a, "--device")) device = std::atoi(val());
        else if (!std::strcmp(a, "--bc")) {
            const char* m = val();
            if (!std::strcmp(m, "periodic")) bc = NUOSC_BC_PERIODIC;
            else if (!std::strcmp(m, "open")) bc = NUOSC_BC_OPEN_Z;
            else { std::printf("Unreconganized parameters %s!\n", m); std::exit(0); }
        }
        else if (!std::strcmp(a, "--zc")) zc = std::atof(val());
        else if (!std::strcmp(a, "--srk3")) srk3 = std::atoi(val());
        else if (!std::strcmp(a, "--ko_order")) ko_order = std::atoi(val());
        else if (!std::strcmp(a, "--nflavor")) nflavor = std::atoi(val());
        else if (!std::strcmp(a, "--adv")) {
            const char* m = val();
            static const char* const names[] = {"cfd4", "cfd2", "upwind2", "lopsided", "muscl", "off"};
            int k = 0;
            while (k <= NUOSC_ADV_OFF && std::strcmp(m, names[k])) ++k;
            if (k > NUOSC_ADV_OFF) { std::printf("Unreconganized parameters %s!\n", m); std::exit(0); }
            adv = k;
        }
        else {
            // nuosc.cpp:135-137 (message and exit status kept)
            std::printf("Unreconganized parameters %s!\n", a);
            std::exit(0);
        }
    }

    // nuosc.cpp:57-61: ghost width 3 for the radius-3 stencils unless given
    if (!gz_set && (ko_order == 5 || adv == NUOSC_ADV_LOPSIDED)) gx = gz = 3;
    try {
        NuOsc state(px, pz, nv_in, nphi, gx, gz, x0, x1, z0, z1, dx, dz, cfl, ko, precision, vgrid, use_graph != 0,
                    device, bc, zc, ko_order, adv, nflavor);
        const long long lpts = state.grid.get_lpts();
        state.set_mu(mu);
        state.set_pmo(pmo);
        state.set_renorm(renorm);
        const int nx = state.grid.nx, nz = state.grid.nz;

        state.fillInitValue((int)ipt, alpha, eps0, sigma, lnue, lnueb, lnuex, lnuebx, seed);
        state.analysis();

        if (DUMP_EVERY <= END_STEP && snap_p3) {
            // the P3 slice of nuosc.cpp:169-175 (x = 0, nx/2, nx-1)
            std::vector<int> vslice;
            for (int v = 0; v < nv_in; ++v) {
                const int k = int((nv_in - 1) / 2) * nv_in + v;
                if (k < state.get_nv()) vslice.push_back(k);
            }
            state.addSnapShotAtXV({"P3"}, "P3_%06d.bin", DUMP_EVERY, {0, nx / 2, nx - 1}, vslice);
            state.checkSnapShot(0);
        }
        std::cout << std::flush;

        double stepms = 0;
        std::chrono::time_point<std::chrono::high_resolution_clock> t1;
        const int cooltime = 3;  // nuosc.cpp:204-237
        for (int t = 1; t <= END_STEP; t++) {
            if (t == cooltime) {
                state.synchronize();
                t1 = std::chrono::high_resolution_clock::now();
            }
            if (srk3) state.step_srk3();
            else state.step_rk4();
            if (t % ANAL_EVERY == 0) state.analysis();
            if (t % DUMP_EVERY == 0) state.checkSnapShot(t);
            if (t == 5 || t == 10 || t == 100 || t == 1000 || t == END_STEP) {
                state.synchronize();
                stepms = std::chrono::duration_cast<std::chrono::milliseconds>(
                             std::chrono::high_resolution_clock::now() - t1).count();
                std::printf("%d Walltime: (Min) %.3f s/T, %.2f ns/step-grid.    (Max) %.3f s/T, %.2f ns/step-grid.\n",
                            t, stepms / state.phy_time / 1000, stepms / (t - cooltime + 1) / lpts * 1e6,
                            stepms / state.phy_time / 1000, stepms / (t - cooltime + 1) / lpts * 1e6);
            }
        }
        const double tmem = max_rss_mb();
        const double ns_per_stepgrid = stepms / (END_STEP - cooltime + 1) / lpts * 1e6;
        std::printf("Completed.\n");
        std::printf("Memory usage (MB) per node: %.2f ~ %.2f\n", tmem, tmem);
        std::printf("[Summ] %d %d %d %d %d %d %f\n", 1, px, pz, nx, nz, state.get_nv(), ns_per_stepgrid);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "nuosc: %s\n", e.what());
        return 1;
    }
    return 0;
}
