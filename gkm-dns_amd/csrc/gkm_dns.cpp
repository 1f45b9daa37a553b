// gkm_dns.cpp — command-line driver over libgkm.so that runs a reference case
// folder the way the weno code's main + Evolve do (weno/src/Main.cpp:10-49,
// update.cpp:41-173), one GPU instead of an MPI Cartesian grid:
//
//   GetParams + PrintParams   read <case>/params, print the parameter block
//   Initialize_DHIT           read the <case>/Init%i%i%i rank files of the
//                             decomposition they were written with
//   HaloTransfer, timestep    ghost fill, dt = cfl dx / max(|u| + c), once
//   while (t <= tmax)         PrintStats line; one step; t += dt (tmax = 7 t0)
//   PrintStats                final line
//   WriteReInit (optional)    restart files of the final state
//
// With --dhit N instead of a case folder, the library generates the
// init_BS.m initial condition itself (gkm_init_dhit: N^3, --k0, --mt, --re,
// --seed, --gamma, density 1), and --write-case DIR writes it out as a case
// folder (params + Init rank files) that the reference can read.
//
// usage: gkm_dns <case-dir> | --dhit N [--k0 k] [--mt M] [--re R] [--seed s]
//        [--gamma g] [--write-case DIR]
//        [--proc-dim a b c] [--tmax T | --steps N]
//        [--cfl C] [--precision fp64|mixed] [--recon symoo|js]
//        [--visc sutherland|power] [--visc-exp e] [--energy T|E] [--pr Pr]
//        [--R R] [--z-fixed] [--scheme weno|gks] [--reinit DIR] [--timing]
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

#include "../../include/gkm.h"
#include "../../include/gkm_ic.h"
#include "../../include/gkm_io.h"

namespace {

void die(const char* what, int rc) {
    std::fprintf(stderr, "gkm_dns: %s failed (%d): %s\n", what, rc, gkm_last_error());
    std::exit(1);
}

#define CALL(expr)                     \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ != GKM_OK) die(#expr, rc_); \
    } while (0)

void usage() {
    std::fprintf(stderr,
                 "usage: gkm_dns <case-dir> | --dhit N [--k0 k] [--mt M] [--re R] [--seed s] [--gamma g]\n"
                 "               [--write-case DIR] [--proc-dim a b c] [--tmax T | --steps N] [--cfl C]\n"
                 "               [--precision fp64|mixed] [--recon symoo|js] [--visc sutherland|power]\n"
                 "               [--visc-exp e] [--energy T|E] [--pr Pr] [--R R] [--z-fixed]\n"
                 "               [--scheme weno|gks] [--reinit DIR] [--timing]\n");
    std::exit(2);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) usage();
    std::string dir;
    int pd[3] = {1, 1, 1};
    double tmax = -1.0, cfl = 0.02;  // cfl as weno/src/param.h:17
    long steps = -1;
    gkm_params par;
    std::memset(&par, 0, sizeof par);
    par.R = 287.0;  // param.h:18
    par.Pr = 0.7;   // param.h:16
    par.visc_exp = 0.76;
    par.visc_law = GKM_VISC_SUTHERLAND;
    par.recon = GKM_RECON_SYMOO;
    par.energy_recon = GKM_ENERGY_T;
    par.zbug_compat = 1;
    par.precision = GKM_PREC_FP64;
    std::string reinit, write_case;
    bool timing = false;
    int dhit_n = 0;
    gkm_ic ic = {1234ull, 4.0, 0.3, 30.0, 1.0};
    double dhit_gamma = 1.4;
    for (int a = 1; a < argc; a++) {
        const std::string s = argv[a];
        auto next = [&]() -> const char* {
            if (a + 1 >= argc) usage();
            return argv[++a];
        };
        if (s == "--proc-dim") {
            pd[0] = std::atoi(next());
            pd[1] = std::atoi(next());
            pd[2] = std::atoi(next());
        } else if (s == "--tmax") {
            tmax = std::atof(next());
        } else if (s == "--steps") {
            steps = std::atol(next());
        } else if (s == "--cfl") {
            cfl = std::atof(next());
        } else if (s == "--precision") {
            const std::string v = next();
            par.precision = v == "mixed" ? GKM_PREC_MIXED : GKM_PREC_FP64;
        } else if (s == "--recon") {
            const std::string v = next();
            par.recon = v == "js" ? GKM_RECON_JS : GKM_RECON_SYMOO;
        } else if (s == "--visc") {
            const std::string v = next();
            par.visc_law = v == "power" ? GKM_VISC_POWER : GKM_VISC_SUTHERLAND;
        } else if (s == "--visc-exp") {
            par.visc_exp = std::atof(next());
        } else if (s == "--energy") {
            const std::string v = next();
            par.energy_recon = v == "E" ? GKM_ENERGY_E : GKM_ENERGY_T;
        } else if (s == "--pr") {
            par.Pr = std::atof(next());
        } else if (s == "--R") {
            par.R = std::atof(next());
        } else if (s == "--scheme") {
            const std::string v = next();
            par.flux_scheme = v == "gks" ? GKM_SCHEME_GKS : GKM_SCHEME_WENO_BGK;
        } else if (s == "--z-fixed") {
            par.zbug_compat = 0;
        } else if (s == "--reinit") {
            reinit = next();
        } else if (s == "--timing") {
            timing = true;
        } else if (s == "--dhit") {
            dhit_n = std::atoi(next());
        } else if (s == "--k0") {
            ic.k0 = std::atof(next());
        } else if (s == "--mt") {
            ic.Mt = std::atof(next());
        } else if (s == "--re") {
            ic.Re = std::atof(next());
        } else if (s == "--seed") {
            ic.seed = std::strtoull(next(), nullptr, 10);
        } else if (s == "--gamma") {
            dhit_gamma = std::atof(next());
        } else if (s == "--write-case") {
            write_case = next();
        } else if (!s.empty() && s[0] == '-') {
            usage();
        } else {
            dir = s;
        }
    }
    if (dir.empty() == (dhit_n == 0)) usage();  // a case folder or --dhit, not both
    const double pi = std::atan(1.0) * 4.0;
    gkm_weno_params wp;
    gkm_t* h = nullptr;
    if (dhit_n > 0) {
        // GetParams + Initialize_DHIT from the library's generator
        wp.nc = dhit_n;
        wp.gam = dhit_gamma;
        par.gamma = dhit_gamma;
        par.K = (int)((5.0 - 3.0 * dhit_gamma) / (dhit_gamma - 1.0));
        par.mu0 = 0.0;  // set from the recipe by gkm_init_dhit
        par.T0 = 1.0;
        gkm_grid g0 = {dhit_n, dhit_n, dhit_n, 3, 2.0 * pi / dhit_n};
        CALL(gkm_create(&g0, &par, &h));
        gkm_ic_info info;
        CALL(gkm_init_dhit(h, &ic, &info));
        CALL(gkm_get_params_box(h, 0, &par));
        wp.den0 = ic.rho0;
        wp.Mt = ic.Mt;
        wp.Re = ic.Re;
        wp.mu0 = info.mu0;
        wp.p0 = info.p0;
        wp.t0 = info.t0;
    } else {
        // GetParams (weno/src/FlowField.cpp:9-57)
        CALL(gkm_read_weno_params((dir + "/params").c_str(), &wp));
        par.gamma = wp.gam;
        par.K = (int)((5.0 - 3.0 * wp.gam) / (wp.gam - 1.0));
        par.mu0 = wp.mu0;
        par.T0 = wp.p0 / (par.R * wp.den0);
    }
    const int nc = wp.nc;
    gkm_grid grid = {nc, nc, nc, 3, 2.0 * pi / nc};
    if (tmax < 0 && steps < 0) tmax = 7.0 * wp.t0;  // tmax = 7 t0 (param.h:62)

    // PrintParams (FlowField.cpp:305-331)
    std::printf("\nGKM DNS DHIT\nInitializing Field .....\n______________________________\n\n");
    std::printf("nc            =  %d\n", nc);
    std::printf("nc_segy       =  %d\nnc_segx       =  %d\nnc_segz       =  %d\n", nc / pd[0], nc / pd[1], nc / pd[2]);
    std::printf("T0            =  %g\nMt            =  %g\nRe            =  %g\ndx            =  %g\n", par.T0, wp.Mt,
                wp.Re, grid.dx);
    std::printf("eddy time     =  %g\np0            =  %g\nden0          =  %g\ngam           =  %g\n", wp.t0, wp.p0,
                wp.den0, wp.gam);
    std::printf("mu0           =  %g\nK             =  %d\nCFL           =  %g\nPrandtl No.   =  %g\n", wp.mu0, par.K,
                cfl, par.Pr);
    std::printf("nprocs    =  %d x %d x %d (read as one box on %s)\n", pd[0], pd[1], pd[2], gkm_version());
    std::printf("\nField Initialized ! \n______________________________\n");

    // Initialize_DHIT (FlowField.cpp:59-112)
    const size_t ncell = (size_t)(nc + 6) * (nc + 6) * (nc + 6);
    std::vector<double> store(5 * ncell, 0.0);
    double* W[5];
    for (int q = 0; q < 5; q++) W[q] = store.data() + q * ncell;
    auto make_dir = [](const std::string& d) {
        std::error_code ec;
        std::filesystem::create_directories(d, ec);
        if (ec) {
            std::fprintf(stderr, "gkm_dns: cannot create %s: %s\n", d.c_str(), ec.message().c_str());
            std::exit(1);
        }
    };
    if (dhit_n > 0) {
        if (!write_case.empty()) {  // the case folder init_BS.m would have written
            make_dir(write_case);
            CALL(gkm_write_weno_params((write_case + "/params").c_str(), &wp));
            CALL(gkm_get_state(h, W));
            CALL(gkm_write_weno_init(write_case.c_str(), nc, pd, wp.gam, W));
        }
    } else {
        CALL(gkm_read_weno_init(dir.c_str(), nc, pd, wp.gam, W));
        CALL(gkm_create(&grid, &par, &h));
        CALL(gkm_set_state(h, W));
    }
    CALL(gkm_fill_ghosts(h));  // HaloTransfer before the loop (update.cpp:106-110)
    double dt = 0.0;
    CALL(gkm_compute_dt(h, cfl, &dt));  // timestep(W), once (update.cpp:112)

    // the Evolve loop (update.cpp:120-156): PrintStats line of the state, one
    // step, t += dt.  gkm_step_print_stats runs the statistics of each state
    // beside the step that reads it; the lines are printed per chunk.
    long n = 0;
    if (steps >= 0) {
        n = steps;
    } else {
        for (double tt = 0.0; tt <= tmax; tt += dt) n++;
    }
    double t = 0.0;
    long iter = 0;
    char line[512];
    gkm_print_stats_t st;
    std::vector<gkm_print_stats_t> sts;
    const auto t0 = std::chrono::steady_clock::now();
    while (iter < n) {
        const int k = (int)std::min<long>(256, n - iter);
        sts.resize((size_t)k);
        CALL(gkm_step_print_stats(h, dt, k, sts.data()));
        for (int s = 0; s < k; s++) {
            CALL(gkm_format_print_stats(&sts[s], t / wp.t0, wp.gam, wp.p0, par.T0, wp.Mt, wp.den0, line, sizeof line));
            std::printf("%s\n", line);
            t += dt;
        }
        iter += k;
    }
    CALL(gkm_print_stats(h, &st));
    const double loop_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    CALL(gkm_format_print_stats(&st, t / wp.t0, wp.gam, wp.p0, par.T0, wp.Mt, wp.den0, line, sizeof line));
    std::printf("%s\n", line);
    std::fflush(stdout);
    if (timing)
        std::fprintf(stderr,
                     "gkm_dns: %ld iterations of %d^3, dt = %.6e, %.3f ms per iteration (PrintStats + step) "
                     "(%.3e cell-updates/s)\n",
                     iter, nc, dt, iter ? loop_ms / iter : 0.0,
                     iter ? (double)nc * nc * nc * iter / (loop_ms * 1e-3) : 0.0);

    if (!reinit.empty()) {  // WriteReInit (update.cpp:1352-1391)
        make_dir(reinit);
        CALL(gkm_get_state(h, W));
        CALL(gkm_write_weno_init(reinit.c_str(), nc, pd, wp.gam, W));
    }
    CALL(gkm_destroy(h));
    return 0;
}
