/*
 * san_driver.cpp — runs the oracle and the library's host-only code under the
 * host sanitizers: the readers and writers of the reference's file formats
 * (gkm-dns_amd/csrc/gkm_io.cpp) and the initial-condition generator
 * (gkm-dns_amd/csrc/gkm_ic.cpp, threaded transforms).
 *
 * TEST INFRASTRUCTURE ONLY.  Built by `make -C oracle san` (AddressSanitizer
 * + UndefinedBehaviorSanitizer, every report fatal) and `make -C oracle tsan`
 * (ThreadSanitizer, generator only, built with -DSAN_IC_ONLY); run by
 * tests/test_sanitize.py.  Nothing in the product links it.
 *
 * The handle entry points gkm_ic.cpp calls (gkm_get_grid, gkm_get_params_box,
 * gkm_set_params_box, gkm_set_state_box, gkm_fill_ghosts) are host stand-ins
 * here that keep a padded FP64 state, so the generator's whole host path runs
 * without a device.  Exit status 0 and a last line "san_driver: ok" when every
 * check passes; a sanitizer report aborts the run.
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../include/gkm.h"
#include "../include/gkm_ic.h"
#ifndef SAN_IC_ONLY
#include "../include/gkm_io.h"
#include "gkm_oracle.h"
#endif

namespace gkm {
// the library's error reporting, which gkm_io.cpp and gkm_ic.cpp call
int report_error(int code, const char* msg) {
    std::fprintf(stderr, "  (reported: %d %s)\n", code, msg);
    return code;
}
}  // namespace gkm

// ---- host stand-in of a one-box handle ---------------------------------------
struct gkm_handle {
    gkm_grid g;
    gkm_params p;
    std::vector<double> W[5];
    int ghost_fills = 0;
};

extern "C" {
int gkm_get_grid(gkm_t* h, gkm_grid* out) {
    *out = ((gkm_handle*)h)->g;
    return GKM_OK;
}
int gkm_get_params_box(gkm_t* h, int box, gkm_params* out) {
    if (box != 0) return GKM_E_ARG;
    *out = ((gkm_handle*)h)->p;
    return GKM_OK;
}
int gkm_set_params_box(gkm_t* h, int box, const gkm_params* p) {
    if (box != 0) return GKM_E_ARG;
    ((gkm_handle*)h)->p = *p;
    return GKM_OK;
}
int gkm_set_state_box(gkm_t* h, int box, const double* const W[5]) {
    gkm_handle* s = (gkm_handle*)h;
    if (box != 0) return GKM_E_ARG;
    const size_t n = (size_t)(s->g.nx + 6) * (s->g.ny + 6) * (s->g.nz + 6);
    for (int q = 0; q < 5; q++) s->W[q].assign(W[q], W[q] + n);
    return GKM_OK;
}
int gkm_fill_ghosts(gkm_t* h) {
    ((gkm_handle*)h)->ghost_fills++;
    return GKM_OK;
}
}

static int failures = 0;
#define CHECK(cond)                                                                 \
    do {                                                                            \
        if (!(cond)) {                                                              \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond); \
            failures++;                                                             \
        }                                                                           \
    } while (0)

static gkm_params gas(int recon, int energy, int scheme) {
    gkm_params p;
    std::memset(&p, 0, sizeof p);
    p.gamma = 1.4;
    p.R = 287.0;
    p.Pr = scheme == GKM_SCHEME_GKS ? 1.0 : 0.7;
    p.mu0 = 1e-3;
    p.T0 = 1.0;
    p.visc_exp = 0.76;
    p.K = 2;
    p.visc_law = GKM_VISC_POWER;
    p.recon = recon;
    p.energy_recon = energy;
    p.zbug_compat = 1;
    p.precision = GKM_PREC_FP64;
    p.flux_scheme = scheme;
    return p;
}

// the generator through the stand-in handle: a padded state of finite values
static void run_ic(int nx, int ny, int nz, gkm_handle& h) {
    h.g.nx = nx;
    h.g.ny = ny;
    h.g.nz = nz;
    h.g.ghost = 3;
    h.g.dx = 2 * M_PI / nx;
    h.p = gas(GKM_RECON_JS, GKM_ENERGY_T, GKM_SCHEME_WENO_BGK);
    gkm_ic ic = {1234ull + (unsigned long long)nx, 4.0, 0.4, 60.0, 1.0};
    gkm_ic_info info;
    CHECK(gkm_init_dhit((gkm_t*)&h, &ic, &info) == GKM_OK);
    CHECK(h.ghost_fills == 1);
    CHECK(std::isfinite(info.t0) && info.t0 > 0);
    for (int q = 0; q < 5; q++)
        for (double x : h.W[q]) CHECK(std::isfinite(x));
    // the generator's own argument checks
    std::vector<double> u((size_t)nx * ny * nz), v(u.size()), w(u.size());
    CHECK(gkm_dhit_velocity(0, ny, nz, &ic, 1.4, u.data(), v.data(), w.data(), nullptr) != GKM_OK);
    CHECK(gkm_dhit_velocity(nx, ny, nz, nullptr, 1.4, u.data(), v.data(), w.data(), nullptr) != GKM_OK);
}

#ifndef SAN_IC_ONLY
static void ghosts(const gkm_grid& g, std::vector<double> (&W)[5]) {
    double* Wp[5];
    for (int q = 0; q < 5; q++) Wp[q] = W[q].data();
    orc_fill_ghosts(&g, Wp);
}

// the oracle's entry points on one box, every scheme switch once
static void run_oracle(gkm_handle& h) {
    const gkm_grid& g = h.g;
    const size_t n = h.W[0].size();
    const int cases[4][3] = {{GKM_RECON_JS, GKM_ENERGY_T, GKM_SCHEME_WENO_BGK},
                             {GKM_RECON_SYMOO, GKM_ENERGY_E, GKM_SCHEME_WENO_BGK},
                             {GKM_RECON_SYMOO, GKM_ENERGY_T, GKM_SCHEME_WENO_BGK},
                             {GKM_RECON_JS, GKM_ENERGY_T, GKM_SCHEME_GKS}};
    for (const auto& cs : cases) {
        gkm_params p = gas(cs[0], cs[1], cs[2]);
        p.mu0 = h.p.mu0;
        p.T0 = h.p.T0;
        std::vector<double> W[5];
        double* Wp[5];
        for (int q = 0; q < 5; q++) {
            W[q] = h.W[q];
            Wp[q] = W[q].data();
        }
        ghosts(g, W);
        const double dt = orc_timestep(&g, &p, Wp, 0.5);
        CHECK(dt > 0 && std::isfinite(dt));
        std::vector<double> F(15 * n, 0.0);
        double* Fp[15];
        for (int q = 0; q < 15; q++) Fp[q] = F.data() + (size_t)q * n;
        orc_fluxes(&g, &p, Wp, dt, Fp, 2);
        orc_step(&g, &p, Wp, dt, 2, 2);
        gkm_stats_t st;
        orc_stats(&g, &p, Wp, &st);
        CHECK(std::isfinite(st.ek) && st.ek > 0);
        gkm_print_stats_t ps;
        orc_print_stats(&g, &p, Wp, &ps);
        CHECK(std::isfinite(ps.tke));
    }
    // point-wise entry points
    const gkm_params p = gas(GKM_RECON_JS, GKM_ENERGY_T, GKM_SCHEME_WENO_BGK);
    double cells[6][5], Wl[5], Wr[5];
    for (int s = 0; s < 6; s++) {
        const double rho = 1 + 0.05 * s, u = 0.1 * s, pr = 1 + 0.02 * s;
        cells[s][0] = rho;
        cells[s][1] = rho * u;
        cells[s][2] = 0.1 * rho;
        cells[s][3] = -0.05 * rho;
        cells[s][4] = pr / 0.4 + 0.5 * rho * (u * u + 0.0125);
    }
    orc_weno(&p, cells, Wl, Wr);
    for (int q = 0; q < 5; q++) CHECK(std::isfinite(Wl[q]) && std::isfinite(Wr[q]));
    double sl[6] = {1, 1.5, 1.2, 3, 2.5, 2}, wl, wr;
    orc_weno_scalar(GKM_RECON_SYMOO, sl, &wl, &wr);
    orc_weno_scalar(GKM_RECON_JS, sl, &wl, &wr);
    double D[5] = {0.01, -0.02, 0.015, 0.0, 0.03}, F[5];
    orc_flux(&p, cells[2], cells[3], Wl, Wr, D, D, D, D, D, D, 1e-3, 1e-3, 0.05, F);
    for (int q = 0; q < 5; q++) CHECK(std::isfinite(F[q]));
    std::vector<double> f(40), Ds(40);
    for (int i = 0; i < 40; i++) f[i] = std::sin(0.3 * i);
    orc_spline_slopes(f.data(), 1, 40, 0.3, Ds.data());
    orc_spline_slopes(f.data(), 1, 2, 0.3, Ds.data());
}

// the file formats: write, read back, compare; and the readers' error paths
static void run_io(const gkm_handle& h, const char* dir) {
    const int nc = h.g.nx;
    const size_t n = (size_t)(nc + 6) * (nc + 6) * (nc + 6);
    gkm_weno_params wp = {nc, 1.4, 1.0, 0.4, 60.0, 1e-3, 0.7, 2.5}, rp;
    const std::string pf = std::string(dir) + "/params";
    CHECK(gkm_write_weno_params(pf.c_str(), &wp) == GKM_OK);
    CHECK(gkm_read_weno_params(pf.c_str(), &rp) == GKM_OK);
    CHECK(rp.nc == wp.nc && rp.Mt == wp.Mt && rp.t0 == wp.t0);
    const double* Win[5];
    for (int q = 0; q < 5; q++) Win[q] = h.W[q].data();
    for (int dec = 0; dec < 2; dec++) {
        const int pd[3] = {dec ? 2 : 1, 1, dec ? 2 : 1};
        CHECK(gkm_write_weno_init(dir, nc, pd, 1.4, Win) == GKM_OK);
        std::vector<double> R[5];
        double* Rp[5];
        for (int q = 0; q < 5; q++) {
            R[q].assign(n, 0.0);
            Rp[q] = R[q].data();
        }
        CHECK(gkm_read_weno_init(dir, nc, pd, 1.4, Rp) == GKM_OK);
        double err = 0, ref = 0;
        for (int j = 3; j < nc + 3; j++)
            for (int i = 3; i < nc + 3; i++)
                for (int k = 3; k < nc + 3; k++) {
                    const size_t c = ((size_t)j * (nc + 6) + i) * (nc + 6) + k;
                    for (int q = 0; q < 5; q++) {
                        err = std::fmax(err, std::fabs(R[q][c] - h.W[q][c]));
                        ref = std::fmax(ref, std::fabs(h.W[q][c]));
                    }
                }
        CHECK(err <= 1e-12 * ref);
    }
    const int bad[3] = {3, 1, 1};
    std::vector<double> R0(n);
    double* Rb[5] = {R0.data(), R0.data(), R0.data(), R0.data(), R0.data()};
    if (nc % 3) CHECK(gkm_read_weno_init(dir, nc, bad, 1.4, Rb) != GKM_OK);
    CHECK(gkm_read_weno_params("/nonexistent/params", &rp) != GKM_OK);
    // Init.dat (FP32 velocities, j fastest)
    gkm_init_dat hd = {nc, 1.0f, 0.4f, 60.0f, 4.0f, 0.2f, 1e-3f, 0.8f, 0.7f, 2.5f}, hr;
    const size_t m = (size_t)nc * nc * nc;
    std::vector<float> u(m), v(m), w(m), ur(m), vr(m), wr(m);
    for (size_t i = 0; i < m; i++) {
        u[i] = 0.1f * std::sin(0.01f * i);
        v[i] = 0.1f * std::cos(0.02f * i);
        w[i] = -0.05f;
    }
    const std::string df = std::string(dir) + "/Init.dat";
    CHECK(gkm_write_init_dat(df.c_str(), &hd, u.data(), v.data(), w.data()) == GKM_OK);
    CHECK(gkm_read_init_dat(df.c_str(), &hr, ur.data(), vr.data(), wr.data()) == GKM_OK);
    CHECK(hr.n == nc && std::memcmp(u.data(), ur.data(), m * sizeof(float)) == 0);
    CHECK(gkm_read_init_dat(df.c_str(), &hr, nullptr, nullptr, nullptr) == GKM_OK);
    std::vector<double> S[5];
    double* Sp[5];
    for (int q = 0; q < 5; q++) {
        S[q].assign(n, 0.0);
        Sp[q] = S[q].data();
    }
    CHECK(gkm_init_dat_state(&hr, ur.data(), vr.data(), wr.data(), 1.4, Sp) == GKM_OK);
}
#endif

int main(int argc, char** argv) {
    const char* dir = argc > 1 ? argv[1] : ".";
    // a box whose edges are not powers of two (Bluestein transforms) and a cube
    gkm_handle a, b;
    run_ic(12, 10, 9, a);
    run_ic(16, 16, 16, b);
#ifndef SAN_IC_ONLY
    run_oracle(a);
    run_io(b, dir);
#else
    (void)dir;
#endif
    if (failures) {
        std::fprintf(stderr, "san_driver: %d check(s) failed\n", failures);
        return 1;
    }
    std::printf("san_driver: ok\n");
    return 0;
}
