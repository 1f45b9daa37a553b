/*
 * ref_flux_harness.cpp — C entry point around the reference's own flux()
 * (weno/src/GKMflux.cpp:11-215), compiled from the reference sources where
 * they lie (oracle/Makefile, target `ref`).  TEST INFRASTRUCTURE ONLY: used
 * to pin the oracle's restatement and to write tests/golden/ fixtures.
 *
 * The reference reads the gas constants from the globals K and gam
 * (weno/src/param.cpp:4-7); this harness sets them before each call.
 */
#include "GKMflux.h"

extern "C" {

/* rec = WL WR Wl Wr DWxl DWxr DWyl DWyr DWzl DWzr (5 each), tau, dt */
void ref_flux_points(int K_, double gam_, int n, const double* rec, double dx, double* F) {
    ::K = K_;
    ::gam = gam_;
    for (int r = 0; r < n; r++) {
        const double* x = rec + (long)r * 52;
        ptype v[10][5];
        for (int a = 0; a < 10; a++)
            for (int q = 0; q < 5; q++) v[a][q] = x[a * 5 + q];
        flux(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], v[9], x[50], x[51], dx, F + (long)r * 5);
    }
}

}
