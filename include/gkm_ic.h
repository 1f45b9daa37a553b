/*
 * gkm_ic.h — decaying isotropic turbulence initial condition for libgkm.so.
 *
 * The reference generates its initial condition offline with a MATLAB script
 * (GKM/CUDA GKM/InitialConditions/init_BS.m:24-157), writes it to files, and
 * reads them in GetParams / Initialize_DHIT (weno/src/FlowField.cpp:9-112).
 * These calls generate the same recipe in the library instead:
 *   u, v, w ~ U[0, 1) per cell, Fourier transformed; projected onto the
 *   solenoidal subspace, u_hat -= (k . u_hat) k / |k|^2; every shell
 *   |k| in [m - 1/2, m + 1/2) rescaled to
 *     E(m) = sqrt(2/pi) (16/3) Mt^2 m^4 exp(-2 m^2 / k0^2) N^2 / k0^5
 *   (N = nx ny nz); transformed back, real part.  Even extents have their
 *   Nyquist planes set to zero so the field stays real.
 *   u_rms = sqrt(<u_i u_i>/3), c0 = sqrt(3) u_rms / Mt, p0 = rho0 c0^2 / gamma,
 *   mu0 = u_rms (2/k0) rho0 / Re, t0 = (2/k0) / u_rms.
 * The random numbers come from SplitMix64 seeded with `seed` (u, then v, then
 * w, each in the [j][i][k] order below), so a seed gives the same field on
 * every machine; the spectrum and the derived scales do not depend on it.
 */
#ifndef GKM_IC_H
#define GKM_IC_H

#include "gkm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gkm_ic {
    unsigned long long seed;
    double k0;    /* peak wavenumber of the spectrum                         */
    double Mt;    /* turbulent Mach number sqrt(<u_i u_i>) / c0              */
    double Re;    /* Reynolds number on u_rms and 2/k0 (the Taylor microscale
                     of this spectrum)                                        */
    double rho0;  /* uniform initial density                                 */
} gkm_ic;

typedef struct gkm_ic_info {
    double q2;    /* <u_i u_i>                                                */
    double urms;  /* sqrt(q2 / 3)                                             */
    double c0;    /* sqrt(3) urms / Mt                                        */
    double p0;    /* rho0 c0^2 / gamma                                        */
    double mu0;   /* urms (2/k0) rho0 / Re                                    */
    double t0;    /* eddy turnover time (2/k0) / urms                         */
} gkm_ic_info;

/* The velocity field on the host: u, v, w receive nx*ny*nz values each in
 * the unpadded order [j][i][k] (j along y slowest, k along z fastest); u is
 * the x component.  info (may be null) receives the derived scales for
 * gamma.  No device is used.  Replaces running init_BS.m. */
int gkm_dhit_velocity(int nx, int ny, int nz, const gkm_ic* ic, double gamma, double* u, double* v, double* w,
                      gkm_ic_info* info);

/* Fill box `box` of a handle with the initial condition (uniform rho0 and p0,
 * the velocity above, conservative variables as weno/src/FlowField.cpp:97-102
 * forms them), fill the ghosts, and set the box's mu0 and T0 = p0 / (R rho0)
 * to the values the recipe derives (what GetParams reads from the params
 * file, FlowField.cpp:9-57).  Replaces GetParams + Initialize_DHIT. */
int gkm_init_dhit_box(gkm_t* h, int box, const gkm_ic* ic, gkm_ic_info* info);
int gkm_init_dhit(gkm_t* h, const gkm_ic* ic, gkm_ic_info* info);

#ifdef __cplusplus
}
#endif

#endif /* GKM_IC_H */
