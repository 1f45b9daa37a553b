/*
 * gkm_oracle.h — CPU oracle of the GKM-DNS time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only as the checker
 * or the timed CPU baseline; the product (libgkm.so) never links it.
 *
 * A plain FP64 C++ restatement of the weno CPU path: flux()
 * (weno/src/GKMflux.cpp:11-308), ApplyWENO SYMOO / JS
 * (weno/src/update.cpp:502-625, 1421-1513), derivsX/Y/Z + update
 * (update.cpp:201-469), PeriodicBC (FlowField.cpp:115-275), timestep
 * (update.cpp:1057-1079) and interior statistics; the GKM tree's smooth-flow
 * scheme (GKM/CUDA GKM/src/CPUupdate.cpp:29-210, GKMflux.cpp:12-280) when
 * params->flux_scheme is GKM_SCHEME_GKS.
 */
#ifndef GKM_ORACLE_H
#define GKM_ORACLE_H

#include "../include/gkm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* BGK interface flux with the reference's argument list; tau on entry is
 * mu/p of the left cell (collision-time blending is added inside). */
void orc_flux(const gkm_params* p, const double WL[5], const double WR[5],
              const double Wl[5], const double Wr[5],
              const double DWxl[5], const double DWxr[5],
              const double DWyl[5], const double DWyr[5],
              const double DWzl[5], const double DWzr[5],
              double tau, double dt, double dx, double F[5]);

/* Smooth-flow gas-kinetic flux of the GKM tree with its argument list
 * (GKM/CUDA GKM/src/GKMflux.cpp:12-280): interface average W0 and slopes in
 * the rotated slot order, viscosity mu, time step dt. */
void orc_gks_flux(const gkm_params* p, const double W0[5], const double DWx[5], const double DWy[5],
                  const double DWz[5], double mu, double dt, double F[5]);

/* Vector WENO at one interface: cells (LLL, LL, L, R, RR, RRR) x 5 fields
 * -> Wl[5], Wr[5].  p->recon selects SYMOO / JS; p->energy_recon selects
 * temperature or total-energy reconstruction of field 4. */
void orc_weno(const gkm_params* p, const double cells[6][5], double Wl[5], double Wr[5]);

/* Scalar WENO line (one field): returns the left / right interface values. */
void orc_weno_scalar(int recon, const double c[6], double* wl, double* wr);

/* The moment helpers, exported for closed-form tests. */
void orc_moment_tables(double U, double lam, int K, double half_pos[7], double half_neg[7],
                       double full[7], double* e2, double* e4);
void orc_slopesolver(const gkm_params* p, const double b[5], const double U[3], double lam, double a[5]);

/* Whole-box operations on padded host arrays (layout as in gkm.h). */
void orc_fill_ghosts(const gkm_grid* g, double* const W[5]);
/* nsteps x (ghost fill, X/Y/Z sweeps, update), then a final ghost fill.
 * nthreads <= 1 runs single-threaded like the reference. */
void orc_step(const gkm_grid* g, const gkm_params* p, double* const W[5], double dt, int nsteps,
              int nthreads);
/* Flux arrays of one step without the update (for kernel-level checks).
 * F must hold 15 padded arrays: Fx0..4, Fy0..4, Fz0..4. */
void orc_fluxes(const gkm_grid* g, const gkm_params* p, double* const W[5], double dt,
                double* const F[15], int nthreads);
double orc_timestep(const gkm_grid* g, const gkm_params* p, double* const W[5], double cfl);
void orc_stats(const gkm_grid* g, const gkm_params* p, double* const W[5], gkm_stats_t* out);

/* Slopes D[0..m-1] at the nodes of the natural cubic spline through the m >= 2
 * values f[0], f[stride], ..., f[(m-1)*stride] spaced h apart. */
void orc_spline_slopes(const double* f, long stride, int m, double h, double* D);
/* PrintStats (weno/src/update.cpp:638-784) with natural-spline gradients over
 * the padded box (after a ghost fill of a copy of W). */
void orc_print_stats(const gkm_grid* g, const gkm_params* p, double* const W[5], gkm_print_stats_t* out);

#ifdef __cplusplus
}
#endif
#endif
