/*
 * gkm.h — C ABI of the MI355X GKM-DNS time step (libgkm.so).
 *
 * One handle owns one periodic box (or one slab) of the conservative state
 * W = (rho, rho*u, rho*v, rho*w, rho*E) on the GPU, with 3 ghost layers per
 * side.  Host arrays exchanged with the library use the reference layout:
 * five separate FP64 arrays, each (ny+6)*(nx+6)*(nz+6) long, index
 *     Is(j, i, k) = j*(nx+6)*(nz+6) + i*(nz+6) + k
 * with j along y (slowest), i along x, k along z (unit stride)
 * (weno/src/param.h:37).  Velocity components follow the axes: W1 = rho*u
 * (x = axis 1), W2 = rho*v (y = axis 0), W3 = rho*w (z = axis 2).
 *
 * Every function returns 0 (GKM_OK) or a negative status; the message of the
 * last failure on the calling thread is returned by gkm_last_error().
 * No C++ exception crosses this boundary.  Calls on one handle are not
 * thread-safe; separate handles may be used from separate threads.
 */
#ifndef GKM_H
#define GKM_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define GKM_OK             0
#define GKM_E_ARG         -1   /* invalid argument (null pointer, bad size)  */
#define GKM_E_HIP         -2   /* a HIP runtime call failed                    */
#define GKM_E_NONPHYSICAL -3   /* rho <= 0, p <= 0 or non-finite state         */
#define GKM_E_NOMEM       -4   /* device allocation failed                     */
#define GKM_E_UNSUPPORTED -5   /* option not available in this build           */

/* ---- option values -------------------------------------------------------- */
#define GKM_VISC_SUTHERLAND 0  /* mu0 (T/T0)^1.5 (T0+110.4)/(T+110.4), weno/src/update.cpp:263 */
#define GKM_VISC_POWER      1  /* mu0 (T/T0)^visc_exp, GKM/CUDA GKM/src/GPUupdate.cu:1327    */

#define GKM_RECON_SYMOO 0      /* WENO-SYMOO, weno/src/update.cpp:502-625          */
#define GKM_RECON_JS    1      /* WENO-JS 5th order, weno/src/update.cpp:1421-1513 */

#define GKM_ENERGY_T 0         /* reconstruct T, rebuild rho*E (CPU path, update.cpp:560-619) */
#define GKM_ENERGY_E 1         /* reconstruct rho*E directly (GPU path, GPUupdate.cu:546)     */

#define GKM_HALO_GATHER 0      /* every ghost cell gathers its periodic image (fused with the
                                  per-cell primitives)                                          */
#define GKM_HALO_PACK   1      /* y faces packed into a send buffer and unpacked from a receive
                                  buffer (the single-GPU stand-in for a slab exchange), then the
                                  x/z ghosts filled in place                                      */

#define GKM_TIMING_OFF  0      /* gkm_set_timing: no events                                     */
#define GKM_TIMING_ALL  1      /* an event around every launch of the step                     */
#define GKM_TIMING_FLUX 2      /* events around the X and Y flux kernels only (least perturbing) */

#define GKM_PREC_FP64  0       /* FP64 everywhere (drop-in form)                               */
#define GKM_PREC_MIXED 1       /* FP32 reconstruction + flux, FP32 hi/lo compensated state     */

#define GKM_SCHEME_WENO_BGK 0  /* WENO reconstruction + BGK-NS flux with half-space moments
                                  (weno/src/update.cpp:194-469, GKMflux.cpp:11-215)              */
#define GKM_SCHEME_GKS      1  /* smooth-flow gas-kinetic flux of the GKM tree: interface
                                  average and central-difference slopes, one Maxwellian
                                  (GKM/CUDA GKM/src/CPUupdate.cpp:29-210, GKMflux.cpp:12-280)    */

/* Grid of one box.  The reference uses one spacing for all axes
 * (dx = 2*pi/nc, weno/src/FlowField.cpp:45). */
typedef struct gkm_grid {
    int    nx, ny, nz;   /* interior cells along x, y, z                      */
    int    ghost;        /* ghost layers per side; must be 3                  */
    double dx;           /* cell size                                         */
} gkm_grid;

/* Gas and scheme parameters.  In the reference these are compile-time
 * macros and globals (weno/src/param.h:11-23, FlowField.cpp:40-49). */
typedef struct gkm_params {
    double gamma;        /* ratio of specific heats (gam)                     */
    double R;            /* gas constant in T = p/(R rho) (param.h:18: 287)   */
    double Pr;           /* Prandtl number applied in the slope solver on a[4]
                            (GKMflux.cpp:289 uses 0.7; GPUupdate.cu:839 uses 1) */
    double mu0;          /* reference viscosity                               */
    double T0;           /* reference temperature                             */
    double visc_exp;     /* exponent of GKM_VISC_POWER (0.76)                 */
    int    K;            /* internal degrees of freedom, (5-3 gamma)/(gamma-1) */
    int    visc_law;     /* GKM_VISC_*                                        */
    int    recon;        /* GKM_RECON_*                                       */
    int    energy_recon; /* GKM_ENERGY_*                                      */
    int    zbug_compat;  /* 1: Z sweep pairs the y-slot slope with d/dx and the
                            z-slot slope with d/dy, as update.cpp:380-423 does;
                            0: physically consistent pairing                  */
    int    precision;    /* GKM_PREC_*                                        */
    int    flux_scheme;  /* GKM_SCHEME_* (0, the weno tree's scheme, when the
                            struct is zero-initialised)                       */
} gkm_params;

/* Statistics of the interior cells, all accumulated in FP64.
 * Means are over the nx*ny*nz interior cells. */
typedef struct gkm_stats_s {
    double tke;        /* <rho u_i u_i>   (update.cpp PrintStats "TKE")        */
    double ek;         /* 0.5 <rho u_i u_i>                                    */
    double q2;         /* <u_i u_i>                                            */
    double mt;         /* sqrt(<u_i u_i>) / <c>, c = sqrt(gamma p / rho)       */
    double rho_mean, rho_rms;
    double p_mean, p_rms;
    double T_mean, T_rms;
    double v_mean, v_rms;       /* specific volume 1/rho                       */
    double c_mean;
    double umax_plus_c;         /* max(|u| + c), the CFL signal speed          */
    /* velocity-gradient statistics, 2nd-order central differences */
    double dudx2, dudx3, dudx4; /* <(du/dx)^n>                                 */
    double dvdy2, dvdy3, dvdy4;
    double dwdz2, dwdz3, dwdz4;
    double skew_u, skew_v, skew_w;
    double kurt_u, kurt_v, kurt_w;
    double enstrophy;           /* <|curl u|^2>                                */
    double dil2;                /* <(div u)^2>                                 */
    double diss_sol;            /* <mu |curl u|^2>                             */
    double diss_dil;            /* 4/3 <mu (div u)^2>                          */
    double diss_ref;            /* <mu du_i/dx_j du_i/dx_j> (PrintStats Diss)  */
    double sijsij;              /* <S_ij S_ij> as PrintStats forms it          */
    double mass, mom_x, mom_y, mom_z, energy; /* interior sums of W0..W4       */
} gkm_stats_t;

/* The quantities PrintStats writes (weno/src/update.cpp:638-784), FP64.
 * Velocity gradients are the slopes at the grid nodes of the natural cubic
 * spline through the padded box's values of u = rho*u/rho (and v, w) along
 * each axis.  The reference evaluates the gradient of a tensor-product cubic
 * B-spline with natural ends built over the padded segment
 * (get_bsplinep, weno/src/Interpolation.cpp:3-31; generateSplines,
 * update.cpp:1215-1231); at grid nodes that gradient is exactly these 1-D
 * slopes.  Means are over the nx*ny*nz interior cells; mu follows the
 * handle's viscosity law (PrintStats uses Sutherland, update.cpp:698).
 * rms values are unnormalised; gkm_format_print_stats applies the
 * reference's scalings. */
typedef struct gkm_print_stats_s {
    double tke;                 /* <rho u_i u_i>                   (TKE)      */
    double diss;                /* <mu du_i/dx_j du_i/dx_j>        (Diss)     */
    double sijsij;              /* <S_ij S_ij>                     (SijSij)   */
    double skew_u, skew_v, skew_w;   /* <(du/dx)^3>/<(du/dx)^2>^1.5 ...    */
    double kurt_u, kurt_v, kurt_w;   /* <(du/dx)^4>/<(du/dx)^2>^2 ...      */
    double p_mean, T_mean, v_mean;   /* v = 1/rho                          */
    double p_rms, T_rms, v_rms;      /* rms about the means                */
    double dudx2, dvdy2, dwdz2;      /* <(du/dx)^2>, <(dv/dy)^2>, <(dw/dz)^2> */
} gkm_print_stats_t;

typedef struct gkm_ctx gkm_t;

/* Create a handle on the current HIP device and allocate the device state.
 * Replaces GetParams/segment/DevAlloc (weno/src/FlowField.cpp:9-57,
 * FlowField.cpp:278-288, GPUAlloc.cu:3-17). */
int gkm_create(const gkm_grid* grid, const gkm_params* params, gkm_t** out);
int gkm_destroy(gkm_t* h);

/* A batch of nbox independent boxes of one grid, advanced by the same
 * launches (one launch per stage covers every box).  params[b] gives box b's
 * gas constants (gamma, R, Pr, mu0, T0, visc_exp, K, visc_law may differ);
 * recon, energy_recon, zbug_compat, precision and flux_scheme must agree.  No reference
 * counterpart: the reference runs one segment per MPI rank.  The *_box calls
 * address one box; the single-box calls above act on box 0. */
int gkm_create_batch(const gkm_grid* grid, const gkm_params* params, int nbox, gkm_t** out);
int gkm_batch_size(gkm_t* h, int* nbox);
int gkm_set_state_box(gkm_t* h, int box, const double* const W[5]);
int gkm_get_state_box(gkm_t* h, int box, double* const W[5]);
/* dt[b] per box; nsteps steps of every box. */
int gkm_step_batch(gkm_t* h, const double* dt, int nsteps);
/* Set dt[b] per box for later steps without stepping (for gkm_group_step). */
int gkm_set_dt_batch(gkm_t* h, const double* dt);
int gkm_compute_dt_box(gkm_t* h, int box, double cfl, double* dt);
int gkm_stats_box(gkm_t* h, int box, gkm_stats_t* out);

/* The handle's grid, and box `box`'s gas parameters. */
int gkm_get_grid(gkm_t* h, gkm_grid* out);
int gkm_get_params_box(gkm_t* h, int box, gkm_params* out);
/* Replace box `box`'s gas constants (gamma, R, Pr, mu0, T0, visc_exp, K,
 * visc_law); recon, energy_recon, zbug_compat, precision and flux_scheme must
 * equal the handle's.  Takes effect from the next step (also in a captured
 * graph, which reads the constants from the device table).  No reference
 * counterpart: the reference's constants are globals set once by GetParams. */
int gkm_set_params_box(gkm_t* h, int box, const gkm_params* p);

/* Capture one step (ghost fill, three sweeps, update) in a hipGraph on first
 * use and replay it for every later step.  dt and the gas constants are read
 * from a device table at replay, so changing them needs no new capture.  Off
 * by default.  Steps timed with gkm_set_timing are launched without the
 * graph. */
int gkm_set_graph(gkm_t* h, int enabled);

/* A group of handles on the current device stepped together, for example an
 * ensemble of uneven box sizes with one batch handle per size.  One captured
 * hipGraph holds one step of every handle, each handle's launches on a branch
 * of their own, so one graph launch advances every box of every handle.  Each
 * handle steps with the dt of its last gkm_step_batch or gkm_set_dt_batch.
 * The handles stay usable on their own between group calls (every group call
 * is ordered after the work already queued on them, and their later work after
 * it); destroy the group before its handles.  No reference counterpart. */
typedef struct gkm_group gkm_group_t;
int gkm_group_create(gkm_t* const* handles, int n, gkm_group_t** out);
int gkm_group_step(gkm_group_t* g, int nsteps);
/* The same steps as gkm_group_step, bit for bit, from one persistent launch:
 * a cooperative kernel whose grid fits on the device at once runs the six
 * phases of every step (WENO and flux X, WENO and flux Y, WENO Z, fused Z flux
 * + update) over every box of every handle, with a grid-wide barrier between
 * phases, for all nsteps steps.  The handles must share precision,
 * reconstruction and energy mode and use the WENO-BGK scheme with
 * GKM_HALO_GATHER (GKM_E_ARG otherwise); at most 16 handles.  The kernel's
 * blocks are 256 threads whatever gkm_set_tuning chose for the flux blocks
 * (its WENO chunk is honoured).  No reference counterpart (the reference runs
 * one realization per process, weno/src/Main.cpp:10-49). */
int gkm_group_step_persistent(gkm_group_t* g, int nsteps);
int gkm_group_sync(gkm_group_t* g);
int gkm_group_destroy(gkm_group_t* g);

/* Launch shapes: threads per block of the X and Y flux kernels (64, 128 or
 * 256) and interfaces per thread of the strided-sweep WENO kernels (4, 8 or
 * 16).  Chosen per box size by the autotuner in gkm_amd/tune.py.  Drops the
 * handle's captured step; groups holding the handle capture theirs again on
 * their next step.  Results never depend on the launch shape.  No reference
 * counterpart (the reference's launch shapes are compile-time constants,
 * weno/src/GPUupdate.cu:65-196). */
int gkm_set_tuning(gkm_t* h, int flux_block, int recon_chunk);
/* The fused Z sweep's launch shape: threads per block (64, 128 or 256;
 * created as 128 mixed, 64 FP64) and the width of the strips of columns its
 * z lines are taken in (4, 8, 16, 32 or 64; created as 16), so that a line's
 * neighbours are read while they are still in one XCD's L2.  Same rules and
 * the same autotuner as gkm_set_tuning.  No reference counterpart. */
int gkm_set_tuning_z(gkm_t* h, int zflux_block, int z_strip);

/* Copy the padded host state in / out (bit-exact round trip for FP64; the
 * mixed form splits each value into an FP32 hi/lo pair and rebuilds it).
 * Replaces Initialize_DHIT's writes into W and the H2D/D2H copies of
 * evolve() (FlowField.cpp:59-112, GPUupdate.cu:74-80,186-190). */
int gkm_set_state(gkm_t* h, const double* const W[5]);
int gkm_get_state(gkm_t* h, double* const W[5]);

/* 3-deep periodic ghost fill on the device (replaces PeriodicBC and the
 * one-rank HaloTransfer, FlowField.cpp:115-275, update.cpp:894-1054). */
int gkm_fill_ghosts(gkm_t* h);

/* Advance nsteps steps of size dt: per step, ghost fill, X/Y/Z WENO + BGK
 * flux sweeps and the conservative update; a final ghost fill leaves the
 * state as the reference leaves it after HaloTransfer.  Replaces
 * evolve(W, dev, dt) / evolveEulerian (GPUupdate.cu:65-196,
 * update.cpp:180-199).  Asynchronous on the handle's stream.
 *
 * Failure detection (the reference checks nothing, GPUupdate.cu:11-20): the
 * update raises a per-box flag when a new cell has rho or T = p/(R rho) not
 * positive and finite.  The flag is sticky until gkm_set_state(_box) replaces
 * that box's state.  gkm_sync, gkm_compute_dt, gkm_stats, gkm_print_stats,
 * gkm_step_print_stats and gkm_group_sync return GKM_E_NONPHYSICAL once the
 * work that raised it has completed; gkm_step / gkm_step_batch /
 * gkm_group_step return it without waiting, and without stepping, once the
 * flag copy queued by an earlier stepping call has landed.  gkm_get_state
 * copies the state either way. */
int gkm_step(gkm_t* h, double dt, int nsteps);

/* dt = cfl*dx / max_interior(|u| + sqrt(gamma p/rho)), FP64 device
 * reduction (replaces timestep, update.cpp:1057-1079). */
int gkm_compute_dt(gkm_t* h, double cfl, double* dt);

/* FP64 device reductions of the interior (replaces the sums of PrintStats,
 * update.cpp:638-784, and the GKM on-device TKE/PRMS, GPUupdate.cu:425-483). */
int gkm_stats(gkm_t* h, gkm_stats_t* out);

/* PrintStats (weno/src/update.cpp:638-784) with its spline gradients
 * (generateSplines + eval_UBspline_3d_d_vg, update.cpp:1215-1231, :679-681):
 * ghost fill, natural-spline slopes of u, v, w along every grid line of the
 * padded box, then FP64 interior reductions.  The first call allocates nine
 * FP64 gradient planes of one box (freed by gkm_destroy). */
int gkm_print_stats(gkm_t* h, gkm_print_stats_t* out);
int gkm_print_stats_box(gkm_t* h, int box, gkm_print_stats_t* out);
/* The body of the reference's Evolve loop (weno/src/update.cpp:131-146:
 * PrintStats of the current state, then the step, t += dt), nsteps times with
 * a fixed dt: out[s] receives PrintStats of the state before step s (nsteps
 * entries).  One-box handles.  The statistics of state s run on a second
 * stream of the handle while step s runs (both read state s; step s writes
 * the other state buffer), so each pair costs about the longer of the two
 * rather than their sum.  The results equal gkm_print_stats then gkm_step
 * called in turn; with GKM_HALO_PACK the two are called in turn.  Returns once
 * the last statistics are on the host (the last step may still run, as after
 * gkm_step). */
int gkm_step_print_stats(gkm_t* h, double dt, int nsteps, gkm_print_stats_t* out);
/* PrintStats' stdout line (update.cpp:778-781), "%.8e" columns:
 *   t/t0  TKE  Diss  SKEWu  SKEWv  SKEWw  KURTu  KURTv  KURTw
 *   Prms/(gamma p0 Mt^2)  Trms/((gamma-1) T0 Mt^2)  Vrms/(Mt^2/den0)
 * with the reference's separators; no newline.  Returns GKM_E_ARG if the
 * line does not fit in len bytes. */
int gkm_format_print_stats(const gkm_print_stats_t* s, double t_over_t0, double gamma, double p0, double T0,
                           double Mt, double den0, char* buf, int len);

/* Wait for all work queued on the handle's stream; GKM_E_NONPHYSICAL if a
 * step left a non-physical state (see gkm_step). */
int gkm_sync(gkm_t* h);

/* Device time of the last gkm_step call, from HIP events recorded on the
 * handle's stream around each launch (first 64 steps of the call):
 * ms[0] BGK flux kernels, ms[1] WENO reconstruction kernels, ms[2] ghost
 * fill (in GKM_HALO_GATHER mode fused with the primitives), ms[3]
 * conservative updates, ms[4] the ghost shell's primitives in
 * GKM_HALO_PACK mode (the fused Z sweep forms the interior's; a full-box pass
 * after a state change counts in ms[2]; 0 in GKM_HALO_GATHER mode); -1 where
 * nothing was recorded.
 * *nsteps receives the number of timed steps.  Timing is off by default.
 * mode: GKM_TIMING_ALL (any other nonzero value too) records every event;
 * GKM_TIMING_FLUX records only those around the X and Y flux kernels, so the
 * step runs with four events instead of thirteen: ms[0] is then those two
 * kernels' time and ms[1..4] are -1 (gkm_last_sweep_timing gives X and Y, Z
 * -1). */
int gkm_set_timing(gkm_t* h, int mode);
int gkm_last_timing(gkm_t* h, double ms[5], int* nsteps);
/* Flux-kernel time of each sweep direction (X, Y, Z) over the timed steps of
 * the last gkm_step call; the Z entry is the flux kernel that also applies
 * the conservative update (so ms[3] of gkm_last_timing is ~0).  -1 where
 * nothing was recorded. */
int gkm_last_sweep_timing(gkm_t* h, double ms[3]);

/* Ghost-fill mode (GKM_HALO_*).  The y-face pack/unpack path is the one a
 * slab decomposition exchanges between ranks (replaces the pack, Sendrecv and
 * unpack of HaloTransfer, weno/src/update.cpp:894-1054, y faces :899-936). */
int gkm_set_halo_mode(gkm_t* h, int mode);

/* Point-wise entry points used to check the device arithmetic against the
 * reference on sampled interfaces.  Replace flux() (GKMflux.cpp:11-215) and
 * ApplyWENO / d_ApplyWENO (update.cpp:502-625, GPUupdate.cu:917-977).
 * Inputs are host arrays of n records; the library copies them to the
 * device, runs the device code and copies the result back.
 *   flux:  in[n*52] = WL[5] WR[5] Wl[5] Wr[5] DWxl[5] DWxr[5] DWyl[5] DWyr[5]
 *                     DWzl[5] DWzr[5] tau dt, with dx given;  F[n*5]
 *   weno:  in[n*30] = six cells x 5 fields (LLL, LL, L, R, RR, RRR);
 *          out[n*10] = Wl[5] Wr[5]                                         */
int gkm_flux_points(const gkm_params* p, int n, const double* in, double dx, double* F);
int gkm_weno_points(const gkm_params* p, int n, const double* in, double* out);
/* The GKM tree's smooth-flow flux on single records (replaces flux(),
 * GKM/CUDA GKM/src/GKMflux.cpp:12-280, same argument list):
 *   in[n*22] = W0[5] DWx[5] DWy[5] DWz[5] mu dt (rotated slot order);  F[n*5] */
int gkm_gks_points(const gkm_params* p, int n, const double* in, double* F);

/* gkm_flux_points with no device allocation per call.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Records
 * and fluxes are staged through a static device array of the library (3.7 MB
 * per GPU, held until the process ends) instead of per-call buffers, on the
 * null stream; the results equal gkm_flux_points' bit for bit.  n may be at
 * most 8192 (GKM_E_ARG beyond).  Calls on one GPU share the array and are
 * serialised by a lock in the library. */
int gkm_flux_points_noalloc(const gkm_params* p, int n, const double* in, double dx, double* F);

const char* gkm_last_error(void);

/* Library version string, e.g. "gkm-mi355x 0.1 gfx950". */
const char* gkm_version(void);

#ifdef __cplusplus
}
#endif
#endif /* GKM_H */
