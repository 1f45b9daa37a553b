/*
 * gkm_io.h — file formats of the reference's initial conditions and
 * restart files, read and written by libgkm.so on the host (no GPU needed).
 *
 * weno tree:
 *   InitialConditions/params   int nc, then 7 FP64: gam, den0, Mt, Re, mu0,
 *                              p0, t0 (GetParams, weno/src/FlowField.cpp:9-57)
 *   InitialConditions/Init%i%i%i, one file per rank (procId[0..2] printed
 *                              back to back): u, v, w, den, p of the rank's
 *                              segment, FP64, each nc_seg^3 values with the
 *                              local j (y) fastest, then i (x), then k (z):
 *                              J = k*nsy*nsx + i*nsy + j
 *                              (Initialize_DHIT, FlowField.cpp:59-112; the
 *                              restart writer WriteReInit, update.cpp:1352-1391,
 *                              writes the same layout)
 *   rank (a, b, c) of a procDim decomposition holds global
 *   j in [a*nsy, (a+1)*nsy), i in [b*nsx, ...), k in [c*nsz, ...)
 *   (segment, FlowField.cpp:276-286).
 * GKM tree:
 *   InitialConditions/Init.dat int N; 9 FP32: rho0, Mt, Re, k0, urms, mu,
 *                              crms, p0, eddy_time; then u, v, w, FP32, N^3
 *                              each, j fastest (GKM/CUDA GKM/InitialConditions/
 *                              init_BS.m:160-175, read by Macro::Initialize_DHIT,
 *                              GKM/CUDA GKM/src/FlowField.cpp:178-235)
 *
 * Whole-box arrays exchanged here use the padded layout of gkm.h
 * (Is(j,i,k) = j*(n+6)*(n+6) + i*(n+6) + k, 3 ghosts); readers leave the
 * ghosts at zero.  Status codes as in gkm.h; GKM_E_IO for a file that cannot
 * be opened, is short, or has the wrong size.
 */
#ifndef GKM_IO_H
#define GKM_IO_H

#ifdef __cplusplus
extern "C" {
#endif

#define GKM_E_IO -6  /* file missing, short or of the wrong size */

/* The weno params file. */
typedef struct gkm_weno_params {
    int    nc;                                      /* cells per edge        */
    double gam, den0, Mt, Re, mu0, p0, t0;          /* file order            */
} gkm_weno_params;

/* Replaces GetParams (weno/src/FlowField.cpp:9-57); the globals it derives
 * are K = int((5 - 3 gam)/(gam - 1)), dx = 2 pi/nc, T0 = p0/(R den0). */
int gkm_read_weno_params(const char* path, gkm_weno_params* out);
int gkm_write_weno_params(const char* path, const gkm_weno_params* in);

/* Initialize_DHIT over every rank file of a proc_dim[0] x proc_dim[1] x
 * proc_dim[2] decomposition of an nc^3 box (each proc_dim must divide nc):
 * reads dir/Init%i%i%i and forms W = (den, u den, v den, w den,
 * den (p/(den (gam-1)) + |u|^2/2)) in the padded whole-box layout.
 * Replaces Initialize_DHIT (FlowField.cpp:59-112). */
int gkm_read_weno_init(const char* dir, int nc, const int proc_dim[3], double gam, double* const W[5]);
/* WriteReInit (weno/src/update.cpp:1352-1391): primitives of the interior
 * cells written as dir/Init%i%i%i rank files. */
int gkm_write_weno_init(const char* dir, int nc, const int proc_dim[3], double gam, const double* const W[5]);

/* The GKM Init.dat header. */
typedef struct gkm_init_dat {
    int   n;
    float rho0, Mt, Re, k0, urms, mu, crms, p0, eddy_time;
} gkm_init_dat;

/* Reads the header and, when u, v, w are not null, the three velocity
 * fields into caller arrays of n^3 floats in file order (j fastest). */
int gkm_read_init_dat(const char* path, gkm_init_dat* hdr, float* u, float* v, float* w);
int gkm_write_init_dat(const char* path, const gkm_init_dat* hdr, const float* u, const float* v, const float* w);
/* Macro::Initialize_DHIT's conversion (GKM/CUDA GKM/src/FlowField.cpp:211-227):
 * uniform density rho0 and pressure p0 of the header, velocities from the
 * file, computed in FP64 from the FP32 values -> padded FP64 W (ghosts 0). */
int gkm_init_dat_state(const gkm_init_dat* hdr, const float* u, const float* v, const float* w, double gam,
                       double* const W[5]);

#ifdef __cplusplus
}
#endif
#endif /* GKM_IO_H */
