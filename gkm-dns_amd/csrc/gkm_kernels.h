// gkm_kernels.h — HIP kernels of the GKM-DNS step (gfx950).
//
// Data layout in HBM: five SoA arrays of the padded box, (ny+6)(nx+6)(nz+6),
// z unit stride (the reference's Is(j,i,k), weno/src/param.h:37).  The FP64
// form stores doubles; the mixed form stores each value as an FP32 hi/lo pair
// (two float planes per field).  Interface arrays (reconstructed states Wl/Wr
// and fluxes F of one sweep direction) use the same padded index, the
// interface i+1/2 being stored at cell i.
#pragma once
#include <hip/hip_runtime.h>

#include "gkm_device.h"

namespace gkm {

// Padded geometry: extents P0 (y), P1 (x), P2 (z) and strides.
struct Geo {
    int P0, P1, P2;
    long s0, s1;  // s2 = 1
    int nx, ny, nz;
};

// A box of cells/interfaces [o, o+n) in padded coordinates, flattened with the
// unit-stride axis fastest.
struct Box3 {
    int o0, o1, o2;
    int n0, n1, n2;
    __host__ __device__ long count() const { return (long)n0 * n1 * n2; }
};

__device__ __forceinline__ long box_index(const Geo& g, const Box3& b, long t) {
    int k = (int)(t % b.n2);
    long r = t / b.n2;
    int i = (int)(r % b.n1);
    int j = (int)(r / b.n1);
    return (long)(j + b.o0) * g.s0 + (long)(i + b.o1) * g.s1 + (k + b.o2);
}

// ---- state storage ---------------------------------------------------------
struct StateD {
    double* w[5];
    __device__ __forceinline__ double get(int q, long c) const { return w[q][c]; }
    __device__ __forceinline__ double cell(int q, long c) const { return w[q][c]; }
};
struct StateM {
    float* hi[5];
    float* lo[5];
    __device__ __forceinline__ double get(int q, long c) const { return (double)hi[q][c] + (double)lo[q][c]; }
    __device__ __forceinline__ float cell(int q, long c) const { return hi[q][c]; }
};

template <class T> struct Planes5 {
    T* p[5];
};
template <class T> struct Planes10 {
    T* p[10];  // Wl[0..4], Wr[0..4]
};
template <class T> struct Planes15 {
    T* p[15];  // Fx[0..4], Fy[0..4], Fz[0..4]
};

struct StepPar {
    double gam, R, invPr, mu0, T0, visc_exp, dt, dx;
    int K, visc_law, energy_recon;
};

}  // namespace gkm
