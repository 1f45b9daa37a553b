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

// Padded geometry: extents P0 (y), P1 (x), P2 (z) and strides.  Cell indices
// are 32-bit (the handle limits a plane to 2^28 cells, so byte offsets of FP64
// planes fit 31 bits): a load is then a 32-bit offset from a plane base held
// in scalar registers, one offset register serving every plane.
struct Geo {
    int P0, P1, P2;
    unsigned s0, s1;  // s2 = 1
    int nx, ny, nz;
    unsigned ncell;   // padded cells of one box; box b's planes start at b * ncell
};

// Blocks are dealt to the 8 XCDs round-robin by blockIdx.x.  The logical block
// number below gives XCD x the x-th contiguous eighth of the blocks, so that
// neighbouring rows of interfaces (or grid lines), which read each other's
// values or share cache lines, run on one XCD and share its L2.  A bijection on [0, gridDim.x).  (For a
// batch, blockIdx.y > 0 rows start on other XCDs when gridDim.x is not a
// multiple of 8; the mapping stays a bijection, only the locality is lost.)
__device__ __forceinline__ unsigned xcd_block() {
    const unsigned nb = gridDim.x, xcd = blockIdx.x & 7u, q = nb >> 3, rem = nb & 7u;
    return xcd * q + min(xcd, rem) + (blockIdx.x >> 3);
}

// Division by a run-time invariant divisor d >= 1 for dividends t < 2^31:
// q = (umulhi(t, m) + t) >> s with s = ceil(log2 d),
// m = floor(2^32 (2^s - d) / d) + 1 (round-up reciprocal, 33-bit magic).
struct FastDiv {
    unsigned d, m, s;
};

inline FastDiv make_fastdiv(unsigned d) {
    FastDiv f;
    f.d = d;
    f.s = 0;
    while ((1ull << f.s) < d) f.s++;
    f.m = (unsigned)((((unsigned long long)1 << 32) * (((unsigned long long)1 << f.s) - d)) / d + 1);
    return f;
}

__device__ __forceinline__ unsigned fdiv(unsigned t, const FastDiv& f) {
    const unsigned hi = __umulhi(t, f.m);
    return (hi + t) >> f.s;
}

// A box of cells/interfaces [o, o+n) in padded coordinates, flattened with the
// unit-stride axis fastest.
struct Box3 {
    int o0, o1, o2;
    int n0, n1, n2;
    FastDiv f1, f2;  // division by n1 and n2
    __host__ __device__ unsigned count() const { return (unsigned)n0 * n1 * n2; }
};

inline Box3 make_box(int o0, int o1, int o2, int n0, int n1, int n2) {
    Box3 b;
    b.o0 = o0; b.o1 = o1; b.o2 = o2;
    b.n0 = n0; b.n1 = n1; b.n2 = n2;
    b.f1 = make_fastdiv((unsigned)n1);
    b.f2 = make_fastdiv((unsigned)n2);
    return b;
}

__device__ __forceinline__ unsigned box_index(const Geo& g, const Box3& b, unsigned t) {
    const unsigned r = fdiv(t, b.f2);
    const unsigned k = t - r * (unsigned)b.n2;
    const unsigned j = fdiv(r, b.f1);
    const unsigned i = r - j * (unsigned)b.n1;
    return (j + b.o0) * g.s0 + (i + b.o1) * g.s1 + (k + b.o2);
}

// element c of a plane, addressed by a 32-bit byte offset
template <class T> __device__ __forceinline__ T ld(const T* base, unsigned c) {
    return *(const T*)((const char*)base + c * (unsigned)sizeof(T));
}
template <class T> __device__ __forceinline__ void put(T* base, unsigned c, T v) {
    *(T*)((char*)base + c * (unsigned)sizeof(T)) = v;
}

// ---- state storage ---------------------------------------------------------
struct StateD {
    double* w[5];
    __device__ __forceinline__ double get(int q, unsigned c) const { return ld(w[q], c); }
    __device__ __forceinline__ double cell(int q, unsigned c) const { return ld(w[q], c); }
};
struct StateM {
    float* hi[5];
    float* lo[5];
    __device__ __forceinline__ double get(int q, unsigned c) const {
        return (double)ld(hi[q], c) + (double)ld(lo[q], c);
    }
    __device__ __forceinline__ float cell(int q, unsigned c) const { return ld(hi[q], c); }
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

// uniform values the kernels use in their arithmetic type, derived once per
// box: the kernels then do no per-thread divisions or FP64 -> FP32 conversions
// of them (each is the same bits as the expression computed in the kernel)
enum : int { kUGam, kUInvPr, kUK, kUKp3, kUIkp3, kUIdx, kUDt, kUDx, kUCR, kUIR, kUMu0, kUIT0, kUVexp, kNU };

struct StepPar {
    double gam, R, invPr, mu0, T0, visc_exp, dt, dx;
    int K, visc_law, energy_recon;
    double ud[kNU];  // gam, 1/Pr, K, K+3, 1/(K+3), 1/dx, dt, dx, R/(gam-1), 1/R, mu0, 1/T0, visc_exp
    float uf[kNU];   // the same rounded to float
};

// fills ud / uf from the fields above
__host__ __device__ inline void derive_uniforms(StepPar& p) {
    p.ud[kUGam] = p.gam;
    p.ud[kUInvPr] = p.invPr;
    p.ud[kUK] = (double)p.K;
    p.ud[kUKp3] = (double)(p.K + 3);
    p.ud[kUIkp3] = 1.0 / (p.K + 3);
    p.ud[kUIdx] = 1.0 / p.dx;
    p.ud[kUDt] = p.dt;
    p.ud[kUDx] = p.dx;
    p.ud[kUCR] = p.R / (p.gam - 1.0);
    p.ud[kUIR] = 1.0 / p.R;
    p.ud[kUMu0] = p.mu0;
    p.ud[kUIT0] = 1.0 / p.T0;
    p.ud[kUVexp] = p.visc_exp;
    for (int i = 0; i < kNU; i++) p.uf[i] = (float)p.ud[i];
}

template <class Real> __device__ __forceinline__ Real upar(const StepPar& p, int i) {
    if constexpr (sizeof(Real) == 8)
        return p.ud[i];
    else
        return p.uf[i];
}

}  // namespace gkm
