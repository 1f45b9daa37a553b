// gkm_device.h — device arithmetic of the GKM-DNS step for gfx950:
// WENO-SYMOO / WENO-JS interface reconstruction and the BGK-NS interface flux,
// templated on the arithmetic type (double for the drop-in form, float for the
// mixed form).
//
// The flux follows weno/src/GKMflux.cpp:11-215 (the same moments, slope solves
// and time-integral coefficients) but is reorganised:
//  * each Maxwellian's moments are taken in its tangentially co-moving frame,
//    where they collapse to a few products of the normal-axis table (see the
//    block comment above Maxw); the reference's MCal instead expands every
//    <u^k v^l w^m psi (a.psi)> in the lab frame (~85 flops, 35 calls per
//    interface);
//  * tau and the gamma / p time-integral coefficients are formed first, then
//    each Maxwellian (left, right, equilibrium) is visited once and its
//    contributions are accumulated into four 5-vectors, so its tables and
//    slopes die as soon as its block ends (the reference keeps ~400 doubles
//    live, GPUupdate.cu:499-520);
//  * only the normal axis carries half-space moments (the reference fills them
//    for all three axes, but MCal reads them only from row 0);
//  * a side's time-slope, time-source and flux moments are summed over their
//    four slope vectors component by component before the table is applied
//    (g_coef / g_mom / g_pair), and adjacent moment orders share a component.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "gkm_erfcx.h"

namespace gkm {

template <class T> struct M;
// FP64: the ROCm device library for exp and pow; reciprocals from v_rcp_f64
// refined by two Newton steps (within an ulp of 1/x for the normal-range
// values that occur here; the IEEE division sequence is about twice as long).
template <> struct M<double> {
    __device__ static __forceinline__ double exp_(double x) { return ::exp(x); }
    __device__ static __forceinline__ double sqrt_(double x) { return ::sqrt(x); }
    __device__ static __forceinline__ double pow_(double x, double y) { return ::pow(x, y); }
    __device__ static __forceinline__ double rcp(double x) {
        double r = __builtin_amdgcn_rcp(x);
        double e = fma(-x, r, 1.0);
        r = fma(r, e, r);
        e = fma(-x, r, 1.0);
        return fma(r, e, r);
    }
    // 1/sqrt(x), x > 0 normal: v_rsq_f64 refined by two Newton steps
    // r <- r (3 - x r^2)/2, each squaring the relative error
    __device__ static __forceinline__ double rsqrt(double x) {
        double r = __builtin_amdgcn_rsq(x);
        const double hx = 0.5 * x;
        r = r * fma(-hx * r, r, 1.5);
        return r * fma(-hx * r, r, 1.5);
    }
};
// FP32 (mixed form): hardware reciprocal, exponential and square root
// (v_rcp_f32, v_exp_f32, v_sqrt_f32, v_rsq_f32, about 1 ulp); erfc from the
// erfcx polynomial below.
template <> struct M<float> {
    __device__ static __forceinline__ float exp_(float x) { return __expf(x); }
    __device__ static __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
    __device__ static __forceinline__ float pow_(float x, float y) { return ::powf(x, y); }
    __device__ static __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    __device__ static __forceinline__ float rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
};

// ---------------------------------------------------------------------------
// WENO (one scalar line of six cells -> values left and right of the centre
// interface).  SYMOO: weno/src/update.cpp:507-557, GPUupdate.cu:917-977.
// JS: update.cpp:1421-1468.
// ---------------------------------------------------------------------------
// Floating-point contraction in the WENO code is statement-local: a multiply
// is fused into an add only within one expression, never across statements.
// Fusing across statements (the compiler's default for HIP) makes the choice
// depend on the code around an inlined call -- how often a product is used --
// so the same indicator formed in two kernels could round differently.  With
// it statement-local, an indicator or weight has the same bits in every
// kernel that forms it, which the carried weights and the Z sweep's lane
// exchange rely on.
#pragma clang fp contract(on)
// Both schemes are written for a thread that walks consecutive interfaces
// along a line.  The stencils of the right value at interface i+1/2 are those of the
// left value at i+3/2 (the cells i-1..i+3 shifted by one), so their
// smoothness weights are carried instead of recomputed:
//   JS:    1/(eps+IS)^2 of IS2n, IS1n, IS0n -> IS0p, IS1p, IS2p of the next interface
//   SYMOO: IS and 1/(eps+IS) of the same three (IS3p, the fourth left stencil,
//          is new cells and is computed each time)
// js_pset / symoo_pset give the left-side weights of the first interface.
template <class T> struct Carry {
    T r[3];   // JS: squared reciprocals; SYMOO: reciprocals
    T is[3];  // SYMOO: smoothness indicators (for the max that replaces IS3)
};

template <class T> __device__ __forceinline__ T is_sq(T a, T b) {
    return T(13) / T(12) * a * a + T(0.25) * b * b;
}

// (written as the previous interface's right-side indicators are, term for
// term, so a carried weight and a freshly computed one are the same bits and
// the results do not depend on where a thread's run of interfaces starts)
template <class T> __device__ __forceinline__ void js_pset(T LLL, T LL, T L, T R, T RR, T eps, Carry<T>& c) {
    T r;
    r = M<T>::rcp(eps + is_sq(LLL - 2 * LL + L, LLL - 4 * LL + 3 * L)); c.r[0] = r * r;
    r = M<T>::rcp(eps + is_sq(LL - 2 * L + R, LL - R));               c.r[1] = r * r;
    r = M<T>::rcp(eps + is_sq(RR - 2 * R + L, RR - 4 * R + 3 * L));   c.r[2] = r * r;
}

// right-side weights of interface i+1/2, stored as the left-side weights of
// interface i+3/2 want them (IS2n, IS1n, IS0n -> IS0p, IS1p, IS2p)
template <class T> __device__ __forceinline__ void js_right(T LL, T L, T R, T RR, T RRR, T eps, Carry<T>& c) {
    T r0, r1, r2;
    r0 = M<T>::rcp(eps + is_sq(RRR - 2 * RR + R, RRR - 4 * RR + 3 * R)); r0 = r0 * r0;
    r1 = M<T>::rcp(eps + is_sq(L - 2 * R + RR, L - RR));                 r1 = r1 * r1;
    r2 = M<T>::rcp(eps + is_sq(LL - 2 * L + R, LL - 4 * L + 3 * R));     r2 = r2 * r2;
    c.r[0] = r2;
    c.r[1] = r1;
    c.r[2] = r0;
}

// both values of interface i+1/2 from its left-side weights (lw) and right-side ones (rw, as js_right stores them)
template <class T>
__device__ __forceinline__ void weno_js_lr(T LLL, T LL, T L, T R, T RR, T RRR, const Carry<T>& lw, const Carry<T>& rw,
                                           T& wl, T& wr) {
    const T c13 = T(1) / T(3), c56 = T(5) / T(6), c16 = T(1) / T(6), c76 = T(7) / T(6), c116 = T(11) / T(6);
    const T P0p = LLL * c13 - LL * c76 + L * c116;
    const T P0n = R * c116 - RR * c76 + RRR * c13;
    const T P1p = -LL * c16 + L * c56 + R * c13;
    const T P1n = L * c13 + R * c56 - RR * c16;
    T A0 = T(0.1) * lw.r[0], A1 = T(0.6) * lw.r[1], A2 = T(0.3) * lw.r[2];
    wl = (A0 * P0p + A1 * P1p + A2 * P1n) * M<T>::rcp(A0 + A1 + A2);
    A0 = T(0.1) * rw.r[2]; A1 = T(0.6) * rw.r[1]; A2 = T(0.3) * rw.r[0];
    wr = (A0 * P0n + A1 * P1n + A2 * P1p) * M<T>::rcp(A0 + A1 + A2);
}

template <class T>
__device__ __forceinline__ void weno_js_carry(T LLL, T LL, T L, T R, T RR, T RRR, T eps, Carry<T>& c, T& wl,
                                              T& wr) {
    Carry<T> rw;
    js_right(LL, L, R, RR, RRR, eps, rw);
    weno_js_lr(LLL, LL, L, R, RR, RRR, c, rw, wl, wr);
    c = rw;
}

template <class T> __device__ __forceinline__ void symoo_pset(T LLL, T LL, T L, T R, T RR, T eps, Carry<T>& c) {
    c.is[0] = is_sq(L - 2 * LL + LLL, -3 * L + 4 * LL - LLL);
    c.is[1] = is_sq(R - 2 * L + LL, LL - R);
    c.is[2] = is_sq(RR - 2 * R + L, RR - 4 * R + 3 * L);
#pragma unroll
    for (int k = 0; k < 3; k++) c.r[k] = M<T>::rcp(eps + c.is[k]);
}

// right-side indicators and weights of interface i+1/2, stored as interface
// i+3/2's left side wants them (IS2n, IS1n, IS0n -> IS0p, IS1p, IS2p)
template <class T> __device__ __forceinline__ void symoo_right(T LL, T L, T R, T RR, T RRR, T eps, Carry<T>& c) {
    const T IS0n = is_sq(RRR - 2 * RR + R, RRR - 4 * RR + 3 * R);
    const T IS1n = is_sq(RR - 2 * R + L, L - RR);
    const T IS2n = is_sq(R - 2 * L + LL, -3 * R + 4 * L - LL);
    c.is[0] = IS2n; c.r[0] = M<T>::rcp(eps + IS2n);
    c.is[1] = IS1n; c.r[1] = M<T>::rcp(eps + IS1n);
    c.is[2] = IS0n; c.r[2] = M<T>::rcp(eps + IS0n);
}

// both values of interface i+1/2 from its left-side (lw) and right-side (rw) indicators and weights; the
// fourth stencil of each side is new cells and formed here
template <class T>
__device__ __forceinline__ void weno_symoo_lr(T LLL, T LL, T L, T R, T RR, T RRR, T eps, const Carry<T>& lw,
                                              const Carry<T>& rw, T& wl, T& wr) {
    const T c13 = T(1) / T(3), c56 = T(5) / T(6), c16 = T(1) / T(6), c76 = T(7) / T(6), c116 = T(11) / T(6);
    const T P0p = LLL * c13 - LL * c76 + L * c116;
    const T P0n = RRR * c13 - RR * c76 + R * c116;
    const T P1p = -LL * c16 + L * c56 + R * c13;
    const T P1n = -RR * c16 + R * c56 + L * c13;
    const T P2p = L * c13 + R * c56 - RR * c16;
    const T P2n = R * c13 + L * c56 - LL * c16;
    const T P3p = R * c116 - RR * c76 + RRR * c13;
    const T P3n = L * c116 - LL * c76 + LLL * c13;
    const T g0 = T(1) / T(20), g1 = T(9) / T(20);
    // left value: IS0p..IS2p carried, IS3p from the new stencil, replaced by the max
    T IS3p = is_sq(R - 2 * RR + RRR, -5 * R + 8 * RR - 3 * RRR);
    IS3p = fmax(fmax(fmax(lw.is[0], lw.is[1]), lw.is[2]), IS3p);
    T A0 = g0 * lw.r[0], A1 = g1 * lw.r[1], A2 = g1 * lw.r[2], A3 = g0 * M<T>::rcp(eps + IS3p);
    wl = (A0 * P0p + A1 * P1p + A2 * P2p + A3 * P3p) * M<T>::rcp(A0 + A1 + A2 + A3);
    T IS3n = is_sq(L - 2 * LL + LLL, -5 * L + 8 * LL - 3 * LLL);
    IS3n = fmax(fmax(fmax(rw.is[2], rw.is[1]), rw.is[0]), IS3n);
    A0 = g0 * rw.r[2]; A1 = g1 * rw.r[1]; A2 = g1 * rw.r[0]; A3 = g0 * M<T>::rcp(eps + IS3n);
    wr = (A0 * P0n + A1 * P1n + A2 * P2n + A3 * P3n) * M<T>::rcp(A0 + A1 + A2 + A3);
}

template <class T>
__device__ __forceinline__ void weno_symoo_carry(T LLL, T LL, T L, T R, T RR, T RRR, T eps, Carry<T>& c, T& wl,
                                                 T& wr) {
    Carry<T> rw;
    symoo_right(LL, L, R, RR, RRR, eps, rw);
    weno_symoo_lr(LLL, LL, L, R, RR, RRR, eps, c, rw, wl, wr);
    c = rw;
}

// Single interface: the first interface's left-side weights, then the
// carried form.  Every path (the strided windows, the Z sweep's lanes, the
// persistent kernel, the device arithmetic checks) therefore forms each
// indicator and weight with the same expression, whichever interface supplied
// it, and gets the same bits.
template <class T>
__device__ __forceinline__ void weno_symoo(T LLL, T LL, T L, T R, T RR, T RRR, T eps, T& wl, T& wr) {
    Carry<T> c;
    symoo_pset(LLL, LL, L, R, RR, eps, c);
    weno_symoo_carry(LLL, LL, L, R, RR, RRR, eps, c, wl, wr);
}

template <class T>
__device__ __forceinline__ void weno_js(T LLL, T LL, T L, T R, T RR, T RRR, T eps, T& wl, T& wr) {
    Carry<T> c;
    js_pset(LLL, LL, L, R, RR, eps, c);
    weno_js_carry(LLL, LL, L, R, RR, RRR, eps, c, wl, wr);
}

template <int RECON, class T>
__device__ __forceinline__ void weno_line_carry(const T c[6], T eps, Carry<T>& k, T& wl, T& wr) {
    if (RECON == 1)
        weno_js_carry(c[0], c[1], c[2], c[3], c[4], c[5], eps, k, wl, wr);
    else
        weno_symoo_carry(c[0], c[1], c[2], c[3], c[4], c[5], eps, k, wl, wr);
}

template <int RECON, class T> __device__ __forceinline__ void weno_pset(const T c[5], T eps, Carry<T>& k) {
    if (RECON == 1)
        js_pset(c[0], c[1], c[2], c[3], c[4], eps, k);
    else
        symoo_pset(c[0], c[1], c[2], c[3], c[4], eps, k);
}

// the right-side weights of a line's interface, and both values from the left- and right-side ones
template <int RECON, class T> __device__ __forceinline__ void weno_right(const T c[6], T eps, Carry<T>& k) {
    if (RECON == 1)
        js_right(c[1], c[2], c[3], c[4], c[5], eps, k);
    else
        symoo_right(c[1], c[2], c[3], c[4], c[5], eps, k);
}
template <int RECON, class T>
__device__ __forceinline__ void weno_line_lr(const T c[6], T eps, const Carry<T>& lw, const Carry<T>& rw, T& wl,
                                             T& wr) {
    if (RECON == 1)
        weno_js_lr(c[0], c[1], c[2], c[3], c[4], c[5], lw, rw, wl, wr);
    else
        weno_symoo_lr(c[0], c[1], c[2], c[3], c[4], c[5], eps, lw, rw, wl, wr);
}

template <int RECON, class T>
__device__ __forceinline__ void weno_line(const T c[6], T eps, T& wl, T& wr) {
    if (RECON == 1)
        weno_js(c[0], c[1], c[2], c[3], c[4], c[5], eps, wl, wr);
    else
        weno_symoo(c[0], c[1], c[2], c[3], c[4], c[5], eps, wl, wr);
}

#pragma clang fp contract(fast)

// ---------------------------------------------------------------------------
// BGK-NS flux
// ---------------------------------------------------------------------------
template <class T> struct Gas {
    T gam;     // ratio of specific heats
    T invPr;   // 1/Pr applied on a[4] in the slope solver (GKMflux.cpp:289)
    T kp3;     // K + 3
    T ikp3;    // 1/(K + 3)
    T K;       // internal degrees of freedom
};

// ---------------------------------------------------------------------------
// Moments in the tangentially co-moving frame.
//
// For a Maxwellian with velocity (U, V, W) and lambda, h = 1/(2 lambda), use
// psi' = (1, u, v', w', (u^2 + v'^2 + w'^2 + xi^2)/2) with v' = v - V,
// w' = w - W.  The tangential moments are then central Gaussian moments
// (<v'> = <v'^3> = 0, <v'^2> = h, <v'^4> = 3h^2) and xi^2 ~ h chi^2_K, so
// every moment the flux needs, <u^k v'^l w'^m psi' (a' . psi')> with l, m in
// {0, 1}, reduces to a few products of the normal-axis table N_j = <u^j>
// (full or half space) and the sums
//   Q_k  = <u^k S>    = (N_{k+2} + (K+2) h N_k) / 2,     S = psi'_4
//   R_k  = <u^k S^2>  = (N_{k+4} + 2 (K+2) h N_{k+2} + (K+2)(K+4) h^2 N_k) / 4
//   Q'_k = <u^k v'^2 S> / h = (N_{k+2} + (K+4) h N_k) / 2.
// Vectors move between psi and psi' by exact linear maps: moments / sources
// with Tinv (to the frame) and T (back to the lab frame), slopes with T^T.
// The slope solve of GKMflux.cpp:280-296 (1/Pr on a[4] included) commutes
// with that change of frame, so slopes are solved directly in the frame with
// velocity (U, 0, 0).  This is the same flux as GKMflux.cpp:11-215; the
// reference's MCal evaluates the lab-frame products term by term instead.
// ---------------------------------------------------------------------------
template <class T> struct Maxw {
    T den, U, V, W, lam, h;
    T ri;  // 1/den
    T P;   // pressure; h = 1/(2 lam) = P/den
};

template <class T> struct NTab {
    T N[7];
    T Q[4], R[3], Qp[2];
};

template <class T>
__device__ __forceinline__ void maxw_from_W(const Gas<T>& g, const T W[5], Maxw<T>& s) {
    s.den = W[0];
    const T ri = M<T>::rcp(W[0]);
    s.ri = ri;
    s.U = W[1] * ri;
    s.V = W[2] * ri;
    s.W = W[3] * ri;
    const T P = (s.den * (g.gam - 1)) * (W[4] * ri - T(0.5) * (s.U * s.U + s.V * s.V + s.W * s.W));
    s.P = P;
    s.lam = T(0.5) * s.den * M<T>::rcp(P);
    s.h = P * ri;
}

template <class T> __device__ __forceinline__ void ntab_sums(const Gas<T>& g, T h, NTab<T>& t) {
    const T c2 = (g.K + 2) * h, c4 = (g.K + 4) * h, c24 = (g.K + 2) * (g.K + 4) * h * h;
#pragma unroll
    for (int k = 0; k < 4; k++) t.Q[k] = T(0.5) * (t.N[k + 2] + c2 * t.N[k]);
#pragma unroll
    for (int k = 0; k < 3; k++) t.R[k] = T(0.25) * (t.N[k + 4] + 2 * c2 * t.N[k + 2] + c24 * t.N[k]);
#pragma unroll
    for (int k = 0; k < 2; k++) t.Qp[k] = T(0.5) * (t.N[k + 2] + c4 * t.N[k]);
}

// full-space normal-axis table <u^j>
template <class T> __device__ __forceinline__ void full_ntab(const Gas<T>& g, T U, T h, NTab<T>& t) {
    t.N[0] = T(1);
    t.N[1] = U;
#pragma unroll
    for (int i = 2; i < 7; i++) t.N[i] = U * t.N[i - 1] + (T(i - 1) * h) * t.N[i - 2];
    ntab_sums(g, h, t);
}

// erfcx(a) = exp(a^2) erfc(a), a >= 0, FP32: the rational P(a)/Q(a) of
// gkm_erfcx.h (deg 4 / deg 5, positive coefficients, so Horner's rule has no
// cancellation), a capped at kErfcxCap where exp(-a^2) is already 0 in FP32.
__device__ __forceinline__ float erfcx_(float a) {
    const float x = fminf(a, float(kErfcxCap));
    float p = kErfcxP[kErfcxDegP], q = kErfcxQ[kErfcxDegQ];
#pragma unroll
    for (int i = kErfcxDegP - 1; i >= 0; i--) p = fmaf(p, x, kErfcxP[i]);
#pragma unroll
    for (int i = kErfcxDegQ - 1; i >= 0; i--) q = fmaf(q, x, kErfcxQ[i]);
    return p * M<float>::rcp(q);
}

// erfcx(a), a >= 0, FP64: the deg 9 / deg 10 rational of gkm_erfcx.h on
// [0, kErfcxCap64] (relative error 1.1e-15 in this evaluation; beyond the cap
// exp(-a^2) underflows in FP64), with the reciprocal of Q within an ulp.  Its
// constants are compile-time literals (scalar operands).
__device__ __forceinline__ double erfcx_(double a) {
    const double x = fmin(a, kErfcxCap64);
    double p = kErfcxP64[kErfcxDegP64], q = kErfcxQ64[kErfcxDegQ64];
#pragma unroll
    for (int i = kErfcxDegP64 - 1; i >= 0; i--) p = fma(p, x, kErfcxP64[i]);
#pragma unroll
    for (int i = kErfcxDegQ64 - 1; i >= 0; i--) q = fma(q, x, kErfcxQ64[i]);
    return p * M<double>::rcp(q);
}

// erfc(y) given e = exp(-y^2): erfcx rational times e, branch-free, in both
// forms (FP64: ~30 instructions against ~130 for the device library's erfc;
// the FP64 flux went from 1697 to 1558 VALU instructions per interface).
template <class T> __device__ __forceinline__ T erfc_gauss(T y, T e) {
    const T r = erfcx_(fabs(y)) * e;
    return y >= T(0) ? r : T(2) - r;
}

// half-space normal-axis table; sgn = +1 for u > 0, -1 for u < 0 (GKMflux.cpp:54-65)
template <class T>
__device__ __forceinline__ void half_ntab(const Gas<T>& g, T U, T lam, T h, int sgn, NTab<T>& t) {
    // sqrt(lam) = lam / sqrt(lam) and 1/sqrt(lam pi) from one reciprocal square root
    const T rs = M<T>::rsqrt(lam);
    const T sl = lam * rs;
    const T half_rsqrt_pi = T(0.28209479177387814347403972578038629);  // 1/(2 sqrt(pi))
    // erfc(y), y = -sgn sqrt(lam) U, shares the Gaussian factor exp(-lam U^2)
    // with N1: erfc(|y|) = erfcx(|y|) exp(-y^2) and erfc(-|y|) = 2 - erfc(|y|)
    const T y = -T(sgn) * sl * U;
    const T e = M<T>::exp_(-lam * U * U);
    t.N[0] = T(0.5) * erfc_gauss(y, e);
    t.N[1] = U * t.N[0] + T(sgn) * half_rsqrt_pi * e * rs;
#pragma unroll
    for (int i = 2; i < 7; i++) t.N[i] = U * t.N[i - 1] + (T(i - 1) * h) * t.N[i - 2];
    ntab_sums(g, h, t);
}

// The moments of a slope vector a (psi' basis) against one normal table:
//   u_k(a) = <u^k psi' (a . psi')>
//          = (a0 N_k + a1 N_k+1 + a4 Q_k,  a0 N_k+1 + a1 N_k+2 + a4 Q_k+1,  h N_k a2,  h N_k a3,
//             a0 Q_k + a1 Q_k+1 + a4 R_k)
//   v_k(a) = <u^k v' psi' (a . psi')> = h (a2 N_k,  a2 N_k+1,  a0 N_k + a1 N_k+1 + a4 Q'_k,  0,  a2 Q'_k)
//   w_k(a) = <u^k w' psi' (a . psi')> = h (a3 N_k,  a3 N_k+1,  0,  a0 N_k + a1 N_k+1 + a4 Q'_k,  a3 Q'_k)
// (the odd tangential Gaussian moments vanish; <v'^2> = h).
// The flux's time-slope, time-source and flux moments are all of the form
//   G_k = <u^k psi' (u ax + v' ay + w' az) . psi'> + <u^k psi' (sv . psi')>
//       = u_{k+1}(ax) + v_k(ay) + w_k(az) + u_k(sv)         (u_k, v_k, w_k above)
// with sv = V ay + W az.  Summed component by component, the four terms share
// their table entries, so G_k is five short dot products of the table with
// coefficients formed once per (ax, ay, az, sv):
//   G_k[0] = a0 N_k + a1 N_k+1 + a2 N_k+2 + a3 Q_k + a4 Q_k+1
//   G_k[1] = the same one entry up (= G_k+1[0])
//   G_k[2] = h (g2 N_k + d2 N_k+1 + ay4 Qp_k)
//   G_k[3] = h (g3 N_k + d3 N_k+1 + az4 Qp_k)
//   G_k[4] = sv0 Q_k + e Qp_k + a1 Q_k+1 + a2 Q_k+2 + a3 R_k + a4 R_k+1
// where e = h (ay2 + az3), a0 = e + sv0, a1 = ax0 + sv1, a2 = ax1, a3 = sv4,
// a4 = ax4, g2 = ay0 + sv2, d2 = ax2 + ay1, g3 = az0 + sv3, d3 = ax3 + az1.
template <class T> struct GCoef {
    T a0, a1, a2, a3, a4, s0, e, g2, d2, y4, g3, d3, z4, h;
};

template <class T>
__device__ __forceinline__ void g_coef(const T ax[5], const T ay[5], const T az[5], const T sv[5], T h,
                                       GCoef<T>& c) {
    c.h = h;
    c.e = h * (ay[2] + az[3]);
    c.s0 = sv[0];
    c.a0 = c.e + sv[0];
    c.a1 = ax[0] + sv[1];
    c.a2 = ax[1];
    c.a3 = sv[4];
    c.a4 = ax[4];
    c.g2 = ay[0] + sv[2];
    c.d2 = ax[2] + ay[1];
    c.y4 = ay[4];
    c.g3 = az[0] + sv[3];
    c.d3 = ax[3] + az[1];
    c.z4 = az[4];
}

// G_k alone
template <int k, class T> __device__ __forceinline__ void g_mom(const NTab<T>& t, const GCoef<T>& c, T o[5]) {
    o[0] = c.a0 * t.N[k] + c.a1 * t.N[k + 1] + c.a2 * t.N[k + 2] + c.a3 * t.Q[k] + c.a4 * t.Q[k + 1];
    o[1] = c.a0 * t.N[k + 1] + c.a1 * t.N[k + 2] + c.a2 * t.N[k + 3] + c.a3 * t.Q[k + 1] + c.a4 * t.Q[k + 2];
    o[2] = c.h * (c.g2 * t.N[k] + c.d2 * t.N[k + 1] + c.y4 * t.Qp[k]);
    o[3] = c.h * (c.g3 * t.N[k] + c.d3 * t.N[k + 1] + c.z4 * t.Qp[k]);
    o[4] = c.s0 * t.Q[k] + c.e * t.Qp[k] + c.a1 * t.Q[k + 1] + c.a2 * t.Q[k + 2] + c.a3 * t.R[k] +
           c.a4 * t.R[k + 1];
}

// G_k and G_k+1, sharing G_k[1] = G_k+1[0]
template <int k, class T>
__device__ __forceinline__ void g_pair(const NTab<T>& t, const GCoef<T>& c, T o0[5], T o1[5]) {
    g_mom<k>(t, c, o0);
    o1[0] = o0[1];
    o1[1] = c.a0 * t.N[k + 2] + c.a1 * t.N[k + 3] + c.a2 * t.N[k + 4] + c.a3 * t.Q[k + 2] + c.a4 * t.Q[k + 3];
    o1[2] = c.h * (c.g2 * t.N[k + 1] + c.d2 * t.N[k + 2] + c.y4 * t.Qp[k + 1]);
    o1[3] = c.h * (c.g3 * t.N[k + 1] + c.d3 * t.N[k + 2] + c.z4 * t.Qp[k + 1]);
    o1[4] = c.s0 * t.Q[k + 1] + c.e * t.Qp[k + 1] + c.a1 * t.Q[k + 2] + c.a2 * t.Q[k + 3] + c.a3 * t.R[k + 1] +
            c.a4 * t.R[k + 2];
}

// u_k(a) = <u^k psi' (a . psi')> alone and with u_k+1(a), sharing u_k[1] = u_k+1[0]
template <int k, class T> __device__ __forceinline__ void u_mom(const NTab<T>& t, T h, const T a[5], T o[5]) {
    o[0] = a[0] * t.N[k] + a[1] * t.N[k + 1] + a[4] * t.Q[k];
    o[1] = a[0] * t.N[k + 1] + a[1] * t.N[k + 2] + a[4] * t.Q[k + 1];
    const T hn = h * t.N[k];
    o[2] = hn * a[2];
    o[3] = hn * a[3];
    o[4] = a[0] * t.Q[k] + a[1] * t.Q[k + 1] + a[4] * t.R[k];
}
template <int k, class T>
__device__ __forceinline__ void u_pair(const NTab<T>& t, T h, const T a[5], T o0[5], T o1[5]) {
    u_mom<k>(t, h, a, o0);
    o1[0] = o0[1];
    o1[1] = a[0] * t.N[k + 2] + a[1] * t.N[k + 3] + a[4] * t.Q[k + 2];
    const T hn = h * t.N[k + 1];
    o1[2] = hn * a[2];
    o1[3] = hn * a[3];
    o1[4] = a[0] * t.Q[k + 1] + a[1] * t.Q[k + 2] + a[4] * t.R[k + 1];
}

// moment / source vector: lab psi basis -> co-moving frame (Tinv)
template <class T> __device__ __forceinline__ void to_frame(const T b[5], T V, T W, T o[5]) {
    o[0] = b[0];
    o[1] = b[1];
    o[2] = b[2] - V * b[0];
    o[3] = b[3] - W * b[0];
    o[4] = b[4] - V * b[2] - W * b[3] + T(0.5) * (V * V + W * W) * b[0];
}

// moment vector: co-moving frame -> lab psi basis (T), added to acc (ADD) or
// stored in it (the first contribution: IEEE arithmetic cannot drop a 0 + x)
template <bool ADD, class T> __device__ __forceinline__ void to_lab(T acc[5], const T m[5], T V, T W) {
    const T m2 = m[2] + V * m[0], m3 = m[3] + W * m[0];
    const T m4 = m[4] + V * m[2] + W * m[3] + T(0.5) * (V * V + W * W) * m[0];
    acc[0] = ADD ? acc[0] + m[0] : m[0];
    acc[1] = ADD ? acc[1] + m[1] : m[1];
    acc[2] = ADD ? acc[2] + m2 : m2;
    acc[3] = ADD ? acc[3] + m3 : m3;
    acc[4] = ADD ? acc[4] + m4 : m4;
}
template <class T> __device__ __forceinline__ void add_to_lab(T acc[5], const T m[5], T V, T W) {
    to_lab<true>(acc, m, V, W);
}

// Slope solve in the co-moving frame (velocity (U, 0, 0)); b, a in psi'
// (GKMflux.cpp:280-296):
//   R2 = b1 - U b0,  a4 = (1/Pr) (2 b4 - b0 q - 2 U R2) 4 lam^2 / (K+3),
//   a3 = 2 lam b3,  a2 = 2 lam b2,  a1 = 2 lam R2 - U a4,  a0 = b0 - U a1 - a4 q / 2
// with q = U^2 + (K+3)/(2 lam).  The Maxwellian's factors are formed once
// (Slv); the solve is linear, so a scale s of the right-hand side is taken
// into its two factors and b0 (solve_scaled) instead of into each b.
template <class T> struct Slv {
    T U, twoU, q, hq, k4, l2;  // U, 2U, q, q/2, 4 lam^2/(Pr (K+3)), 2 lam
};

template <class T>
__device__ __forceinline__ void slv_make(const Gas<T>& g, T U, T lam, T h, Slv<T>& S) {
    S.U = U;
    S.twoU = 2 * U;
    S.q = U * U + g.kp3 * h;  // U^2 + (K+3)/(2 lam)
    S.hq = T(0.5) * S.q;
    S.l2 = 2 * lam;
    S.k4 = (g.invPr * g.ikp3) * (S.l2 * S.l2);
}

// a = solve of s b, given k4s = s k4 and l2s = s l2
template <class T>
__device__ __forceinline__ void solve_scaled(const Slv<T>& S, T s, T k4s, T l2s, const T b[5], T a[5]) {
    const T R2 = b[1] - S.U * b[0];
    a[4] = k4s * (2 * b[4] - b[0] * S.q - S.twoU * R2);
    a[3] = l2s * b[3];
    a[2] = l2s * b[2];
    a[1] = l2s * R2 - S.U * a[4];
    a[0] = s * b[0] - S.U * a[1] - S.hq * a[4];
}

// a = solve of b
template <class T> __device__ __forceinline__ void solve_frame(const Slv<T>& S, const T b[5], T a[5]) {
    const T R2 = b[1] - S.U * b[0];
    a[4] = S.k4 * (2 * b[4] - b[0] * S.q - S.twoU * R2);
    a[3] = S.l2 * b[3];
    a[2] = S.l2 * b[2];
    a[1] = S.l2 * R2 - S.U * a[4];
    a[0] = b[0] - S.U * a[1] - S.hq * a[4];
}

// Contribution of one side (left: normal table u > 0, right: u < 0).
// Accumulates, in the lab frame: by, bz (sources of the equilibrium state's
// tangential slopes), B (source of its time slope) and F (the flux).
template <bool FIRST, class T>
__device__ __forceinline__ void side_block(const Gas<T>& g, const Maxw<T>& s, const NTab<T>& nh, const T DWx[5],
                                           const T DWy[5], const T DWz[5], T sx, T st, T c45, T c5, T p3, T p4,
                                           T p5, T by[5], T bz[5], T B[5], T F[5]) {
    const T h = s.h;
    const T rx = sx * s.ri, rt = st * s.ri;  // slope scale / rho
    Slv<T> S;
    slv_make(g, s.U, s.lam, h, S);
    T b[5], bp[5], ax[5], ay[5], az[5], sv[5];
    to_frame(DWx, s.V, s.W, bp);
    solve_scaled(S, rx, S.k4 * rx, S.l2 * rx, bp, ax);
    const T k4t = S.k4 * rt, l2t = S.l2 * rt;
    to_frame(DWy, s.V, s.W, bp);
    solve_scaled(S, rt, k4t, l2t, bp, ay);
    to_frame(DWz, s.V, s.W, bp);
    solve_scaled(S, rt, k4t, l2t, bp, az);
#pragma unroll
    for (int q = 0; q < 5; q++) sv[q] = s.V * ay[q] + s.W * az[q];
    GCoef<T> gc;
    g_coef(ax, ay, az, sv, h, gc);
    // time slope: A = -(<u psi ax> + <v psi ay> + <w psi az>) over the full
    // space = -solve(G_0 of the full table); nA = -A
    T nA[5];
    {
        NTab<T> nf;
        full_ntab(g, s.U, h, nf);
        g_mom<0>(nf, gc, b);
        solve_frame(S, b, nA);
    }
    T m[5];
    // equilibrium tangential-slope sources <psi ay>_half, <psi az>_half
    u_mom<0>(nh, h, ay, m);
    to_lab<!FIRST>(by, m, s.V, s.W);
    u_mom<0>(nh, h, az, m);
    to_lab<!FIRST>(bz, m, s.V, s.W);
    const T d = s.den;
    // time-slope source: (gm4+gm5) rho (<u psi ax> + <v psi ay> + <w psi az>) + gm5 rho <psi A>
    //   = rho (gm4+gm5) G_0 + rho gm5 u_0(A)
    // flux: p3 rho <u psi> + p4 rho (<u^2 psi ax> + <uv psi ay> + <uw psi az>) + p5 rho <u psi A>
    //   = rho p4 G_1 + rho p5 u_1(A) + rho p3 (N_1, N_2, 0, 0, Q_1)
    T G0[5], G1[5], uA0[5], uA1[5];
    g_pair<0>(nh, gc, G0, G1);
    u_pair<0>(nh, h, nA, uA0, uA1);
    const T dc45 = d * c45, dc5 = d * c5, dp3 = d * p3, dp4 = d * p4, dp5 = d * p5;
#pragma unroll
    for (int q = 0; q < 5; q++) m[q] = dc45 * G0[q] - dc5 * uA0[q];
    add_to_lab(B, m, s.V, s.W);
#pragma unroll
    for (int q = 0; q < 5; q++) m[q] = dp4 * G1[q] - dp5 * uA1[q];
    m[0] += dp3 * nh.N[1];
    m[1] += dp3 * nh.N[2];
    m[4] += dp3 * nh.Q[1];
    to_lab<!FIRST>(F, m, s.V, s.W);
}

// Time-integral coefficients of the BGK solution (GKMflux.cpp:155-157,
// 186-190) at x = dt/tau: gm1, gm2, gm4 of the time slope, and P, Q, p0..p2
// (gm3 = -gm1, gm5 = tau gm3, p3 = P, p4 = -Q - tau P, p5 = -tau P follow).
//
// FP64: the reference's expressions, in its order.
//
// FP32: the same expressions lose every digit as x -> 0: gm0 = dt - tau(1 -
// e^-x) = tau (x^2/2 - ...) is a difference of two O(dt) terms each rounded to
// eps dt/x, so gm0 carries a relative error ~ 2 eps/x^2 (3e-4 at x = 0.16, the
// reference's CFL 0.02 on a 256^3 box; 80 % at x = 0.01), and gm1, gm2, gm4,
// gm5 divide by it.  Below x = 1/2 they are formed instead from the
// exponential-integrator functions phi_k(-x) = sum_n (-x)^n / (n+k)!, which
// have no cancellation (8 terms each: the first dropped term is < 3e-8 of
// the sum at x = 1/2):
//   gm1 = -phi1/(dt phi2)            gm2 = x (2 phi3 - phi2)/phi2
//   gm4 = (phi2 - phi1)/phi2         P = dt phi1       Q = dt^2 (phi1 - phi2)
//   p0 = dt x phi2   p1 = dt^2 x phi3   p2 = dt^2 x ((2 + x) phi3 - 1/2)
// which are the reference's expressions rewritten with e^-x = 1 - x phi1,
// phi1 = 1 - x phi2, phi2 = 1/2 - x phi3 (exact identities).  Worst relative
// error of the ten coefficients in FP32: 2.8e-7 below x = 1/2, 1.9e-6 just
// above it (the reference form).
template <class T> struct TimeCoef {
    T gm1, gm2, gm4, P, Q, p0, p1, p2;
};

__device__ __forceinline__ void time_coef(double dt, double tau, TimeCoef<double>& k) {
    const double ex = M<double>::exp_(-dt * M<double>::rcp(tau));
    const double om = 1.0 - ex;
    const double gm0 = dt - tau * om;
    const double igm0 = M<double>::rcp(gm0);
    k.gm1 = -om * igm0;
    k.gm2 = (-dt + 2 * tau * om - dt * ex) * igm0;
    k.gm4 = (dt * ex - tau * om) * igm0;
    k.P = tau * om;
    k.Q = -tau * dt * ex + tau * tau * om;
    k.p0 = dt - k.P;
    k.p1 = 0.5 * dt * dt - tau * k.p0;
    k.p2 = -tau * k.p0 + k.Q;
}

__device__ __forceinline__ void time_coef(float dt, float tau, TimeCoef<float>& k) {
    const float x = dt * M<float>::rcp(tau);
    if (x < 0.5f) {
        // Horner in x of sum_{n<8} (-x)^n/(n+k)!
        const float f1 = 1.f + x * (-1.f / 2 + x * (1.f / 6 + x * (-1.f / 24 + x * (1.f / 120 + x * (-1.f / 720 +
                         x * (1.f / 5040 + x * (-1.f / 40320)))))));
        const float f2 = 1.f / 2 + x * (-1.f / 6 + x * (1.f / 24 + x * (-1.f / 120 + x * (1.f / 720 +
                         x * (-1.f / 5040 + x * (1.f / 40320 + x * (-1.f / 362880)))))));
        const float f3 = 1.f / 6 + x * (-1.f / 24 + x * (1.f / 120 + x * (-1.f / 720 + x * (1.f / 5040 +
                         x * (-1.f / 40320 + x * (1.f / 362880 + x * (-1.f / 3628800)))))));
        const float if2 = M<float>::rcp(f2);
        const float dt2 = dt * dt;
        k.gm1 = -f1 * if2 * M<float>::rcp(dt);
        k.gm2 = x * (2.f * f3 - f2) * if2;
        k.gm4 = (f2 - f1) * if2;
        k.P = dt * f1;
        k.Q = dt2 * (f1 - f2);
        k.p0 = dt * x * f2;
        k.p1 = dt2 * x * f3;
        k.p2 = dt2 * x * ((2.f + x) * f3 - 0.5f);
        return;
    }
    const float ex = M<float>::exp_(-x);
    const float om = 1.f - ex;
    const float gm0 = dt - tau * om;
    const float igm0 = M<float>::rcp(gm0);
    k.gm1 = -om * igm0;
    k.gm2 = (-dt + 2 * tau * om - dt * ex) * igm0;
    k.gm4 = (dt * ex - tau * om) * igm0;
    k.P = tau * om;
    k.Q = -tau * dt * ex + tau * tau * om;
    k.p0 = dt - k.P;
    k.p1 = 0.5f * dt * dt - tau * k.p0;
    k.p2 = -tau * k.p0 + k.Q;
}

// BGK-NS interface flux, time-integrated over dt (GKMflux.cpp:11-215).
// All vectors are in the rotated slot order (slot 1 = normal momentum).
// tau0 = mu/p of the left cell.  slopes(side, Dx, Dy, Dz) yields the slopes
// left (side 0) or right (side 1) of the interface along the normal (x) and
// the two tangential (y, z) directions, divided by sx (normal) and st
// (tangential): the scale factors are applied together with 1/rho.
template <class T, class LoadSlopes>
__device__ __forceinline__ void bgk_flux(const Gas<T>& g, const T WL[5], const T WR[5], const T Wl[5],
                                         const T Wr[5], LoadSlopes&& slopes, T sx, T st, T tau0, T dt, T idx,
                                         T F[5]) {
    Maxw<T> l, r;
    maxw_from_W(g, Wl, l);
    maxw_from_W(g, Wr, r);
    // collision time with the pressure-jump blending (GKMflux.cpp:149-150);
    // rho/lam = 2p, and the factor 2 cancels
    const T tau = tau0 + dt * fabs(l.P - r.P) * M<T>::rcp(fabs(l.P + r.P));
    TimeCoef<T> tc;
    time_coef(dt, tau, tc);
    const T gm1 = tc.gm1, gm2 = tc.gm2, gm4 = tc.gm4;
    const T gm3 = -gm1;
    const T gm5 = tau * gm3;
    const T P = tc.P, Q = tc.Q;
    const T p0 = tc.p0, p1 = tc.p1, p2 = tc.p2;
    const T p3 = P;
    const T p4 = -Q - tau * P;
    const T p5 = -tau * P;

    // half-space normal tables of the two sides and the equilibrium state We
    NTab<T> hl, hr;
    half_ntab(g, l.U, l.lam, l.h, +1, hl);
    half_ntab(g, r.U, r.lam, r.h, -1, hr);
    // We = rho_l <psi>_+ + rho_r <psi>_-, each side's moments rho (N_0, N_1, 0, 0, Q_0)
    T We[5];
    {
        T m[5] = {l.den * hl.N[0], l.den * hl.N[1], T(0), T(0), l.den * hl.Q[0]};
        const T vl = T(0.5) * (l.V * l.V + l.W * l.W);
        We[0] = m[0];
        We[1] = m[1];
        We[2] = l.V * m[0];
        We[3] = l.W * m[0];
        We[4] = m[4] + vl * m[0];
        m[0] = r.den * hr.N[0];
        m[1] = r.den * hr.N[1];
        m[4] = r.den * hr.Q[0];
        const T vr = T(0.5) * (r.V * r.V + r.W * r.W);
        We[0] += m[0];
        We[1] += m[1];
        We[2] += r.V * m[0];
        We[3] += r.W * m[0];
        We[4] += m[4] + vr * m[0];
    }
    Maxw<T> e;
    maxw_from_W(g, We, e);

    T by[5], bz[5], B[5], Fa[5];
#pragma unroll
    for (int q = 0; q < 5; q++) B[q] = gm3 * We[q];
    const T c45 = gm4 + gm5;
    {
        T Dx[5], Dy[5], Dz[5];
        slopes(0, Dx, Dy, Dz);  // left side
        side_block<true>(g, l, hl, Dx, Dy, Dz, sx, st, c45, gm5, p3, p4, p5, by, bz, B, Fa);
    }
    {
        T Dx[5], Dy[5], Dz[5];
        slopes(1, Dx, Dy, Dz);  // right side
        side_block<false>(g, r, hr, Dx, Dy, Dz, sx, st, c45, gm5, p3, p4, p5, by, bz, B, Fa);
    }

    // equilibrium state, in its own co-moving frame
    const T he = e.h;
    T axl[5], axr[5], ay[5], az[5], A[5], b[5], bp[5], sv[5];
    Slv<T> S;
    slv_make(g, e.U, e.lam, he, S);
    const T s2 = T(2) * e.ri * idx;  // normal slopes over half a cell, / rho
    const T k4x = S.k4 * s2, l2x = S.l2 * s2;
#pragma unroll
    for (int q = 0; q < 5; q++) b[q] = We[q] - WL[q];
    to_frame(b, e.V, e.W, bp);
    solve_scaled(S, s2, k4x, l2x, bp, axl);
#pragma unroll
    for (int q = 0; q < 5; q++) b[q] = WR[q] - We[q];
    to_frame(b, e.V, e.W, bp);
    solve_scaled(S, s2, k4x, l2x, bp, axr);
    to_frame(by, e.V, e.W, bp);
    solve_frame(S, bp, ay);
    to_frame(bz, e.V, e.W, bp);
    solve_frame(S, bp, az);
#pragma unroll
    for (int q = 0; q < 5; q++) sv[q] = e.V * ay[q] + e.W * az[q];
    const T de = e.den;
    // The u < 0 half table is the full table minus the u > 0 one, and every
    // moment is linear in the table, so <.. axl>_+ + <.. axr>_- is taken as
    // <.. axr>_full + <.. (axl - axr)>_+ (one erfc/exp table fewer).
    NTab<T> ep, ef;
    half_ntab(g, e.U, e.lam, he, +1, ep);
    full_ntab(g, e.U, he, ef);
    T dax[5];
#pragma unroll
    for (int q = 0; q < 5; q++) dax[q] = axl[q] - axr[q];
    // time-slope source of the equilibrium part, added in the frame:
    //   rho gm2 (G_0(axr, ay, az, sv) + u_1(axl - axr)_+) + rho gm1 (N_0, N_1, 0, 0, Q_0)
    GCoef<T> gc;
    g_coef(axr, ay, az, sv, he, gc);
    T G0[5], G1[5], ud1[5], ud2[5];
    g_pair<0>(ef, gc, G0, G1);
    u_pair<1>(ep, he, dax, ud1, ud2);
    to_frame(B, e.V, e.W, bp);
    {
        const T dg1 = de * gm1, dg2 = de * gm2;
#pragma unroll
        for (int q = 0; q < 5; q++) bp[q] += dg2 * (G0[q] + ud1[q]);
        bp[0] += dg1 * ef.N[0];
        bp[1] += dg1 * ef.N[1];
        bp[4] += dg1 * ef.Q[0];
    }
    const T ide = e.ri;
    solve_scaled(S, ide, S.k4 * ide, S.l2 * ide, bp, A);
    {
        // flux: rho p2 (G_1 + u_2(axl - axr)_+) + rho p1 u_1(A) + rho p0 (N_1, N_2, 0, 0, Q_1)
        T m[5], uA[5];
        u_mom<1>(ef, he, A, uA);
        const T dp0 = de * p0, dp1 = de * p1, dp2 = de * p2;
#pragma unroll
        for (int q = 0; q < 5; q++) m[q] = dp2 * (G1[q] + ud2[q]) + dp1 * uA[q];
        m[0] += dp0 * ef.N[1];
        m[1] += dp0 * ef.N[2];
        m[4] += dp0 * ef.Q[1];
        add_to_lab(Fa, m, e.V, e.W);
    }
#pragma unroll
    for (int q = 0; q < 5; q++) F[q] = Fa[q];
}

// Smooth-flow gas-kinetic flux of the GKM tree (GKM/CUDA GKM/src/GKMflux.cpp:12-280):
// one Maxwellian at the interface average W0 with slopes ax, ay, az, the time
// slope A from <psi (A.psi)> = -(<u psi ax> + <v psi ay> + <w psi az>), and
//   F = rho0 ((p1 + p2) <u psi A> + p1 (<u^2 psi ax> + <uv psi ay> + <uw psi az>) + p0 <u psi>)
// with p0 = dt, p1 = -tau dt, p2 = dt^2/2 (:269-278).  Taken in the Maxwellian's
// tangentially co-moving frame like bgk_flux: v = v' + V and w = w' + W turn
// <v psi ay> into <v' psi' ay> + V <psi' ay> (the sv terms).
template <class T>
__device__ __forceinline__ void gks_flux(const Gas<T>& g, const T W0[5], const T DWx[5], const T DWy[5],
                                         const T DWz[5], T tau, T dt, T F[5]) {
    Maxw<T> s;
    maxw_from_W(g, W0, s);
    const T h = s.h;
    NTab<T> nf;
    full_ntab(g, s.U, h, nf);
    Slv<T> S;
    slv_make(g, s.U, s.lam, h, S);
    const T ri = s.ri, k4r = S.k4 * ri, l2r = S.l2 * ri;
    T bp[5], ax[5], ay[5], az[5], nA[5], sv[5];
    to_frame(DWx, s.V, s.W, bp);
    solve_scaled(S, ri, k4r, l2r, bp, ax);
    to_frame(DWy, s.V, s.W, bp);
    solve_scaled(S, ri, k4r, l2r, bp, ay);
    to_frame(DWz, s.V, s.W, bp);
    solve_scaled(S, ri, k4r, l2r, bp, az);
#pragma unroll
    for (int q = 0; q < 5; q++) sv[q] = s.V * ay[q] + s.W * az[q];
    // A = -solve(G_0), nA = -A; F = rho (p1 G_1 + (p1 + p2) u_1(A) + p0 (N_1, N_2, 0, 0, Q_1))
    GCoef<T> gc;
    g_coef(ax, ay, az, sv, h, gc);
    T G0[5], G1[5], uA[5], m[5];
    g_pair<0>(nf, gc, G0, G1);
    solve_frame(S, G0, nA);
    u_mom<1>(nf, h, nA, uA);
    const T p0 = dt, p1 = -tau * dt, p2 = T(0.5) * dt * dt;
    const T dp0 = s.den * p0, dp1 = s.den * p1, dp12 = s.den * (p1 + p2);
#pragma unroll
    for (int q = 0; q < 5; q++) m[q] = dp1 * G1[q] - dp12 * uA[q];
    m[0] += dp0 * nf.N[1];
    m[1] += dp0 * nf.N[2];
    m[4] += dp0 * nf.Q[1];
    to_lab<false>(F, m, s.V, s.W);
}

}  // namespace gkm
