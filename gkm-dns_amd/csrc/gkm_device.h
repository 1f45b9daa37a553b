// gkm_device.h — device arithmetic of the GKM-DNS step for gfx950:
// WENO-SYMOO / WENO-JS interface reconstruction and the BGK-NS interface flux,
// templated on the arithmetic type (double for the drop-in form, float for the
// mixed form).
//
// The flux follows weno/src/GKMflux.cpp:11-215 (the same moments, slope solves
// and time-integral coefficients) but is reorganised for register pressure:
// the reference keeps ~400 doubles live (seven [3][7] tables and ~50
// five-vectors, GPUupdate.cu:499-520).  Here tau, the gamma and p coefficients
// are formed first, then each Maxwellian (left, right, equilibrium) is visited
// once and its contributions are accumulated into four 5-vectors (the
// tangential slope sources of the equilibrium state, the time-slope source and
// the flux), so a state's tables and slopes die as soon as its block ends.
// Only the normal axis carries half-space moments (the reference fills them
// for all three axes, but MCal reads them only from row 0).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace gkm {

template <class T> struct M;
// FP64: the ROCm device library for erfc/exp/pow; reciprocals from v_rcp_f64
// refined by two Newton steps (within an ulp of 1/x for the normal-range
// values that occur here; the IEEE division sequence is about twice as long).
template <> struct M<double> {
    __device__ static __forceinline__ double erfc_(double x) { return ::erfc(x); }
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
};
// FP32 (mixed form): hardware reciprocal, exponential and square root
// (v_rcp_f32, v_exp_f32, v_sqrt_f32, about 1 ulp); erfc from the device library.
template <> struct M<float> {
    __device__ static __forceinline__ float erfc_(float x) { return ::erfcf(x); }
    __device__ static __forceinline__ float exp_(float x) { return __expf(x); }
    __device__ static __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
    __device__ static __forceinline__ float pow_(float x, float y) { return ::powf(x, y); }
    __device__ static __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
};

// ---------------------------------------------------------------------------
// WENO (one scalar line of six cells -> values left and right of the centre
// interface).  SYMOO: weno/src/update.cpp:507-557, GPUupdate.cu:917-977.
// JS: update.cpp:1421-1468.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void weno_symoo(T LLL, T LL, T L, T R, T RR, T RRR, T eps, T& wl, T& wr) {
    const T c13 = T(1) / T(3), c56 = T(5) / T(6), c16 = T(1) / T(6), c76 = T(7) / T(6), c116 = T(11) / T(6);
    const T c1312 = T(13) / T(12);
    T P0p = LLL * c13 - LL * c76 + L * c116;
    T P0n = RRR * c13 - RR * c76 + R * c116;
    T P1p = -LL * c16 + L * c56 + R * c13;
    T P1n = -RR * c16 + R * c56 + L * c13;
    T P2p = L * c13 + R * c56 - RR * c16;
    T P2n = R * c13 + L * c56 - LL * c16;
    T P3p = R * c116 - RR * c76 + RRR * c13;
    T P3n = L * c116 - LL * c76 + LLL * c13;
    T a, b;
    a = LLL - 2 * LL + L; b = LLL - 4 * LL + 3 * L; T IS0p = c1312 * a * a + T(0.25) * b * b;
    a = RRR - 2 * RR + R; b = RRR - 4 * RR + 3 * R; T IS0n = c1312 * a * a + T(0.25) * b * b;
    a = LL - 2 * L + R;   b = R - LL;               T IS1p = c1312 * a * a + T(0.25) * b * b;
    a = RR - 2 * R + L;   b = L - RR;               T IS1n = c1312 * a * a + T(0.25) * b * b;
    a = L - 2 * R + RR;   b = -3 * L + 4 * R - RR;  T IS2p = c1312 * a * a + T(0.25) * b * b;
    a = R - 2 * L + LL;   b = -3 * R + 4 * L - LL;  T IS2n = c1312 * a * a + T(0.25) * b * b;
    a = R - 2 * RR + RRR; b = -5 * R + 8 * RR - 3 * RRR; T IS3p = c1312 * a * a + T(0.25) * b * b;
    a = L - 2 * LL + LLL; b = -5 * L + 8 * LL - 3 * LLL; T IS3n = c1312 * a * a + T(0.25) * b * b;
    // the downwind indicator is replaced by the largest of the four
    IS3p = fmax(fmax(fmax(IS0p, IS1p), IS2p), IS3p);
    IS3n = fmax(fmax(fmax(IS0n, IS1n), IS2n), IS3n);
    const T g0 = T(1) / T(20), g1 = T(9) / T(20);
    T A0 = g0 * M<T>::rcp(eps + IS0p), A1 = g1 * M<T>::rcp(eps + IS1p);
    T A2 = g1 * M<T>::rcp(eps + IS2p), A3 = g0 * M<T>::rcp(eps + IS3p);
    wl = (A0 * P0p + A1 * P1p + A2 * P2p + A3 * P3p) * M<T>::rcp(A0 + A1 + A2 + A3);
    A0 = g0 * M<T>::rcp(eps + IS0n); A1 = g1 * M<T>::rcp(eps + IS1n);
    A2 = g1 * M<T>::rcp(eps + IS2n); A3 = g0 * M<T>::rcp(eps + IS3n);
    wr = (A0 * P0n + A1 * P1n + A2 * P2n + A3 * P3n) * M<T>::rcp(A0 + A1 + A2 + A3);
}

template <class T>
__device__ __forceinline__ void weno_js(T LLL, T LL, T L, T R, T RR, T RRR, T eps, T& wl, T& wr) {
    const T c13 = T(1) / T(3), c56 = T(5) / T(6), c16 = T(1) / T(6), c76 = T(7) / T(6), c116 = T(11) / T(6);
    const T c1312 = T(13) / T(12);
    T P0p = LLL * c13 - LL * c76 + L * c116;
    T P0n = R * c116 - RR * c76 + RRR * c13;
    T P1p = -LL * c16 + L * c56 + R * c13;
    T P1n = L * c13 + R * c56 - RR * c16;
    T P2p = P1n;  // L/3 + 5R/6 - RR/6
    T P2n = P1p;  // -LL/6 + 5L/6 + R/3
    T a, b;
    a = LLL - 2 * LL + L; b = LLL - 4 * LL + 3 * L; T IS0p = c1312 * a * a + T(0.25) * b * b;
    a = RRR - 2 * RR + R; b = RRR - 4 * RR + 3 * R; T IS0n = c1312 * a * a + T(0.25) * b * b;
    a = LL - 2 * L + R;   b = LL - R;               T IS1p = c1312 * a * a + T(0.25) * b * b;
    a = L - 2 * R + RR;   b = L - RR;               T IS1n = c1312 * a * a + T(0.25) * b * b;
    b = 3 * L - 4 * R + RR;                         T IS2p = c1312 * a * a + T(0.25) * b * b;
    a = LL - 2 * L + R;   b = LL - 4 * L + 3 * R;   T IS2n = c1312 * a * a + T(0.25) * b * b;
    T r;
    r = M<T>::rcp(eps + IS0p); T A0 = T(0.1) * r * r;
    r = M<T>::rcp(eps + IS1p); T A1 = T(0.6) * r * r;
    r = M<T>::rcp(eps + IS2p); T A2 = T(0.3) * r * r;
    wl = (A0 * P0p + A1 * P1p + A2 * P2p) * M<T>::rcp(A0 + A1 + A2);
    r = M<T>::rcp(eps + IS0n); A0 = T(0.1) * r * r;
    r = M<T>::rcp(eps + IS1n); A1 = T(0.6) * r * r;
    r = M<T>::rcp(eps + IS2n); A2 = T(0.3) * r * r;
    wr = (A0 * P0n + A1 * P1n + A2 * P2n) * M<T>::rcp(A0 + A1 + A2);
}

template <int RECON, class T>
__device__ __forceinline__ void weno_line(const T c[6], T eps, T& wl, T& wr) {
    if (RECON == 1)
        weno_js(c[0], c[1], c[2], c[3], c[4], c[5], eps, wl, wr);
    else
        weno_symoo(c[0], c[1], c[2], c[3], c[4], c[5], eps, wl, wr);
}

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

// <u^k v^l w^m psi>, psi = (1, u, v, w, (u^2+v^2+w^2+xi^2)/2)  (GKMflux.cpp:266-275)
template <int k, int l, int m, class T>
__device__ __forceinline__ void mom(T Mo[5], const T* n, const T* v, const T* w, T e2) {
    Mo[0] = n[k] * v[l] * w[m];
    Mo[1] = n[1 + k] * v[l] * w[m];
    Mo[2] = n[k] * v[1 + l] * w[m];
    Mo[3] = n[k] * v[l] * w[1 + m];
    Mo[4] = T(0.5) * (n[2 + k] * v[l] * w[m] + n[k] * v[2 + l] * w[m] + n[k] * v[l] * w[2 + m] +
                      n[k] * v[l] * w[m] * e2);
}

// <u^k v^l w^m psi (a . psi)>  (GKMflux.cpp:225-264), accumulated as acc += s * M
template <int k, int l, int m, class T>
__device__ __forceinline__ void mom_a_acc(T acc[5], T s, const T* n, const T* v, const T* w, T e2, T e4,
                                          const T a[5]) {
    const T v0 = T(0.5) * (n[2 + k] * v[l] * w[m] + n[k] * v[2 + l] * w[m] + n[k] * v[l] * w[2 + m] +
                           n[k] * v[l] * w[m] * e2);
    const T v1 = T(0.5) * (n[3 + k] * v[l] * w[m] + n[1 + k] * v[2 + l] * w[m] + n[1 + k] * v[l] * w[2 + m] +
                           n[1 + k] * v[l] * w[m] * e2);
    const T v2 = T(0.5) * (n[2 + k] * v[1 + l] * w[m] + n[k] * v[3 + l] * w[m] + n[k] * v[1 + l] * w[2 + m] +
                           n[k] * v[1 + l] * w[m] * e2);
    const T v3 = T(0.5) * (n[2 + k] * v[l] * w[1 + m] + n[k] * v[2 + l] * w[1 + m] + n[k] * v[l] * w[3 + m] +
                           n[k] * v[l] * w[1 + m] * e2);
    T m0 = a[0] * (n[k] * v[l] * w[m]) + a[1] * (n[1 + k] * v[l] * w[m]) + a[2] * (n[k] * v[1 + l] * w[m]) +
           a[3] * (n[k] * v[l] * w[1 + m]) + a[4] * v0;
    T m1 = a[0] * (n[1 + k] * v[l] * w[m]) + a[1] * (n[2 + k] * v[l] * w[m]) +
           a[2] * (n[1 + k] * v[1 + l] * w[m]) + a[3] * (n[1 + k] * v[l] * w[1 + m]) + a[4] * v1;
    T m2 = a[0] * (n[k] * v[1 + l] * w[m]) + a[1] * (n[1 + k] * v[1 + l] * w[m]) +
           a[2] * (n[k] * v[2 + l] * w[m]) + a[3] * (n[k] * v[1 + l] * w[1 + m]) + a[4] * v2;
    T m3 = a[0] * (n[k] * v[l] * w[1 + m]) + a[1] * (n[1 + k] * v[l] * w[1 + m]) +
           a[2] * (n[k] * v[1 + l] * w[1 + m]) + a[3] * (n[k] * v[l] * w[2 + m]) + a[4] * v3;
    T m4 = T(0.25) * a[4] *
               (n[4 + k] * v[l] * w[m] + n[k] * v[4 + l] * w[m] + n[k] * v[l] * w[4 + m] + n[k] * v[l] * w[m] * e4 +
                2 * n[2 + k] * v[2 + l] * w[m] + 2 * n[2 + k] * v[l] * w[2 + m] + 2 * n[2 + k] * v[l] * w[m] * e2 +
                2 * n[k] * v[2 + l] * w[2 + m] + 2 * n[k] * v[2 + l] * w[m] * e2 + 2 * n[k] * v[l] * w[2 + m] * e2) +
           a[0] * v0 + a[1] * v1 + a[2] * v2 + a[3] * v3;
    acc[0] += s * m0;
    acc[1] += s * m1;
    acc[2] += s * m2;
    acc[3] += s * m3;
    acc[4] += s * m4;
}

// Closed-form inverse of the Maxwellian moment matrix: b -> a (GKMflux.cpp:280-296);
// h = 1/(2 lam).
template <class T>
__device__ __forceinline__ void slope_solve(const Gas<T>& g, const T b[5], const T U[3], T lam, T h, T a[5]) {
    const T uu = U[0] * U[0] + U[1] * U[1] + U[2] * U[2];
    const T R2 = b[1] - U[0] * b[0];
    const T R3 = b[2] - U[1] * b[0];
    const T R4 = b[3] - U[2] * b[0];
    const T q = g.kp3 * h;  // (K+3)/(2 lam)
    const T R5 = 2 * b[4] - b[0] * (uu + q);
    a[4] = g.invPr * (R5 - 2 * U[0] * R2 - 2 * U[1] * R3 - 2 * U[2] * R4) * (4 * lam * lam) * g.ikp3;
    a[3] = 2 * lam * R4 - U[2] * a[4];
    a[2] = 2 * lam * R3 - U[1] * a[4];
    a[1] = 2 * lam * R2 - U[0] * a[4];
    a[0] = b[0] - a[1] * U[0] - a[2] * U[1] - a[3] * U[2] - T(0.5) * a[4] * (uu + q);
}

// One Maxwellian: primitive variables and the full-space tables along the
// three axes (index 0..6).
template <class T> struct Maxw {
    T den, U[3], lam, h;  // h = 1/(2 lam)
    T e2, e4;
};

template <class T>
__device__ __forceinline__ void maxw_from_W(const Gas<T>& g, const T W[5], Maxw<T>& s) {
    s.den = W[0];
    const T ri = M<T>::rcp(W[0]);
    s.U[0] = W[1] * ri;
    s.U[1] = W[2] * ri;
    s.U[2] = W[3] * ri;
    const T P = (s.den * (g.gam - 1)) * (W[4] * ri - T(0.5) * (s.U[0] * s.U[0] + s.U[1] * s.U[1] + s.U[2] * s.U[2]));
    s.lam = T(0.5) * s.den * M<T>::rcp(P);
    s.h = M<T>::rcp(2 * s.lam);
    s.e2 = g.K * s.h;                                  // K/(2 lam)
    s.e4 = (3 * g.K + g.K * (g.K - 1)) * s.h * s.h;    // (3K + K(K-1))/(4 lam^2)
}

template <class T> __device__ __forceinline__ void full_table(T U, T h, T f[7]) {
    f[0] = T(1);
    f[1] = U;
#pragma unroll
    for (int i = 2; i < 7; i++) f[i] = U * f[i - 1] + f[i - 2] * T(i - 1) * h;
}

// half-space moments along the normal axis; sgn = +1 for u > 0, -1 for u < 0
template <class T> __device__ __forceinline__ void half_table(T U, T lam, T h, int sgn, T f[7]) {
    const T sl = M<T>::sqrt_(lam);
    const T pi = T(3.141592653589793238462643383279502884);
    f[0] = T(0.5) * M<T>::erfc_(-T(sgn) * sl * U);
    f[1] = U * f[0] + T(sgn) * T(0.5) * M<T>::exp_(-lam * U * U) * M<T>::rcp(M<T>::sqrt_(lam * pi));
#pragma unroll
    for (int i = 2; i < 7; i++) f[i] = U * f[i - 1] + f[i - 2] * T(i - 1) * h;
}

// Contribution of one side (left = sgn +1, right = sgn -1) to the flux.
//   W     : reconstructed interface state of that side
//   DWx/y/z: slopes of that side (normal, tangential-1, tangential-2)
// Accumulates: by, bz (sources of the equilibrium tangential slopes),
//   B (source of the equilibrium time slope), F (the flux).
template <class T>
__device__ __forceinline__ void side_block(const Gas<T>& g, const Maxw<T>& s, const T* hs, const T DWx[5],
                                           const T DWy[5], const T DWz[5], T c45, T c5, T p3, T p4, T p5, T by[5],
                                           T bz[5], T B[5], T F[5]) {
    T fn[7], fv[7], fw[7];
    full_table(s.U[0], s.h, fn);
    full_table(s.U[1], s.h, fv);
    full_table(s.U[2], s.h, fw);
    const T ri = M<T>::rcp(s.den);
    T b[5], ax[5], ay[5], az[5], A[5];
#pragma unroll
    for (int q = 0; q < 5; q++) b[q] = DWx[q] * ri;
    slope_solve(g, b, s.U, s.lam, s.h, ax);
#pragma unroll
    for (int q = 0; q < 5; q++) b[q] = DWy[q] * ri;
    slope_solve(g, b, s.U, s.lam, s.h, ay);
#pragma unroll
    for (int q = 0; q < 5; q++) b[q] = DWz[q] * ri;
    slope_solve(g, b, s.U, s.lam, s.h, az);
    // time slope of this side: A = -(<u psi ax> + <v psi ay> + <w psi az>) over the full space
#pragma unroll
    for (int q = 0; q < 5; q++) b[q] = T(0);
    mom_a_acc<1, 0, 0>(b, T(-1), fn, fv, fw, s.e2, s.e4, ax);
    mom_a_acc<0, 1, 0>(b, T(-1), fn, fv, fw, s.e2, s.e4, ay);
    mom_a_acc<0, 0, 1>(b, T(-1), fn, fv, fw, s.e2, s.e4, az);
    slope_solve(g, b, s.U, s.lam, s.h, A);
    // equilibrium tangential-slope sources: <psi ay>_half, <psi az>_half
    mom_a_acc<0, 0, 0>(by, T(1), hs, fv, fw, s.e2, s.e4, ay);
    mom_a_acc<0, 0, 0>(bz, T(1), hs, fv, fw, s.e2, s.e4, az);
    const T d = s.den;
    // time-slope source: (gm4+gm5) rho (<u psi ax> + <v psi ay> + <w psi az>)_half + gm5 rho <psi A>_half
    const T dc45 = d * c45;
    mom_a_acc<1, 0, 0>(B, dc45, hs, fv, fw, s.e2, s.e4, ax);
    mom_a_acc<0, 1, 0>(B, dc45, hs, fv, fw, s.e2, s.e4, ay);
    mom_a_acc<0, 0, 1>(B, dc45, hs, fv, fw, s.e2, s.e4, az);
    mom_a_acc<0, 0, 0>(B, d * c5, hs, fv, fw, s.e2, s.e4, A);
    // flux: p3 rho <u psi> + p4 rho (<u^2 psi ax> + <uv psi ay> + <uw psi az>) + p5 rho <u psi A>
    T m[5];
    mom<1, 0, 0>(m, hs, fv, fw, s.e2);
    const T dp3 = d * p3;
#pragma unroll
    for (int q = 0; q < 5; q++) F[q] += dp3 * m[q];
    const T dp4 = d * p4;
    mom_a_acc<2, 0, 0>(F, dp4, hs, fv, fw, s.e2, s.e4, ax);
    mom_a_acc<1, 1, 0>(F, dp4, hs, fv, fw, s.e2, s.e4, ay);
    mom_a_acc<1, 0, 1>(F, dp4, hs, fv, fw, s.e2, s.e4, az);
    mom_a_acc<1, 0, 0>(F, d * p5, hs, fv, fw, s.e2, s.e4, A);
}

// BGK-NS interface flux, time-integrated over dt (GKMflux.cpp:11-215).
// All vectors are in the rotated slot order (slot 1 = normal momentum).
// tau0 = mu/p of the left cell.  DW(l|r) = slopes left / right of the
// interface along the normal (x) and the two tangential (y, z) directions.
template <class T, class LoadSlopes>
__device__ __forceinline__ void bgk_flux(const Gas<T>& g, const T WL[5], const T WR[5], const T Wl[5],
                                         const T Wr[5], LoadSlopes&& slopes, T tau0, T dt, T dx, T F[5]) {
    Maxw<T> l, r;
    maxw_from_W(g, Wl, l);
    maxw_from_W(g, Wr, r);
    // collision time with the pressure-jump blending (GKMflux.cpp:149-150)
    const T rl = l.den * M<T>::rcp(l.lam), rr = r.den * M<T>::rcp(r.lam);
    const T tau = tau0 + dt * fabs(rl - rr) * M<T>::rcp(fabs(rl + rr));
    const T ex = M<T>::exp_(-dt * M<T>::rcp(tau));
    const T om = T(1) - ex;
    const T gm0 = dt - tau * om;
    const T igm0 = M<T>::rcp(gm0);
    const T gm1 = -om * igm0;
    const T gm2 = (-dt + 2 * tau * om - dt * ex) * igm0;
    const T gm3 = -gm1;
    const T gm4 = (dt * ex - tau * om) * igm0;
    const T gm5 = tau * gm3;
    const T P = tau * om;
    const T Q = -tau * dt * ex + tau * tau * om;
    const T p0 = dt - P;
    const T p1 = T(0.5) * dt * dt - tau * p0;
    const T p2 = -tau * p0 + Q;
    const T p3 = P;
    const T p4 = -Q - tau * P;
    const T p5 = -tau * P;

    // half-space normal tables of the two sides and the equilibrium state
    T hl[7], hr[7];
    half_table(l.U[0], l.lam, l.h, +1, hl);
    half_table(r.U[0], r.lam, r.h, -1, hr);
    T We[5];
    {
        T fv[7], fw[7], m[5];
        full_table(l.U[1], l.h, fv);
        full_table(l.U[2], l.h, fw);
        mom<0, 0, 0>(m, hl, fv, fw, l.e2);
#pragma unroll
        for (int q = 0; q < 5; q++) We[q] = l.den * m[q];
        full_table(r.U[1], r.h, fv);
        full_table(r.U[2], r.h, fw);
        mom<0, 0, 0>(m, hr, fv, fw, r.e2);
#pragma unroll
        for (int q = 0; q < 5; q++) We[q] += r.den * m[q];
    }
    Maxw<T> e;
    maxw_from_W(g, We, e);

    T by[5], bz[5], B[5], Fa[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        by[q] = T(0);
        bz[q] = T(0);
        B[q] = gm3 * We[q];
        Fa[q] = T(0);
    }
    const T c45 = gm4 + gm5;
    {
        T Dx[5], Dy[5], Dz[5];
        slopes(0, Dx, Dy, Dz);  // left side
        side_block(g, l, hl, Dx, Dy, Dz, c45, gm5, p3, p4, p5, by, bz, B, Fa);
    }
    {
        T Dx[5], Dy[5], Dz[5];
        slopes(1, Dx, Dy, Dz);  // right side
        side_block(g, r, hr, Dx, Dy, Dz, c45, gm5, p3, p4, p5, by, bz, B, Fa);
    }

    // equilibrium state: normal slopes from the cell averages, tangential slopes from by, bz
    T he_p[7], he_n[7], fn[7], fv[7], fw[7];
    half_table(e.U[0], e.lam, e.h, +1, he_p);
    half_table(e.U[0], e.lam, e.h, -1, he_n);
    full_table(e.U[0], e.h, fn);
    full_table(e.U[1], e.h, fv);
    full_table(e.U[2], e.h, fw);
    T axl[5], axr[5], ay[5], az[5], A[5], b[5];
    const T s2 = T(2) * M<T>::rcp(e.den * dx);
#pragma unroll
    for (int q = 0; q < 5; q++) b[q] = (We[q] - WL[q]) * s2;
    slope_solve(g, b, e.U, e.lam, e.h, axl);
#pragma unroll
    for (int q = 0; q < 5; q++) b[q] = (WR[q] - We[q]) * s2;
    slope_solve(g, b, e.U, e.lam, e.h, axr);
    slope_solve(g, by, e.U, e.lam, e.h, ay);
    slope_solve(g, bz, e.U, e.lam, e.h, az);
    const T de = e.den;
    {
        T m[5];
        mom<0, 0, 0>(m, fn, fv, fw, e.e2);
        const T dg1 = de * gm1;
#pragma unroll
        for (int q = 0; q < 5; q++) B[q] += dg1 * m[q];
    }
    const T dg2 = de * gm2;
    mom_a_acc<1, 0, 0>(B, dg2, he_p, fv, fw, e.e2, e.e4, axl);
    mom_a_acc<1, 0, 0>(B, dg2, he_n, fv, fw, e.e2, e.e4, axr);
    mom_a_acc<0, 1, 0>(B, dg2, fn, fv, fw, e.e2, e.e4, ay);
    mom_a_acc<0, 0, 1>(B, dg2, fn, fv, fw, e.e2, e.e4, az);
    const T ide = M<T>::rcp(de);
#pragma unroll
    for (int q = 0; q < 5; q++) B[q] *= ide;
    slope_solve(g, B, e.U, e.lam, e.h, A);
    {
        T m[5];
        mom<1, 0, 0>(m, fn, fv, fw, e.e2);
        const T dp0 = de * p0;
#pragma unroll
        for (int q = 0; q < 5; q++) Fa[q] += dp0 * m[q];
    }
    mom_a_acc<1, 0, 0>(Fa, de * p1, fn, fv, fw, e.e2, e.e4, A);
    const T dp2 = de * p2;
    mom_a_acc<2, 0, 0>(Fa, dp2, he_p, fv, fw, e.e2, e.e4, axl);
    mom_a_acc<2, 0, 0>(Fa, dp2, he_n, fv, fw, e.e2, e.e4, axr);
    mom_a_acc<1, 1, 0>(Fa, dp2, fn, fv, fw, e.e2, e.e4, ay);
    mom_a_acc<1, 0, 1>(Fa, dp2, fn, fv, fw, e.e2, e.e4, az);
#pragma unroll
    for (int q = 0; q < 5; q++) F[q] = Fa[q];
}

}  // namespace gkm
