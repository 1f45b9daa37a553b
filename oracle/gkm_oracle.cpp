/*
 * gkm_oracle.cpp — CPU oracle of the GKM-DNS time step (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libgkm.so and the CPU baseline
 * of bench.py.  Nothing in the product links or calls this file.
 *
 * Each function restates the weno CPU path and keeps the reference's
 * expression order, so that it can be compared bit-for-bit with the
 * reference flux compiled from weno/src/GKMflux.cpp (oracle/Makefile).
 * Built with -O2 -ffp-contract=off (no FMA contraction, as the reference's
 * x86-64 build).
 *
 * Parity pinning (which check fixes which function):
 *   orc_flux, orc_slopesolver, moment tables  -> bit-compared with the
 *        compiled reference flux() on random and jump states
 *        (tests/golden/flux_golden.npz) plus closed forms (uniform state ->
 *        dt x Euler flux; Gaussian moments; inverse identity at Pr = 1).
 *   orc_weno / orc_weno_scalar -> an independent rebuild of SYMOO and JS from
 *        their definitions (candidate quadratics solved from the cell-average
 *        integrals, smoothness indicators integrated exactly over the upwind
 *        cell, the published ideal weights and the downwind maximum) on rough,
 *        step, spike and linear lines (tests/test_weno_pin.py), plus closed
 *        forms (constant data, quadratic reproduction, smooth limit, mirror
 *        symmetry).
 *   orc_step -> telescoping conservation, uniform state fixed bitwise,
 *        X/Y axis-permutation symmetry (Z too with zbug_compat = 0), and
 *        composition of the pinned parts on single interfaces.
 *   orc_stats -> closed-form fields (Fourier modes with their exact central
 *        differences, Taylor-Green cells without dilatation, a density wave,
 *        a skewed two-mode line) in tests/test_stats_pin.py.
 *   orc_spline_slopes / orc_print_stats -> scipy's natural CubicSpline and a
 *        numpy restatement of PrintStats (tests/test_print_stats.py).
 */
#include "gkm_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* Reference macros pi and eps (weno/src/param.h:14-15), evaluated the same way. */
inline double ref_pi() { return std::atan(1.0) * 4.0; }
inline double ref_eps() { return std::pow(10.0, -10.0); }

/* Conservative -> primitive inside the flux (GKMflux.cpp:301-308). */
inline void c2p_flux(double gam, const double W[5], double& den, double U[3], double& P) {
    den = W[0];
    U[0] = W[1] / den;
    U[1] = W[2] / den;
    U[2] = W[3] / den;
    P = (den * (gam - 1)) * (W[4] / den - 0.5 * (U[0] * U[0] + U[1] * U[1] + U[2] * U[2]));
}

/* Moment tables of one Maxwellian along one axis (GKMflux.cpp:48-66):
 * hp / hn are the half-space (u > 0 / u < 0) moments, fu the full ones. */
struct Axis {
    double hp[7], hn[7], fu[7];
};

inline void axis_tables(double U, double lam, Axis& a, bool want_pos, bool want_neg) {
    a.fu[0] = 1.0;
    a.fu[1] = U;
    if (want_pos) {
        a.hp[0] = 0.5 * (std::erfc(-std::sqrt(lam) * U));
        a.hp[1] = U * a.hp[0] + 0.5 * std::exp(-lam * U * U) / std::sqrt(lam * std::atan(1.0) * 4.0);
    }
    if (want_neg) {
        a.hn[0] = 0.5 * (std::erfc(std::sqrt(lam) * U));
        a.hn[1] = U * a.hn[0] - 0.5 * std::exp(-lam * U * U) / std::sqrt(lam * std::atan(1.0) * 4.0);
    }
    for (int i = 2; i < 7; i++) {
        if (want_pos) a.hp[i] = U * a.hp[i - 1] + a.hp[i - 2] * (i - 1) / (2 * lam);
        a.fu[i] = U * a.fu[i - 1] + a.fu[i - 2] * (i - 1) / (2 * lam);
        if (want_neg) a.hn[i] = U * a.hn[i - 1] + a.hn[i - 2] * (i - 1) / (2 * lam);
    }
}

/* A Maxwellian state: density, velocity, lambda = rho/(2p) and the
 * per-axis tables plus the internal-energy moments <xi^2>, <xi^4>. */
struct State {
    double den, U[3], P, lam;
    Axis ax[3];
    double e2, e4;
};

inline void make_state(double gam, int K, const double W[5], State& s, bool pos, bool neg) {
    c2p_flux(gam, W, s.den, s.U, s.P);
    s.lam = 0.5 * s.den / s.P;
    s.e2 = K / (2 * s.lam);
    s.e4 = 3 * K / (4 * s.lam * s.lam) + K * (K - 1) / (4 * s.lam * s.lam);
    /* only the normal axis needs the half-space tables; the reference also
       fills them for the tangential axes, where they are never read */
    for (int j = 0; j < 3; j++) axis_tables(s.U[j], s.lam, s.ax[j], pos && j == 0, neg && j == 0);
}

/* <u^k v^l w^m psi> with psi = (1, u, v, w, (u^2+v^2+w^2+xi^2)/2); the
 * normal-axis table is n, the tangential ones v, w (GKMflux.cpp:266-275). */
inline void mom(double M[5], const double* n, const double* v, const double* w, double e2,
                int k, int l, int m) {
    double val0 = 0.5 * (n[2 + k] * v[l] * w[m] + n[k] * v[2 + l] * w[m] + n[k] * v[l] * w[2 + m] +
                         n[k] * v[l] * w[m] * e2);
    M[0] = n[k] * v[l] * w[m];
    M[1] = n[1 + k] * v[l] * w[m];
    M[2] = n[k] * v[1 + l] * w[m];
    M[3] = n[k] * v[l] * w[1 + m];
    M[4] = val0;
}

/* <u^k v^l w^m psi (a . psi)> (GKMflux.cpp:225-264). */
inline void mom_a(double M[5], const double* n, const double* v, const double* w, double e2,
                  double e4, int k, int l, int m, const double a[5]) {
    double val0 = 0.5 * (n[2 + k] * v[l] * w[m] + n[k] * v[2 + l] * w[m] + n[k] * v[l] * w[2 + m] +
                         n[k] * v[l] * w[m] * e2);
    M[0] = a[0] * (n[k] * v[l] * w[m]) + a[1] * (n[1 + k] * v[l] * w[m]) + a[2] * (n[k] * v[1 + l] * w[m]) +
           a[3] * (n[k] * v[l] * w[1 + m]) + a[4] * val0;
    double val1 = 0.5 * (n[3 + k] * v[l] * w[m] + n[1 + k] * v[2 + l] * w[m] + n[1 + k] * v[l] * w[2 + m] +
                         n[1 + k] * v[l] * w[m] * e2);
    M[1] = a[0] * (n[1 + k] * v[l] * w[m]) + a[1] * (n[2 + k] * v[l] * w[m]) +
           a[2] * (n[1 + k] * v[1 + l] * w[m]) + a[3] * (n[1 + k] * v[l] * w[1 + m]) + a[4] * val1;
    double val2 = 0.5 * (n[2 + k] * v[1 + l] * w[m] + n[k] * v[3 + l] * w[m] + n[k] * v[1 + l] * w[2 + m] +
                         n[k] * v[1 + l] * w[m] * e2);
    M[2] = a[0] * (n[k] * v[1 + l] * w[m]) + a[1] * (n[1 + k] * v[1 + l] * w[m]) +
           a[2] * (n[k] * v[2 + l] * w[m]) + a[3] * (n[k] * v[1 + l] * w[1 + m]) + a[4] * val2;
    double val3 = 0.5 * (n[2 + k] * v[l] * w[1 + m] + n[k] * v[2 + l] * w[1 + m] + n[k] * v[l] * w[3 + m] +
                         n[k] * v[l] * w[1 + m] * e2);
    M[3] = a[0] * (n[k] * v[l] * w[1 + m]) + a[1] * (n[1 + k] * v[l] * w[1 + m]) +
           a[2] * (n[k] * v[1 + l] * w[1 + m]) + a[3] * (n[k] * v[l] * w[2 + m]) + a[4] * val3;
    M[4] = 0.25 * a[4] *
               (n[4 + k] * v[l] * w[m] + n[k] * v[4 + l] * w[m] + n[k] * v[l] * w[4 + m] + n[k] * v[l] * w[m] * e4 +
                2 * n[2 + k] * v[2 + l] * w[m] + 2 * n[2 + k] * v[l] * w[2 + m] + 2 * n[2 + k] * v[l] * w[m] * e2 +
                2 * n[k] * v[2 + l] * w[2 + m] + 2 * n[k] * v[2 + l] * w[m] * e2 + 2 * n[k] * v[l] * w[2 + m] * e2) +
           a[0] * val0 + a[1] * val1 + a[2] * val2 + a[3] * val3;
}

/* Closed-form inverse of the Maxwellian moment matrix, b -> a
 * (GKMflux.cpp:280-296); the 1/Pr factor on a[4] is the CPU's (D4). */
inline void slopesolver(int K, double invPr, const double b[5], const double U[3], double lam, double a[5]) {
    double R2 = b[1] - U[0] * b[0];
    double R3 = b[2] - U[1] * b[0];
    double R4 = b[3] - U[2] * b[0];
    double R5 = 2 * b[4] - b[0] * (U[0] * U[0] + U[1] * U[1] + U[2] * U[2] + (K + 3) / (2 * lam));
    a[4] = invPr * (R5 - 2 * U[0] * R2 - 2 * U[1] * R3 - 2 * U[2] * R4) * (4 * lam * lam) / (K + 3);
    a[3] = 2 * lam * R4 - U[2] * a[4];
    a[2] = 2 * lam * R3 - U[1] * a[4];
    a[1] = 2 * lam * R2 - U[0] * a[4];
    a[0] = b[0] - a[1] * U[0] - a[2] * U[1] - a[3] * U[2] -
           .5 * a[4] * (U[0] * U[0] + U[1] * U[1] + U[2] * U[2] + (K + 3) / (2 * lam));
}

/* BGK-NS flux of one interface (GKMflux.cpp:11-215). */
void bgk_flux(const gkm_params& gp, const double WL[5], const double WR[5], const double Wl[5],
              const double Wr[5], const double DWxl[5], const double DWxr[5], const double DWyl[5],
              const double DWyr[5], const double DWzl[5], const double DWzr[5], double tau, double dt,
              double dx, double F[5]) {
    const double gam = gp.gamma;
    const int K = gp.K;
    const double invPr = 1 / gp.Pr;
    State l, r, e;
    make_state(gam, K, Wl, l, true, false);
    make_state(gam, K, Wr, r, false, true);
    const double* lp = l.ax[0].hp;  /* left state, u > 0 half   */
    const double* rn = r.ax[0].hn;  /* right state, u < 0 half  */
    const double* lf = l.ax[0].fu;
    const double* rf = r.ax[0].fu;
    const double *lv = l.ax[1].fu, *lw = l.ax[2].fu, *rv = r.ax[1].fu, *rw = r.ax[2].fu;

    double Mpgl000[5], Mngr000[5], We[5];
    mom(Mpgl000, lp, lv, lw, l.e2, 0, 0, 0);
    mom(Mngr000, rn, rv, rw, r.e2, 0, 0, 0);
    for (int q = 0; q < 5; q++) We[q] = l.den * Mpgl000[q] + r.den * Mngr000[q];

    make_state(gam, K, We, e, true, true);
    const double *ep = e.ax[0].hp, *en = e.ax[0].hn, *ef = e.ax[0].fu, *ev = e.ax[1].fu, *ew = e.ax[2].fu;

    double bxl[5], bxr[5], byl[5], byr[5], bzl[5], bzr[5], bx_l[5], bx_r[5];
    for (int i = 0; i < 5; i++) {
        bxl[i] = DWxl[i] / l.den;
        bxr[i] = DWxr[i] / r.den;
        byl[i] = DWyl[i] / l.den;
        byr[i] = DWyr[i] / r.den;
        bzl[i] = DWzl[i] / l.den;
        bzr[i] = DWzr[i] / r.den;
        bx_l[i] = 2 * (We[i] - WL[i]) / (e.den * dx);
        bx_r[i] = 2 * (WR[i] - We[i]) / (e.den * dx);
    }
    double axl[5], axr[5], ayl[5], ayr[5], azl[5], azr[5], ax_l[5], ax_r[5];
    slopesolver(K, invPr, bxl, l.U, l.lam, axl);
    slopesolver(K, invPr, bxr, r.U, r.lam, axr);
    slopesolver(K, invPr, byl, l.U, l.lam, ayl);
    slopesolver(K, invPr, byr, r.U, r.lam, ayr);
    slopesolver(K, invPr, bzl, l.U, l.lam, azl);
    slopesolver(K, invPr, bzr, r.U, r.lam, azr);
    slopesolver(K, invPr, bx_l, e.U, e.lam, ax_l);
    slopesolver(K, invPr, bx_r, e.U, e.lam, ax_r);

    double M1[5], M2[5], M3[5], M4[5], M5[5], M6[5], M7[5], M8[5], M9[5], M10[5];
    mom_a(M1, lp, lv, lw, l.e2, l.e4, 0, 0, 0, ayl);
    mom_a(M2, rn, rv, rw, r.e2, r.e4, 0, 0, 0, ayr);
    mom_a(M3, lp, lv, lw, l.e2, l.e4, 0, 0, 0, azl);
    mom_a(M4, rn, rv, rw, r.e2, r.e4, 0, 0, 0, azr);
    mom_a(M5, lf, lv, lw, l.e2, l.e4, 1, 0, 0, axl);
    mom_a(M6, rf, rv, rw, r.e2, r.e4, 1, 0, 0, axr);
    mom_a(M7, lf, lv, lw, l.e2, l.e4, 0, 1, 0, ayl);
    mom_a(M8, rf, rv, rw, r.e2, r.e4, 0, 1, 0, ayr);
    mom_a(M9, lf, lv, lw, l.e2, l.e4, 0, 0, 1, azl);
    mom_a(M10, rf, rv, rw, r.e2, r.e4, 0, 0, 1, azr);
    double by_[5], bz_[5], Bl[5], Br[5];
    for (int i = 0; i < 5; i++) {
        by_[i] = M1[i] + M2[i];
        bz_[i] = M3[i] + M4[i];
        Bl[i] = -M5[i] - M7[i] - M9[i];
        Br[i] = -M6[i] - M8[i] - M10[i];
    }
    double ay_[5], az_[5], Al[5], Ar[5];
    slopesolver(K, invPr, by_, e.U, e.lam, ay_);
    slopesolver(K, invPr, bz_, e.U, e.lam, az_);
    slopesolver(K, invPr, Bl, l.U, l.lam, Al);
    slopesolver(K, invPr, Br, r.U, r.lam, Ar);

    /* collision time: tau = mu/p + dt * |rho_l/lam_l - rho_r/lam_r| / |... + ...| */
    double w = std::abs(l.den / l.lam - r.den / r.lam) / (std::abs(l.den / l.lam + r.den / r.lam));
    tau = tau + dt * w;

    double gm0 = dt - tau * (1 - std::exp(-dt / tau));
    double gm1 = -(1 - std::exp(-dt / tau)) / gm0;
    double gm2 = (-dt + 2 * tau * (1 - std::exp(-dt / tau)) - dt * std::exp(-dt / tau)) / gm0;
    double gm3 = -gm1;
    double gm4 = (dt * std::exp(-dt / tau) - tau * (1 - std::exp(-dt / tau))) / gm0;
    double gm5 = tau * gm3;

    double Mfge000[5], Mpe100[5], Mne100[5], Mfe010[5], Mfe001[5], Mpl100a[5], Mnr100a[5], Mpl010[5],
        Mnr010[5], Mpl001[5], Mnr001[5], MplA[5], MnrA[5];
    mom(Mfge000, ef, ev, ew, e.e2, 0, 0, 0);
    mom_a(Mpe100, ep, ev, ew, e.e2, e.e4, 1, 0, 0, ax_l);
    mom_a(Mne100, en, ev, ew, e.e2, e.e4, 1, 0, 0, ax_r);
    mom_a(Mfe010, ef, ev, ew, e.e2, e.e4, 0, 1, 0, ay_);
    mom_a(Mfe001, ef, ev, ew, e.e2, e.e4, 0, 0, 1, az_);
    mom_a(Mpl100a, lp, lv, lw, l.e2, l.e4, 1, 0, 0, axl);
    mom_a(Mnr100a, rn, rv, rw, r.e2, r.e4, 1, 0, 0, axr);
    mom_a(Mpl010, lp, lv, lw, l.e2, l.e4, 0, 1, 0, ayl);
    mom_a(Mnr010, rn, rv, rw, r.e2, r.e4, 0, 1, 0, ayr);
    mom_a(Mpl001, lp, lv, lw, l.e2, l.e4, 0, 0, 1, azl);
    mom_a(Mnr001, rn, rv, rw, r.e2, r.e4, 0, 0, 1, azr);
    mom_a(MplA, lp, lv, lw, l.e2, l.e4, 0, 0, 0, Al);
    mom_a(MnrA, rn, rv, rw, r.e2, r.e4, 0, 0, 0, Ar);

    double B_[5], A_[5];
    const double dene = e.den, denl = l.den, denr = r.den;
    for (int i = 0; i < 5; i++) {
        B_[i] = gm1 * dene * Mfge000[i] + gm2 * dene * (Mpe100[i] + Mne100[i] + Mfe010[i] + Mfe001[i]) +
                gm3 * (denl * Mpgl000[i] + denr * Mngr000[i]) +
                (gm4 + gm5) * (denl * (Mpl100a[i] + Mpl010[i] + Mpl001[i]) + denr * (Mnr100a[i] + Mnr010[i] + Mnr001[i])) +
                gm5 * (denl * MplA[i] + denr * MnrA[i]);
        B_[i] = B_[i] / dene;
    }
    slopesolver(K, invPr, B_, e.U, e.lam, A_);

    double P = -tau * (std::exp(-dt / tau) - 1);
    double Q = -tau * dt * std::exp(-dt / tau) - tau * tau * (std::exp(-dt / tau) - 1);
    double p0 = (dt - P);
    double p1 = 0.5 * dt * dt - tau * (p0);
    double p2 = -tau * (p0) + Q;
    double p3 = P;
    double p4 = -Q - tau * P;
    double p5 = -tau * P;

    double Mpl100[5], Mnr100[5], Mfe100[5], MfeA[5], Mpe200[5], Mne200[5], Mfe110[5], Mfe101[5], Mpl200[5],
        Mnr200[5], Mpl110[5], Mnr110[5], Mpl101[5], Mnr101[5], MplA1[5], MnrA1[5];
    mom(Mpl100, lp, lv, lw, l.e2, 1, 0, 0);
    mom(Mnr100, rn, rv, rw, r.e2, 1, 0, 0);
    mom(Mfe100, ef, ev, ew, e.e2, 1, 0, 0);
    mom_a(MfeA, ef, ev, ew, e.e2, e.e4, 1, 0, 0, A_);
    mom_a(Mpe200, ep, ev, ew, e.e2, e.e4, 2, 0, 0, ax_l);
    mom_a(Mne200, en, ev, ew, e.e2, e.e4, 2, 0, 0, ax_r);
    mom_a(Mfe110, ef, ev, ew, e.e2, e.e4, 1, 1, 0, ay_);
    mom_a(Mfe101, ef, ev, ew, e.e2, e.e4, 1, 0, 1, az_);
    mom_a(Mpl200, lp, lv, lw, l.e2, l.e4, 2, 0, 0, axl);
    mom_a(Mnr200, rn, rv, rw, r.e2, r.e4, 2, 0, 0, axr);
    mom_a(Mpl110, lp, lv, lw, l.e2, l.e4, 1, 1, 0, ayl);
    mom_a(Mnr110, rn, rv, rw, r.e2, r.e4, 1, 1, 0, ayr);
    mom_a(Mpl101, lp, lv, lw, l.e2, l.e4, 1, 0, 1, azl);
    mom_a(Mnr101, rn, rv, rw, r.e2, r.e4, 1, 0, 1, azr);
    mom_a(MplA1, lp, lv, lw, l.e2, l.e4, 1, 0, 0, Al);
    mom_a(MnrA1, rn, rv, rw, r.e2, r.e4, 1, 0, 0, Ar);

    for (int i = 0; i < 5; i++) {
        F[i] = p0 * dene * Mfe100[i] + p1 * dene * MfeA[i] +
               p2 * dene * (Mpe200[i] + Mne200[i] + Mfe110[i] + Mfe101[i]) +
               p3 * (denl * Mpl100[i] + denr * Mnr100[i]) +
               p4 * (denl * (Mpl200[i] + Mpl110[i] + Mpl101[i]) + denr * (Mnr200[i] + Mnr110[i] + Mnr101[i])) +
               p5 * (denl * MplA1[i] + denr * MnrA1[i]);
    }
}

/* ---- WENO ---------------------------------------------------------------- */

/* WENO-SYMOO on one line (update.cpp:507-557): 4 candidate 3-point
 * stencils, ideal weights 1/20, 9/20, 9/20, 1/20, the downwind smoothness
 * indicator replaced by the largest of the four, alpha = gamma/(eps + IS). */
inline void symoo(const double c[6], double& wl, double& wr) {
    const double gam0 = 1.0 / 20.0, gam1 = 9.0 / 20.0, gam2 = 9.0 / 20.0, gam3 = 1.0 / 20.0;
    const double eps = ref_eps();
    const double LLL = c[0], LL = c[1], L = c[2], R = c[3], RR = c[4], RRR = c[5];
    double P0p = LLL / 3 - 7 * LL / 6 + 11 * L / 6;
    double P0n = RRR / 3 - 7 * RR / 6 + 11 * R / 6;
    double P1p = -LL / 6 + 5 * L / 6 + R / 3;
    double P1n = -RR / 6 + 5 * R / 6 + L / 3;
    double P2p = L / 3 + 5 * R / 6 - RR / 6;
    double P2n = R / 3 + 5 * L / 6 - LL / 6;
    double P3p = 11 * R / 6 - 7 * RR / 6 + RRR / 3;
    double P3n = 11 * L / 6 - 7 * LL / 6 + LLL / 3;
    auto sq = [](double x) { return std::pow(x, 2); };
    double IS0p = 13 * sq(LLL - 2 * LL + L) / 12 + sq(LLL - 4 * LL + 3 * L) / 4;
    double IS0n = 13 * sq(RRR - 2 * RR + R) / 12 + sq(RRR - 4 * RR + 3 * R) / 4;
    double IS1p = 13 * sq(LL - 2 * L + R) / 12 + sq(-LL + R) / 4;
    double IS1n = 13 * sq(RR - 2 * R + L) / 12 + sq(-RR + L) / 4;
    double IS2p = 13 * sq(L - 2 * R + RR) / 12 + sq(-3 * L + 4 * R - RR) / 4;
    double IS2n = 13 * sq(R - 2 * L + LL) / 12 + sq(-3 * R + 4 * L - LL) / 4;
    double IS3p = 13 * sq(R - 2 * RR + RRR) / 12 + sq(-5 * R + 8 * RR - 3 * RRR) / 4;
    double IS3n = 13 * sq(L - 2 * LL + LLL) / 12 + sq(-5 * L + 8 * LL - 3 * LLL) / 4;
    IS3p = std::max(std::max(std::max(IS0p, IS1p), IS2p), IS3p);
    IS3n = std::max(std::max(std::max(IS0n, IS1n), IS2n), IS3n);
    double A0p = gam0 * (1 / (eps + IS0p)), A0n = gam0 * (1 / (eps + IS0n));
    double A1p = gam1 * (1 / (eps + IS1p)), A1n = gam1 * (1 / (eps + IS1n));
    double A2p = gam2 * (1 / (eps + IS2p)), A2n = gam2 * (1 / (eps + IS2n));
    double A3p = gam3 * (1 / (eps + IS3p)), A3n = gam3 * (1 / (eps + IS3n));
    double Sp = A0p + A1p + A2p + A3p;
    double Sn = A0n + A1n + A2n + A3n;
    wl = (A0p / Sp * P0p + A1p / Sp * P1p + A2p / Sp * P2p + A3p / Sp * P3p);
    wr = (A0n / Sn * P0n + A1n / Sn * P1n + A2n / Sn * P2n + A3n / Sn * P3n);
}

/* WENO-JS, 3 stencils, d = 0.1/0.6/0.3, alpha = d/(eps+IS)^2
 * (update.cpp:1421-1468). */
inline void wenojs(const double c[6], double& wl, double& wr) {
    const double gam0 = 0.1, gam1 = 0.6, gam2 = 0.3;
    const double eps = ref_eps();
    const double LLL = c[0], LL = c[1], L = c[2], R = c[3], RR = c[4], RRR = c[5];
    double P0p = LLL / 3 - 7 * LL / 6 + 11 * L / 6;
    double P0n = 11 * R / 6 - 7 * RR / 6 + RRR / 3;
    double P1p = -LL / 6 + 5 * L / 6 + R / 3;
    double P1n = L / 3 + 5 * R / 6 - RR / 6;
    double P2p = L / 3 + 5 * R / 6 - RR / 6;
    double P2n = -LL / 6 + 5 * L / 6 + R / 3;
    auto sq = [](double x) { return std::pow(x, 2); };
    double IS0p = 13 * sq(LLL - 2 * LL + L) / 12 + sq(LLL - 4 * LL + 3 * L) / 4;
    double IS0n = 13 * sq(RRR - 2 * RR + R) / 12 + sq(RRR - 4 * RR + 3 * R) / 4;
    double IS1p = 13 * sq(LL - 2 * L + R) / 12 + sq(LL - R) / 4;
    double IS1n = 13 * sq(L - 2 * R + RR) / 12 + sq(L - RR) / 4;
    double IS2p = 13 * sq(L - 2 * R + RR) / 12 + sq(3 * L - 4 * R + RR) / 4;
    double IS2n = 13 * sq(LL - 2 * L + R) / 12 + sq(LL - 4 * L + 3 * R) / 4;
    double A0p = gam0 * sq(1 / (eps + IS0p)), A0n = gam0 * sq(1 / (eps + IS0n));
    double A1p = gam1 * sq(1 / (eps + IS1p)), A1n = gam1 * sq(1 / (eps + IS1n));
    double A2p = gam2 * sq(1 / (eps + IS2p)), A2n = gam2 * sq(1 / (eps + IS2n));
    double Sp = A0p + A1p + A2p;
    double Sn = A0n + A1n + A2n;
    wl = (A0p / Sp * P0p + A1p / Sp * P1p + A2p / Sp * P2p);
    wr = (A0n / Sn * P0n + A1n / Sn * P1n + A2n / Sn * P2n);
}

inline void weno_line(int recon, const double c[6], double& wl, double& wr) {
    if (recon == GKM_RECON_JS)
        wenojs(c, wl, wr);
    else
        symoo(c, wl, wr);
}

/* Cell temperature as ApplyWENO forms it (update.cpp:560-565). */
inline double cell_T(double gam, double R, const double W[5]) {
    return (gam - 1) * (W[4] / W[0] - 0.5 * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]) / (W[0] * W[0])) / R;
}

void weno_vec(const gkm_params& gp, const double cells[6][5], double Wl[5], double Wr[5]) {
    const int nlin = gp.energy_recon == GKM_ENERGY_E ? 5 : 4;
    for (int q = 0; q < nlin; q++) {
        double c[6];
        for (int s = 0; s < 6; s++) c[s] = cells[s][q];
        weno_line(gp.recon, c, Wl[q], Wr[q]);
    }
    if (gp.energy_recon == GKM_ENERGY_E) return;
    double T[6];
    for (int s = 0; s < 6; s++) T[s] = cell_T(gp.gamma, gp.R, cells[s]);
    double Tl, Tr;
    weno_line(gp.recon, T, Tl, Tr);
    Wl[4] = Wl[0] * gp.R * Tl / (gp.gamma - 1.0) + 0.5 * (Wl[1] * Wl[1] + Wl[2] * Wl[2] + Wl[3] * Wl[3]) / Wl[0];
    Wr[4] = Wr[0] * gp.R * Tr / (gp.gamma - 1.0) + 0.5 * (Wr[1] * Wr[1] + Wr[2] * Wr[2] + Wr[3] * Wr[3]) / Wr[0];
}

/* Viscosity law of the left cell (update.cpp:262-264, GPUupdate.cu:1327). */
inline double viscosity(const gkm_params& gp, double T) {
    if (gp.visc_law == GKM_VISC_POWER) return gp.mu0 * std::pow((T / gp.T0), gp.visc_exp);
    return gp.mu0 * std::pow((T / gp.T0), 1.5) * (gp.T0 + 110.4) / (T + 110.4);
}

/* Smooth-flow gas-kinetic flux of the GKM tree (GKM/CUDA GKM/src/GKMflux.cpp:12-280):
 * one Maxwellian at the interface average W0, slopes from the central
 * differences, the time slope from the compatibility condition, and
 *   F = rho0 ((p1 + p2) <u psi A> + p1 (<u^2 psi ax> + <uv psi ay> + <uw psi az>) + p0 <u psi>)
 * with p0 = dt, p1 = -tau dt, p2 = dt^2/2, tau = mu/p0 (:47, :269-278).  The
 * moments are the full-space ones of all three axes (:28-42). */
void gks_flux(const gkm_params& gp, const double W0[5], const double DWx[5], const double DWy[5],
              const double DWz[5], double mu, double dt, double F[5]) {
    State s;
    make_state(gp.gamma, gp.K, W0, s, false, false);
    const double den0 = s.den, lam0 = s.lam;
    const double* f0 = s.ax[0].fu;
    const double* f1 = s.ax[1].fu;
    const double* f2 = s.ax[2].fu;
    const double tau = mu / s.P;
    const double invPr = 1.0 / gp.Pr;
    double bx[5], by[5], bz[5], ax[5], ay[5], az[5];
    for (int i = 0; i < 5; i++) {
        bx[i] = DWx[i] / den0;
        by[i] = DWy[i] / den0;
        bz[i] = DWz[i] / den0;
    }
    slopesolver(gp.K, invPr, bx, s.U, lam0, ax);
    slopesolver(gp.K, invPr, by, s.U, lam0, ay);
    slopesolver(gp.K, invPr, bz, s.U, lam0, az);
    double Mu_ax[5], Mu2_ax[5], Muv_ay[5], Muw_az[5], Mv_ay[5], Mw_az[5], M[5], Mu_A[5], B[5], A[5];
    mom_a(Mu_ax, f0, f1, f2, s.e2, s.e4, 1, 0, 0, ax);
    mom_a(Mu2_ax, f0, f1, f2, s.e2, s.e4, 2, 0, 0, ax);
    mom_a(Muv_ay, f0, f1, f2, s.e2, s.e4, 1, 1, 0, ay);
    mom_a(Muw_az, f0, f1, f2, s.e2, s.e4, 1, 0, 1, az);
    mom_a(Mv_ay, f0, f1, f2, s.e2, s.e4, 0, 1, 0, ay);
    mom_a(Mw_az, f0, f1, f2, s.e2, s.e4, 0, 0, 1, az);
    mom(M, f0, f1, f2, s.e2, 1, 0, 0);
    for (int i = 0; i < 5; i++) B[i] = -Mu_ax[i] - Mv_ay[i] - Mw_az[i];
    slopesolver(gp.K, invPr, B, s.U, lam0, A);
    mom_a(Mu_A, f0, f1, f2, s.e2, s.e4, 1, 0, 0, A);
    const double C0 = den0;
    const double p0 = dt, p1 = -tau * dt, p2 = dt * dt / 2;
    for (int i = 0; i < 5; i++)
        F[i] = C0 * ((p1 + p2) * Mu_A[i] + p1 * Mu2_ax[i] + p1 * Muv_ay[i] + p1 * Muw_az[i] + p0 * M[i]);
}

/* ---- whole-box helpers ----------------------------------------------------- */

struct Box {
    int P[3];      /* padded extents along axis 0 (y), 1 (x), 2 (z) */
    long s[3];     /* strides */
    long n;
    double dx;
    explicit Box(const gkm_grid& g) {
        P[0] = g.ny + 6;
        P[1] = g.nx + 6;
        P[2] = g.nz + 6;
        s[2] = 1;
        s[1] = P[2];
        s[0] = (long)P[1] * P[2];
        n = (long)P[0] * P[1] * P[2];
        dx = g.dx;
    }
    long at(int j, int i, int k) const { return j * s[0] + i * s[1] + k; }
};

/* Per-sweep geometry: normal axis, tangential axes paired with the "y" and
 * "z" slots of the rotated flux, and the slot -> field map
 * (update.cpp:201-453). */
struct Sweep {
    int an, t1, t2;
    int slot[5];
};

Sweep make_sweep(int dir, int zbug) {
    Sweep s;
    if (dir == 0) { /* X: normal = axis 1; N/S = j+-1, F/B = k+-1 */
        s.an = 1; s.t1 = 0; s.t2 = 2;
        int m[5] = {0, 1, 2, 3, 4};
        std::memcpy(s.slot, m, sizeof m);
    } else if (dir == 1) { /* Y: normal = axis 0; N/S = i-+1 (sign folded), F/B = k+-1 */
        s.an = 0; s.t1 = 1; s.t2 = 2;
        int m[5] = {0, 2, 1, 3, 4};
        std::memcpy(s.slot, m, sizeof m);
    } else { /* Z: normal = axis 2; reference pairs N/S = i+-1, F/B = j+-1 */
        s.an = 2;
        if (zbug) { s.t1 = 1; s.t2 = 0; } else { s.t1 = 0; s.t2 = 1; }
        int m[5] = {0, 3, 2, 1, 4};
        std::memcpy(s.slot, m, sizeof m);
    }
    return s;
}

/* One sweep: WENO on the extended interface set (each interface once; the
 * reference recomputes the four transverse neighbours' lines, which gives
 * the same values), then the flux on the core set. */
void sweep(const Box& B, const gkm_params& gp, double* const W[5], double dt, int dir, double* const F[5],
           int nthreads) {
    const Sweep S = make_sweep(dir, gp.zbug_compat);
    const long sn = B.s[S.an], st1 = B.s[S.t1], st2 = B.s[S.t2];
    std::vector<double> cache((size_t)B.n * 10, 0.0);
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = 2;
        hi[a] = B.P[a] - 3; /* tangential: [2, P-3]; normal: [2, P-4] */
    }
    hi[S.an] = B.P[S.an] - 4;
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
    for (int j = lo[0]; j <= hi[0]; j++)
        for (int i = lo[1]; i <= hi[1]; i++)
            for (int k = lo[2]; k <= hi[2]; k++) {
                long c = B.at(j, i, k);
                double cells[6][5];
                for (int s = 0; s < 6; s++)
                    for (int q = 0; q < 5; q++) cells[s][q] = W[q][c + (s - 2) * sn];
                double* out = &cache[(size_t)c * 10];
                weno_vec(gp, cells, out, out + 5);
            }
    for (int a = 0; a < 3; a++) {
        lo[a] = 3;
        hi[a] = B.P[a] - 4;
    }
    lo[S.an] = 2;
    const double dx = B.dx;
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
    for (int j = lo[0]; j <= hi[0]; j++)
        for (int i = lo[1]; i <= hi[1]; i++)
            for (int k = lo[2]; k <= hi[2]; k++) {
                long c = B.at(j, i, k);
                double WL[5], WR[5], Wl[5], Wr[5], Dxl[5], Dxr[5], Dyl[5], Dyr[5], Dzl[5], Dzr[5];
                const double* own = &cache[(size_t)c * 10];
                const double* n1p = &cache[(size_t)(c + st1) * 10];
                const double* n1m = &cache[(size_t)(c - st1) * 10];
                const double* n2p = &cache[(size_t)(c + st2) * 10];
                const double* n2m = &cache[(size_t)(c - st2) * 10];
                for (int q = 0; q < 5; q++) {
                    WL[q] = W[q][c];
                    WR[q] = W[q][c + sn];
                    Wl[q] = own[q];
                    Wr[q] = own[5 + q];
                    Dxl[q] = 2 * (Wl[q] - WL[q]) / dx;
                    Dxr[q] = 2 * (WR[q] - Wr[q]) / dx;
                    Dyl[q] = 0.5 * (n1p[q] - n1m[q]) / dx;
                    Dyr[q] = 0.5 * (n1p[5 + q] - n1m[5 + q]) / dx;
                    Dzl[q] = 0.5 * (n2p[q] - n2m[q]) / dx;
                    Dzr[q] = 0.5 * (n2p[5 + q] - n2m[5 + q]) / dx;
                }
                /* left-cell collision time from the unrotated state (update.cpp:262-264, 882-891) */
                double deni = WL[0];
                double ui = WL[1] / deni, vi = WL[2] / deni, wi = WL[3] / deni;
                double Vsqr = (ui * ui + vi * vi + wi * wi);
                double Pi = (deni * (gp.gamma - 1)) * (WL[4] / deni - 0.5 * Vsqr);
                double Ti = Pi / (gp.R * deni);
                double tau = viscosity(gp, Ti) / Pi;
                double rWL[5], rWR[5], rWl[5], rWr[5], rxl[5], rxr[5], ryl[5], ryr[5], rzl[5], rzr[5];
                for (int s = 0; s < 5; s++) {
                    int q = S.slot[s];
                    rWL[s] = WL[q]; rWR[s] = WR[q]; rWl[s] = Wl[q]; rWr[s] = Wr[q];
                    rxl[s] = Dxl[q]; rxr[s] = Dxr[q]; ryl[s] = Dyl[q]; ryr[s] = Dyr[q];
                    rzl[s] = Dzl[q]; rzr[s] = Dzr[q];
                }
                double Fs[5];
                bgk_flux(gp, rWL, rWR, rWl, rWr, rxl, rxr, ryl, ryr, rzl, rzr, tau, dt, dx, Fs);
                for (int s = 0; s < 5; s++) F[S.slot[s]][c] = Fs[s];
            }
}

/* One sweep of the GKM tree's scheme (GKM/CUDA GKM/src/CPUupdate.cpp:29-210):
 * W0 = (W_i + W_i+1)/2, DWx = (W_i+1 - W_i)/dx, DWy / DWz = half the difference
 * of the interface averages of the two neighbour rows along the tangential
 * axes paired with the y / z slots (the same pairing, Z-sweep quirk included,
 * as the weno sweeps), mu(T) of the interface average (:62-65); core set:
 * normal [2, P-4], tangential [3, P-4]. */
void sweep_gks(const Box& B, const gkm_params& gp, double* const W[5], double dt, int dir, double* const F[5],
               int nthreads) {
    const Sweep S = make_sweep(dir, gp.zbug_compat);
    const long sn = B.s[S.an], st1 = B.s[S.t1], st2 = B.s[S.t2];
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = 3;
        hi[a] = B.P[a] - 4;
    }
    lo[S.an] = 2;
    const double dx = B.dx;
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
    for (int j = lo[0]; j <= hi[0]; j++)
        for (int i = lo[1]; i <= hi[1]; i++)
            for (int k = lo[2]; k <= hi[2]; k++) {
                const long c = B.at(j, i, k);
                double W0[5], DWx[5], DWy[5], DWz[5];
                for (int q = 0; q < 5; q++) {
                    const double* w = W[q];
                    W0[q] = .5 * (w[c + sn] + w[c]);
                    const double W0N = .5 * (w[c + st1 + sn] + w[c + st1]);
                    const double W0S = .5 * (w[c - st1 + sn] + w[c - st1]);
                    const double W0F = .5 * (w[c + st2 + sn] + w[c + st2]);
                    const double W0B = .5 * (w[c - st2 + sn] + w[c - st2]);
                    DWx[q] = (w[c + sn] - w[c]) / dx;
                    DWy[q] = 0.5 * (W0N - W0S) / dx;
                    DWz[q] = 0.5 * (W0F - W0B) / dx;
                }
                const double deni = W0[0];
                const double ui = W0[1] / deni, vi = W0[2] / deni, wi = W0[3] / deni;
                const double Vsqr = (ui * ui + vi * vi + wi * wi);
                const double Pi = (deni * (gp.gamma - 1)) * (W0[4] / deni - .5 * Vsqr);
                const double mu = viscosity(gp, Pi / (gp.R * deni));
                double r0[5], rx[5], ry[5], rz[5];
                for (int s = 0; s < 5; s++) {
                    const int q = S.slot[s];
                    r0[s] = W0[q]; rx[s] = DWx[q]; ry[s] = DWy[q]; rz[s] = DWz[q];
                }
                double Fs[5];
                gks_flux(gp, r0, rx, ry, rz, mu, dt, Fs);
                for (int s = 0; s < 5; s++) F[S.slot[s]][c] = Fs[s];
            }
}

void fill_ghosts(const Box& B, double* const W[5]) {
    /* per axis, over the full padded extent of the others (FlowField.cpp:115-275) */
    for (int a = 0; a < 3; a++) {
        const int Pa = B.P[a];
        const int b = (a + 1) % 3, c = (a + 2) % 3;
        for (int u = 0; u < B.P[b]; u++)
            for (int v = 0; v < B.P[c]; v++) {
                long base = u * B.s[b] + v * B.s[c];
                for (int q = 0; q < 5; q++) {
                    double* w = W[q];
                    for (int g = 0; g < 3; g++) {
                        w[base + g * B.s[a]] = w[base + (Pa - 6 + g) * B.s[a]];
                    }
                    for (int g = 0; g < 3; g++) {
                        w[base + (Pa - 3 + g) * B.s[a]] = w[base + (3 + g) * B.s[a]];
                    }
                }
            }
    }
}

void update(const Box& B, double* const W[5], double* const F[15]) {
    const long s0 = B.s[0], s1 = B.s[1];
    for (int j = 3; j < B.P[0] - 3; j++)
        for (int i = 3; i < B.P[1] - 3; i++)
            for (int k = 3; k < B.P[2] - 3; k++) {
                long c = B.at(j, i, k);
                for (int q = 0; q < 5; q++) {
                    const double *Fx = F[q], *Fy = F[5 + q], *Fz = F[10 + q];
                    W[q][c] -= (1 / B.dx) * (Fx[c] - Fx[c - s1] + Fy[c] - Fy[c - s0] + Fz[c] - Fz[c - 1]);
                }
            }
}

}  // namespace

extern "C" {

void orc_flux(const gkm_params* p, const double WL[5], const double WR[5], const double Wl[5], const double Wr[5],
              const double DWxl[5], const double DWxr[5], const double DWyl[5], const double DWyr[5],
              const double DWzl[5], const double DWzr[5], double tau, double dt, double dx, double F[5]) {
    bgk_flux(*p, WL, WR, Wl, Wr, DWxl, DWxr, DWyl, DWyr, DWzl, DWzr, tau, dt, dx, F);
}

void orc_gks_flux(const gkm_params* p, const double W0[5], const double DWx[5], const double DWy[5],
                  const double DWz[5], double mu, double dt, double F[5]) {
    gks_flux(*p, W0, DWx, DWy, DWz, mu, dt, F);
}

void orc_weno(const gkm_params* p, const double cells[6][5], double Wl[5], double Wr[5]) {
    weno_vec(*p, cells, Wl, Wr);
}

void orc_weno_scalar(int recon, const double c[6], double* wl, double* wr) {
    weno_line(recon, c, *wl, *wr);
}

void orc_moment_tables(double U, double lam, int K, double half_pos[7], double half_neg[7], double full[7],
                       double* e2, double* e4) {
    Axis a;
    axis_tables(U, lam, a, true, true);
    for (int i = 0; i < 7; i++) {
        half_pos[i] = a.hp[i];
        half_neg[i] = a.hn[i];
        full[i] = a.fu[i];
    }
    *e2 = K / (2 * lam);
    *e4 = 3 * K / (4 * lam * lam) + K * (K - 1) / (4 * lam * lam);
}

void orc_slopesolver(const gkm_params* p, const double b[5], const double U[3], double lam, double a[5]) {
    slopesolver(p->K, 1 / p->Pr, b, U, lam, a);
}

void orc_fill_ghosts(const gkm_grid* g, double* const W[5]) { fill_ghosts(Box(*g), W); }

void orc_fluxes(const gkm_grid* g, const gkm_params* p, double* const W[5], double dt, double* const F[15],
                int nthreads) {
    Box B(*g);
    for (int dir = 0; dir < 3; dir++) {
        if (p->flux_scheme == GKM_SCHEME_GKS) sweep_gks(B, *p, W, dt, dir, F + 5 * dir, nthreads);
        else sweep(B, *p, W, dt, dir, F + 5 * dir, nthreads);
    }
}

void orc_step(const gkm_grid* g, const gkm_params* p, double* const W[5], double dt, int nsteps, int nthreads) {
    Box B(*g);
    std::vector<double> store((size_t)B.n * 15, 0.0);
    double* F[15];
    for (int q = 0; q < 15; q++) F[q] = &store[(size_t)q * B.n];
    for (int s = 0; s < nsteps; s++) {
        fill_ghosts(B, W);
        for (int dir = 0; dir < 3; dir++) {
            if (p->flux_scheme == GKM_SCHEME_GKS) sweep_gks(B, *p, W, dt, dir, F + 5 * dir, nthreads);
            else sweep(B, *p, W, dt, dir, F + 5 * dir, nthreads);
        }
        update(B, W, F);
    }
    fill_ghosts(B, W);
}

double orc_timestep(const gkm_grid* g, const gkm_params* p, double* const W[5], double cfl) {
    Box B(*g);
    double mx = 0.0;
    for (int j = 3; j < B.P[0] - 3; j++)
        for (int i = 3; i < B.P[1] - 3; i++)
            for (int k = 3; k < B.P[2] - 3; k++) {
                long c = B.at(j, i, k);
                double deni = W[0][c];
                double ui = W[1][c] / deni, vi = W[2][c] / deni, wi = W[3][c] / deni;
                double Vsqr = (ui * ui + vi * vi + wi * wi);
                double Pi = (deni * (p->gamma - 1)) * (W[4][c] / deni - 0.5 * Vsqr);
                double val = std::sqrt(Vsqr) + std::sqrt(p->gamma * Pi / deni);
                if (mx < val) mx = val;
            }
    return cfl * B.dx / mx;
}

void orc_stats(const gkm_grid* g, const gkm_params* p, double* const Win[5], gkm_stats_t* out) {
    Box B(*g);
    std::vector<double> store((size_t)B.n * 5);
    double* W[5];
    for (int q = 0; q < 5; q++) {
        W[q] = &store[(size_t)q * B.n];
        std::memcpy(W[q], Win[q], sizeof(double) * B.n);
    }
    fill_ghosts(B, W);
    const double gam = p->gamma, R = p->R;
    const double N = (double)g->nx * g->ny * g->nz;
    gkm_stats_t s;
    std::memset(&s, 0, sizeof s);
    auto vel = [&](long c, double u[3]) {
        u[0] = W[1][c] / W[0][c];
        u[1] = W[2][c] / W[0][c];
        u[2] = W[3][c] / W[0][c];
    };
    /* pass 1: means, sums, max */
    double sum_rho = 0, sum_p = 0, sum_T = 0, sum_v = 0, sum_c = 0, sum_q2 = 0, sum_tke = 0;
    double mx = 0;
    for (int j = 3; j < B.P[0] - 3; j++)
        for (int i = 3; i < B.P[1] - 3; i++)
            for (int k = 3; k < B.P[2] - 3; k++) {
                long c = B.at(j, i, k);
                double rho = W[0][c], u[3];
                vel(c, u);
                double q2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
                double P = (rho * (gam - 1)) * (W[4][c] / rho - 0.5 * q2);
                double T = P / (R * rho);
                double cs = std::sqrt(gam * P / rho);
                sum_rho += rho; sum_p += P; sum_T += T; sum_v += 1 / rho; sum_c += cs;
                sum_q2 += q2; sum_tke += rho * q2;
                mx = std::max(mx, std::sqrt(q2) + cs);
                s.mass += W[0][c]; s.mom_x += W[1][c]; s.mom_y += W[2][c]; s.mom_z += W[3][c]; s.energy += W[4][c];
                /* velocity gradients, 2nd-order central differences */
                double up[3], um[3], d[3][3];
                const long st[3] = {B.s[1], B.s[0], B.s[2]}; /* d/dx, d/dy, d/dz */
                for (int a = 0; a < 3; a++) {
                    vel(c + st[a], up);
                    vel(c - st[a], um);
                    for (int comp = 0; comp < 3; comp++) d[comp][a] = (up[comp] - um[comp]) / (2 * B.dx);
                }
                double mu = viscosity(*p, T);
                double dudx = d[0][0], dudy = d[0][1], dudz = d[0][2];
                double dvdx = d[1][0], dvdy = d[1][1], dvdz = d[1][2];
                double dwdx = d[2][0], dwdy = d[2][1], dwdz = d[2][2];
                double ox = dwdy - dvdz, oy = dudz - dwdx, oz = dvdx - dudy;
                double om2 = ox * ox + oy * oy + oz * oz;
                double th = dudx + dvdy + dwdz;
                double gg = 0;
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) gg += d[a][b] * d[a][b];
                s.enstrophy += om2;
                s.dil2 += th * th;
                s.diss_sol += mu * om2;
                s.diss_dil += mu * th * th;
                s.diss_ref += mu * gg;
                s.sijsij += dudx * dudx + 0.5 * (dudy + dvdx) * (dudy + dvdx) + 0.5 * (dudz + dwdx) * (dudz + dwdx) +
                            0.5 * (dwdy + dvdz) * (dwdy + dvdz) + dvdy * dvdy + dwdz * dwdz;
                s.dudx2 += dudx * dudx; s.dudx3 += dudx * dudx * dudx; s.dudx4 += dudx * dudx * dudx * dudx;
                s.dvdy2 += dvdy * dvdy; s.dvdy3 += dvdy * dvdy * dvdy; s.dvdy4 += dvdy * dvdy * dvdy * dvdy;
                s.dwdz2 += dwdz * dwdz; s.dwdz3 += dwdz * dwdz * dwdz; s.dwdz4 += dwdz * dwdz * dwdz * dwdz;
            }
    s.rho_mean = sum_rho / N; s.p_mean = sum_p / N; s.T_mean = sum_T / N; s.v_mean = sum_v / N;
    s.c_mean = sum_c / N; s.q2 = sum_q2 / N; s.tke = sum_tke / N; s.ek = 0.5 * s.tke;
    s.mt = std::sqrt(s.q2) / s.c_mean;
    s.umax_plus_c = mx;
    s.enstrophy /= N; s.dil2 /= N; s.diss_sol /= N; s.diss_dil = (4.0 / 3.0) * s.diss_dil / N;
    s.diss_ref /= N; s.sijsij /= N;
    s.dudx2 /= N; s.dudx3 /= N; s.dudx4 /= N; s.dvdy2 /= N; s.dvdy3 /= N; s.dvdy4 /= N;
    s.dwdz2 /= N; s.dwdz3 /= N; s.dwdz4 /= N;
    s.skew_u = s.dudx3 / std::pow(s.dudx2, 1.5); s.skew_v = s.dvdy3 / std::pow(s.dvdy2, 1.5);
    s.skew_w = s.dwdz3 / std::pow(s.dwdz2, 1.5);
    s.kurt_u = s.dudx4 / (s.dudx2 * s.dudx2); s.kurt_v = s.dvdy4 / (s.dvdy2 * s.dvdy2);
    s.kurt_w = s.dwdz4 / (s.dwdz2 * s.dwdz2);
    /* pass 2: rms about the means */
    double vr = 0, vp = 0, vT = 0, vv = 0;
    for (int j = 3; j < B.P[0] - 3; j++)
        for (int i = 3; i < B.P[1] - 3; i++)
            for (int k = 3; k < B.P[2] - 3; k++) {
                long c = B.at(j, i, k);
                double rho = W[0][c], u[3];
                vel(c, u);
                double q2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
                double P = (rho * (gam - 1)) * (W[4][c] / rho - 0.5 * q2);
                double T = P / (R * rho);
                vr += (rho - s.rho_mean) * (rho - s.rho_mean);
                vp += (P - s.p_mean) * (P - s.p_mean);
                vT += (T - s.T_mean) * (T - s.T_mean);
                vv += (1 / rho - s.v_mean) * (1 / rho - s.v_mean);
            }
    s.rho_rms = std::sqrt(vr / N); s.p_rms = std::sqrt(vp / N); s.T_rms = std::sqrt(vT / N);
    s.v_rms = std::sqrt(vv / N);
    *out = s;
}


/* Natural cubic spline slopes by the Thomas algorithm on the slope form of the
 * continuity conditions (second derivative zero at both ends):
 *   2 D_0 + D_1 = 3 (f_1 - f_0)/h
 *   D_{i-1} + 4 D_i + D_{i+1} = 3 (f_{i+1} - f_{i-1})/h,  0 < i < m-1
 *   D_{m-2} + 2 D_{m-1} = 3 (f_{m-1} - f_{m-2})/h
 * At grid nodes the gradient of the tensor-product interpolating B-spline with
 * natural ends that einspline builds per axis (get_bsplinep,
 * weno/src/Interpolation.cpp:3-31) equals these 1-D slopes along the line. */
void orc_spline_slopes(const double* f, long stride, int m, double h, double* D) {
    std::vector<double> cp((size_t)m), y((size_t)m);
    for (int i = 0; i < m; i++) {
        double r;
        if (i == 0) r = 3.0 * (f[stride] - f[0]) / h;
        else if (i == m - 1) r = 3.0 * (f[(long)(m - 1) * stride] - f[(long)(m - 2) * stride]) / h;
        else r = 3.0 * (f[(long)(i + 1) * stride] - f[(long)(i - 1) * stride]) / h;
        const double b = (i == 0 || i == m - 1) ? 2.0 : 4.0;
        const double den = i == 0 ? b : b - cp[i - 1];
        cp[i] = 1.0 / den;
        y[i] = i == 0 ? r / den : (r - y[i - 1]) / den;
    }
    D[m - 1] = y[m - 1];
    for (int i = m - 2; i >= 0; i--) D[i] = y[i] - cp[i] * D[i + 1];
}

/* PrintStats (weno/src/update.cpp:638-784): first pass means of p, T, 1/rho
 * and the gradient sums (:658-707), second pass rms about the means
 * (:751-763); velocities u = W1/rho etc. over the padded segment
 * (generateSplines, :1215-1231); gradients from the natural spline. */
void orc_print_stats(const gkm_grid* g, const gkm_params* p, double* const Win[5], gkm_print_stats_t* out) {
    Box B(*g);
    std::vector<double> store((size_t)B.n * 5);
    double* W[5];
    for (int q = 0; q < 5; q++) {
        W[q] = &store[(size_t)q * B.n];
        std::memcpy(W[q], Win[q], sizeof(double) * B.n);
    }
    fill_ghosts(B, W);
    const double gam = p->gamma, R = p->R;
    const double N = (double)g->nx * g->ny * g->nz;
    /* velocity fields of the whole padded box */
    std::vector<double> vel((size_t)B.n * 3);
    for (long c = 0; c < B.n; c++) {
        const double den = W[0][c];
        vel[c] = W[1][c] / den;
        vel[B.n + c] = W[2][c] / den;
        vel[2 * B.n + c] = W[3][c] / den;
    }
    /* grad[(3 * a + comp) * n + c]: slope of velocity comp along storage axis a
     * (0 = y, 1 = x, 2 = z), on lines through interior nodes */
    std::vector<double> grad((size_t)B.n * 9, 0.0);
    std::vector<double> line;
    for (int a = 0; a < 3; a++) {
        const int b1 = a == 0 ? 1 : 0, b2 = a == 2 ? 1 : 2;
        const int m = B.P[a];
        line.resize((size_t)m);
        for (int u1 = 3; u1 < B.P[b1] - 3; u1++)
            for (int u2 = 3; u2 < B.P[b2] - 3; u2++) {
                const long base = u1 * B.s[b1] + u2 * B.s[b2];
                for (int comp = 0; comp < 3; comp++) {
                    const double* f = &vel[(size_t)comp * B.n + base];
                    orc_spline_slopes(f, B.s[a], m, B.dx, line.data());
                    double* G = &grad[(size_t)(3 * a + comp) * B.n + base];
                    for (int t = 0; t < m; t++) G[t * B.s[a]] = line[t];
                }
            }
    }
    auto gr = [&](int a, int comp, long c) { return grad[(size_t)(3 * a + comp) * B.n + c]; };
    double sTKE = 0, sDiss = 0, sS = 0, sP = 0, sT = 0, sV = 0;
    double u2 = 0, u3 = 0, u4 = 0, v2 = 0, v3 = 0, v4 = 0, w2 = 0, w3 = 0, w4 = 0;
    for (int j = 3; j < B.P[0] - 3; j++)
        for (int i = 3; i < B.P[1] - 3; i++)
            for (int k = 3; k < B.P[2] - 3; k++) {
                const long c = B.at(j, i, k);
                const double deni = W[0][c];
                const double ui = W[1][c] / deni, vi = W[2][c] / deni, wi = W[3][c] / deni;
                const double Vsqr = (ui * ui + vi * vi + wi * wi);
                const double Pi = (deni * (gam - 1)) * (W[4][c] / deni - 0.5 * Vsqr);
                const double Ti = Pi / (R * deni);
                sP += Pi;
                sT += Ti;
                sV += 1 / deni;
                const double dudx = gr(1, 0, c), dudy = gr(0, 0, c), dudz = gr(2, 0, c);
                const double dvdx = gr(1, 1, c), dvdy = gr(0, 1, c), dvdz = gr(2, 1, c);
                const double dwdx = gr(1, 2, c), dwdy = gr(0, 2, c), dwdz = gr(2, 2, c);
                const double mu = viscosity(*p, Ti);
                sDiss += mu * (dudx * dudx + dudy * dudy + dudz * dudz + dvdx * dvdx + dvdy * dvdy + dvdz * dvdz +
                               dwdx * dwdx + dwdy * dwdy + dwdz * dwdz);
                sTKE += deni * (ui * ui + vi * vi + wi * wi);
                sS += (dudx * dudx + 0.5 * std::pow(dudy + dvdx, 2) + 0.5 * std::pow(dudz + dwdx, 2) +
                       0.5 * std::pow(dwdy + dvdz, 2) + dvdy * dvdy + dwdz * dwdz);
                u2 += dudx * dudx; u3 += dudx * dudx * dudx; u4 += dudx * dudx * dudx * dudx;
                v2 += dvdy * dvdy; v3 += dvdy * dvdy * dvdy; v4 += dvdy * dvdy * dvdy * dvdy;
                w2 += dwdz * dwdz; w3 += dwdz * dwdz * dwdz; w4 += dwdz * dwdz * dwdz * dwdz;
            }
    gkm_print_stats_t s;
    std::memset(&s, 0, sizeof s);
    s.tke = sTKE / N;
    s.diss = sDiss / N;
    s.sijsij = sS / N;
    u2 /= N; u3 /= N; u4 /= N; v2 /= N; v3 /= N; v4 /= N; w2 /= N; w3 /= N; w4 /= N;
    s.skew_u = u3 / std::pow(u2, 1.5); s.skew_v = v3 / std::pow(v2, 1.5); s.skew_w = w3 / std::pow(w2, 1.5);
    s.kurt_u = u4 / std::pow(u2, 2); s.kurt_v = v4 / std::pow(v2, 2); s.kurt_w = w4 / std::pow(w2, 2);
    s.dudx2 = u2; s.dvdy2 = v2; s.dwdz2 = w2;
    s.p_mean = sP / N; s.T_mean = sT / N; s.v_mean = sV / N;
    double rp = 0, rT = 0, rV = 0;
    for (int j = 3; j < B.P[0] - 3; j++)
        for (int i = 3; i < B.P[1] - 3; i++)
            for (int k = 3; k < B.P[2] - 3; k++) {
                const long c = B.at(j, i, k);
                const double deni = W[0][c];
                const double ui = W[1][c] / deni, vi = W[2][c] / deni, wi = W[3][c] / deni;
                const double Vsqr = (ui * ui + vi * vi + wi * wi);
                const double Pi = (deni * (gam - 1)) * (W[4][c] / deni - 0.5 * Vsqr);
                const double Ti = Pi / (R * deni);
                rp += std::pow((s.p_mean - Pi), 2);
                rT += std::pow((s.T_mean - Ti), 2);
                rV += std::pow((s.v_mean - 1 / deni), 2);
            }
    s.p_rms = std::sqrt(rp / N);
    s.T_rms = std::sqrt(rT / N);
    s.v_rms = std::sqrt(rV / N);
    *out = s;
}

}  // extern "C"
