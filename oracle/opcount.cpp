/*
 * opcount.cpp — floating-point operation count of the oracle (FP64 restatement
 * of the reference CPU step), per interface and per cell-update.
 *
 * TEST INFRASTRUCTURE ONLY (measurement): built as oracle/_build/libopcount.so
 * and read by tests/test_opcount.py and bench.py.  Nothing in the product
 * links it.
 *
 * gkm_oracle.cpp is compiled a second time with every `double` replaced by
 * CReal, a value type that does the same FP64 arithmetic and counts it:
 *   add/sub, mul, div, sqrt   one operation each (pow(x, 2) counts as a mul,
 *                             as the reference's pow(.., 2) computes x*x);
 *   exp, erfc, pow            counted apart (transcendentals);
 *   fabs, max, min, compares  not counted.
 * The counts follow the reference's own expression structure (the oracle keeps
 * its order of operations, GKMflux.cpp:11-308, update.cpp:201-469, 502-625,
 * 1421-1468), so they are the algorithmic work of the reference scheme with
 * each WENO line evaluated once per interface.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace opc {
struct Counts {
    uint64_t add, mul, div, sqrt_, exp_, erfc_, pow_;
};
thread_local Counts cnt = {0, 0, 0, 0, 0, 0, 0};
}  // namespace opc

struct CReal {
    double v;
    CReal() = default;
    CReal(double x) : v(x) {}  // NOLINT: implicit, as a double literal is
    explicit operator double() const { return v; }
    explicit operator int() const { return (int)v; }
    explicit operator long() const { return (long)v; }
    CReal& operator+=(CReal o) { opc::cnt.add++; v += o.v; return *this; }
    CReal& operator-=(CReal o) { opc::cnt.add++; v -= o.v; return *this; }
    CReal& operator*=(CReal o) { opc::cnt.mul++; v *= o.v; return *this; }
    CReal& operator/=(CReal o) { opc::cnt.div++; v /= o.v; return *this; }
    CReal operator-() const { return CReal(-v); }
    CReal operator+() const { return *this; }
};
inline CReal operator+(CReal a, CReal b) { opc::cnt.add++; return CReal(a.v + b.v); }
inline CReal operator-(CReal a, CReal b) { opc::cnt.add++; return CReal(a.v - b.v); }
inline CReal operator*(CReal a, CReal b) { opc::cnt.mul++; return CReal(a.v * b.v); }
inline CReal operator/(CReal a, CReal b) { opc::cnt.div++; return CReal(a.v / b.v); }
inline bool operator<(CReal a, CReal b) { return a.v < b.v; }
inline bool operator>(CReal a, CReal b) { return a.v > b.v; }
inline bool operator<=(CReal a, CReal b) { return a.v <= b.v; }
inline bool operator>=(CReal a, CReal b) { return a.v >= b.v; }
inline bool operator==(CReal a, CReal b) { return a.v == b.v; }
inline bool operator!=(CReal a, CReal b) { return a.v != b.v; }

namespace std {
inline CReal sqrt(CReal a) { opc::cnt.sqrt_++; return CReal(std::sqrt(a.v)); }
inline CReal exp(CReal a) { opc::cnt.exp_++; return CReal(std::exp(a.v)); }
inline CReal erfc(CReal a) { opc::cnt.erfc_++; return CReal(std::erfc(a.v)); }
inline CReal pow(CReal a, CReal b) {
    if (b.v == 2.0) {
        opc::cnt.mul++;
        return CReal(a.v * a.v);
    }
    opc::cnt.pow_++;
    return CReal(std::pow(a.v, b.v));
}
inline CReal pow(CReal a, int b) { return pow(a, CReal((double)b)); }
inline CReal abs(CReal a) { return CReal(std::fabs(a.v)); }
inline CReal fabs(CReal a) { return CReal(std::fabs(a.v)); }
inline CReal atan(CReal a) { return CReal(std::atan(a.v)); }
inline bool isfinite(CReal a) { return std::isfinite(a.v); }
}  // namespace std

#define double CReal
#include "gkm_oracle.cpp"
#undef double

namespace {
opc::Counts take() {
    opc::Counts c = opc::cnt;
    opc::cnt = {0, 0, 0, 0, 0, 0, 0};
    return c;
}
void put(const opc::Counts& c, unsigned long long* out) {
    out[0] = c.add;
    out[1] = c.mul;
    out[2] = c.div;
    out[3] = c.sqrt_;
    out[4] = c.exp_;
    out[5] = c.erfc_;
    out[6] = c.pow_;
}
}  // namespace

extern "C" {

/* Counts of one interface's WENO (weno_vec: the five fields of the six-cell
 * stencil, temperature mode or not), of one interface's flux (the sweep's
 * slope differences, the left-cell collision time and bgk_flux) and of one
 * cell's update, for the parameters given; the inputs are a smooth, physical
 * state (the counts do not depend on the values).  out[3][7]: add, mul, div,
 * sqrt, exp, erfc, pow for each of the three. */
void orc_opcount(const gkm_params* pin, unsigned long long* out) {
    gkm_params p;
    std::memcpy(&p, pin, sizeof p);
    (void)take();
    /* WENO of one interface */
    CReal cells[6][5];
    for (int s = 0; s < 6; s++) {
        const CReal rho = 1.0 + 0.01 * s, u = 0.1 - 0.02 * s, v = 0.05 * s, w = -0.03, pr = 0.7 + 0.01 * s;
        cells[s][0] = rho.v;
        cells[s][1] = rho.v * u.v;
        cells[s][2] = rho.v * v.v;
        cells[s][3] = rho.v * w.v;
        cells[s][4] = pr.v / (p.gamma.v - 1) + 0.5 * rho.v * (u.v * u.v + v.v * v.v + w.v * w.v);
    }
    (void)take();
    CReal Wl[5], Wr[5];
    weno_vec(p, cells, Wl, Wr);
    put(take(), out);
    /* flux of one interface, as sweep() forms its inputs */
    CReal WL[5], WR[5], Dn1[5], Dn2[5], Dt1[5], Dt2[5];
    for (int q = 0; q < 5; q++) {
        WL[q] = cells[2][q].v;
        WR[q] = cells[3][q].v;
        Dn1[q] = 0.01 * q;
        Dn2[q] = -0.02 * q;
        Dt1[q] = 0.015 * q;
        Dt2[q] = 0.005 * q;
    }
    const CReal dx = 0.05, dt = 1e-3;
    (void)take();
    CReal Dxl[5], Dxr[5], Dyl[5], Dyr[5], Dzl[5], Dzr[5];
    for (int q = 0; q < 5; q++) {
        Dxl[q] = 2 * (Wl[q] - WL[q]) / dx;
        Dxr[q] = 2 * (WR[q] - Wr[q]) / dx;
        Dyl[q] = 0.5 * (Dn1[q] - Dn2[q]) / dx;
        Dyr[q] = 0.5 * (Dn1[q] - Dn2[q]) / dx;
        Dzl[q] = 0.5 * (Dt1[q] - Dt2[q]) / dx;
        Dzr[q] = 0.5 * (Dt1[q] - Dt2[q]) / dx;
    }
    CReal deni = WL[0];
    CReal ui = WL[1] / deni, vi = WL[2] / deni, wi = WL[3] / deni;
    CReal Vsqr = (ui * ui + vi * vi + wi * wi);
    CReal Pi = (deni * (p.gamma - 1)) * (WL[4] / deni - 0.5 * Vsqr);
    CReal Ti = Pi / (p.R * deni);
    CReal tau = viscosity(p, Ti) / Pi;
    CReal F[5];
    bgk_flux(p, WL, WR, Wl, Wr, Dxl, Dxr, Dyl, Dyr, Dzl, Dzr, tau, dt, dx, F);
    put(take(), out + 7);
    /* update of the 27 cells of a 3^3 box (update.cpp:455-469), per cell */
    gkm_grid g;
    g.nx = g.ny = g.nz = 3;
    g.ghost = 3;
    g.dx = dx.v;
    Box B(g);
    std::vector<CReal> store((size_t)B.n * 20, CReal(0.5));
    CReal* Wp[5];
    CReal* Fp[15];
    for (int q = 0; q < 5; q++) Wp[q] = &store[(size_t)q * B.n];
    for (int q = 0; q < 15; q++) Fp[q] = &store[(size_t)(5 + q) * B.n];
    (void)take();
    update(B, Wp, Fp);
    opc::Counts u = take();
    u.add /= 27;
    u.mul /= 27;
    u.div /= 27;
    put(u, out + 14);
}

}  // extern "C"
