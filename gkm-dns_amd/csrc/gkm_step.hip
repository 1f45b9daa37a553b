// gkm_step.hip — kernels of one GKM-DNS time step on gfx950 and their
// launchers: periodic ghost fill, WENO reconstruction per sweep, BGK flux per
// sweep, conservative update, state split/merge for the mixed form.
#include <hip/hip_runtime.h>

#include "gkm_internal.h"

namespace gkm {

// ---------------------------------------------------------------------------
// Periodic ghost fill (replaces PeriodicBC / one-rank HaloTransfer,
// weno/src/FlowField.cpp:115-275, update.cpp:894-1054) fused with the per-cell
// primitives every sweep needs.  Every ghost cell takes the value of its
// periodic image in the interior, which is what the three full-extent axis
// passes of the reference produce, corners included.  Then, for every padded
// cell, from the FP64 state:
//   Tw  = (gam-1)(E/rho - |m|^2/(2 rho^2))/R   the WENO temperature input
//         (update.cpp:560-565)
//   tau = mu(Ti)/p, Ti = p/(R rho)            the left-cell collision time
//         (update.cpp:262-264, 882-891)
// so that the sweeps read them instead of recomputing them per interface.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double viscosity(const StepPar& par, double T) {
    if (par.visc_law == 1) return par.mu0 * pow(T / par.T0, par.visc_exp);
    return par.mu0 * (T / par.T0) * sqrt(T / par.T0) * (par.T0 + 110.4) / (T + 110.4);
}

template <int NPL, class T, class Real, class St>
__global__ void __launch_bounds__(256) k_ghost_prim(Geo g, Planes10<T> f, St st, Real* Tw, Real* tau, StepPar par) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)g.P0 * g.P1 * g.P2;
    if (t >= n) return;
    int k = (int)(t % g.P2);
    long r = t / g.P2;
    int i = (int)(r % g.P1);
    int j = (int)(r / g.P1);
    const bool gj = j < 3 || j >= g.ny + 3, gi = i < 3 || i >= g.nx + 3, gk = k < 3 || k >= g.nz + 3;
    long src = t;
    if (gj || gi || gk) {
        int js = j < 3 ? j + g.ny : (j >= g.ny + 3 ? j - g.ny : j);
        int is = i < 3 ? i + g.nx : (i >= g.nx + 3 ? i - g.nx : i);
        int ks = k < 3 ? k + g.nz : (k >= g.nz + 3 ? k - g.nz : k);
        src = (long)js * g.s0 + (long)is * g.s1 + ks;
#pragma unroll
        for (int q = 0; q < NPL; q++) f.p[q][t] = f.p[q][src];
    }
    if (!Tw) return;
    const double r0 = st.get(0, src), m1 = st.get(1, src), m2 = st.get(2, src), m3 = st.get(3, src),
                 E = st.get(4, src);
    Tw[t] = (Real)((par.gam - 1) * (E / r0 - 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / (r0 * r0)) / par.R);
    const double u = m1 / r0, v = m2 / r0, w = m3 / r0;
    const double P = (r0 * (par.gam - 1)) * (E / r0 - 0.5 * (u * u + v * v + w * w));
    const double Ti = P / (par.R * r0);
    tau[t] = (Real)(viscosity(par, Ti) / P);
}

// ---------------------------------------------------------------------------
// WENO reconstruction of one sweep on the extended interface set (normal
// index [2, P-4], tangential [2, P-3]): each interface once.  The reference
// evaluates the line of every transverse neighbour again inside each
// interface (5 lines per interface, update.cpp:247-251); the flux kernel
// reads the neighbours' values from here instead, which are the same numbers.
// ---------------------------------------------------------------------------
template <class Real, class St, int RECON>
__global__ void __launch_bounds__(256) k_recon(Geo g, Box3 b, long sn, St st, const Real* __restrict__ Tw,
                                               Planes10<Real> wi, StepPar par) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.count()) return;
    const long c = box_index(g, b, t);
    const Real eps = Real(1e-10);
    Real wl[5], wr[5];
    const int nlin = par.energy_recon == 1 ? 5 : 4;
    for (int q = 0; q < nlin; q++) {
        Real v[6];
#pragma unroll
        for (int s = 0; s < 6; s++) v[s] = (Real)st.cell(q, c + (s - 2) * sn);
        weno_line<RECON>(v, eps, wl[q], wr[q]);
    }
    if (par.energy_recon == 0) {
        // temperature reconstruction (update.cpp:560-621), cell temperatures from the plane
        Real Tc[6];
#pragma unroll
        for (int s = 0; s < 6; s++) Tc[s] = Tw[c + (s - 2) * sn];
        Real Tl, Tr;
        weno_line<RECON>(Tc, eps, Tl, Tr);
        const Real cR = Real(par.R / (par.gam - 1.0));
        wl[4] = wl[0] * cR * Tl + Real(0.5) * (wl[1] * wl[1] + wl[2] * wl[2] + wl[3] * wl[3]) * M<Real>::rcp(wl[0]);
        wr[4] = wr[0] * cR * Tr + Real(0.5) * (wr[1] * wr[1] + wr[2] * wr[2] + wr[3] * wr[3]) * M<Real>::rcp(wr[0]);
    }
#pragma unroll
    for (int q = 0; q < 5; q++) {
        wi.p[q][c] = wl[q];
        wi.p[5 + q][c] = wr[q];
    }
}

// slot -> field map of the rotated flux frame (update.cpp:340-352, 429-441)
template <int DIR> __device__ __forceinline__ int slot_field(int s) {
    if (DIR == 1) return s == 1 ? 2 : (s == 2 ? 1 : s);
    if (DIR == 2) return s == 1 ? 3 : (s == 3 ? 1 : s);
    return s;
}

// ---------------------------------------------------------------------------
// BGK flux of one sweep on the core interface set (normal [2, P-4],
// tangential [3, P-4]); writes the five flux planes of that direction.
// st1 / st2 are the strides of the tangential axes paired with the y / z
// slots of the rotated frame.
// ---------------------------------------------------------------------------
template <class Real, class St, int DIR>
__global__ void __launch_bounds__(256) k_flux(Geo g, Box3 b, long sn, long st1, long st2, St st,
                                              const Real* __restrict__ tau, Planes10<Real> wi, Planes5<Real> F,
                                              StepPar par) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.count()) return;
    const long c = box_index(g, b, t);
    Gas<Real> gas;
    gas.gam = (Real)par.gam;
    gas.invPr = (Real)par.invPr;
    gas.K = (Real)par.K;
    gas.kp3 = (Real)(par.K + 3);
    gas.ikp3 = (Real)(1.0 / (par.K + 3));
    const Real idx = Real(1.0 / par.dx);

    // collision time mu/p of the left cell (update.cpp:262-264), from the plane
    const Real tau0 = tau[c];
    Real WL[5], WR[5], Wl[5], Wr[5];
#pragma unroll
    for (int s = 0; s < 5; s++) {
        const int q = slot_field<DIR>(s);
        WL[s] = (Real)st.cell(q, c);
        WR[s] = (Real)st.cell(q, c + sn);
        Wl[s] = wi.p[q][c];
        Wr[s] = wi.p[5 + q][c];
    }
    auto slopes = [&](int side, Real Dx[5], Real Dy[5], Real Dz[5]) {
#pragma unroll
        for (int s = 0; s < 5; s++) {
            const int q = slot_field<DIR>(s);
            const int o = side == 0 ? q : 5 + q;
            Dx[s] = side == 0 ? 2 * (Wl[s] - WL[s]) * idx : 2 * (WR[s] - Wr[s]) * idx;
            Dy[s] = Real(0.5) * (wi.p[o][c + st1] - wi.p[o][c - st1]) * idx;
            Dz[s] = Real(0.5) * (wi.p[o][c + st2] - wi.p[o][c - st2]) * idx;
        }
    };
    Real Fs[5];
    bgk_flux(gas, WL, WR, Wl, Wr, slopes, tau0, (Real)par.dt, (Real)par.dx, Fs);
#pragma unroll
    for (int s = 0; s < 5; s++) F.p[slot_field<DIR>(s)][c] = Fs[s];
}

// ---------------------------------------------------------------------------
// Conservative update on the interior (update.cpp:455-469): the flux
// differences are summed in the reference's order.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_update_d(Geo g, Box3 b, StateD st, Planes15<double> F, double dx) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.count()) return;
    const long c = box_index(g, b, t);
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const double* Fx = F.p[q];
        const double* Fy = F.p[5 + q];
        const double* Fz = F.p[10 + q];
        st.w[q][c] -= (1 / dx) * (Fx[c] - Fx[c - g.s1] + Fy[c] - Fy[c - g.s0] + Fz[c] - Fz[c - 1]);
    }
}

// Mixed form: the FP32 flux divergence is added to the hi/lo pair with an
// error-free TwoSum followed by a Fast2Sum renormalisation.
__device__ __forceinline__ void two_sum_add(float& hi, float& lo, float d) {
    const float s = hi + d;
    const float bp = s - hi;
    const float e = (hi - (s - bp)) + (d - bp);
    const float l2 = lo + e;
    const float h2 = s + l2;
    lo = l2 - (h2 - s);
    hi = h2;
}

__global__ void __launch_bounds__(256) k_update_m(Geo g, Box3 b, StateM st, Planes15<float> F, float idx) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.count()) return;
    const long c = box_index(g, b, t);
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const float* Fx = F.p[q];
        const float* Fy = F.p[5 + q];
        const float* Fz = F.p[10 + q];
        const float R = Fx[c] - Fx[c - g.s1] + Fy[c] - Fy[c - g.s0] + Fz[c] - Fz[c - 1];
        float hi = st.hi[q][c], lo = st.lo[q][c];
        two_sum_add(hi, lo, -idx * R);
        st.hi[q][c] = hi;
        st.lo[q][c] = lo;
    }
}

// ---- split / merge of the mixed state --------------------------------------
__global__ void __launch_bounds__(256) k_split(long n, const double* src, float* hi, float* lo) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double v = src[t];
    const float h = (float)v;
    hi[t] = h;
    lo[t] = (float)(v - (double)h);
}

__global__ void __launch_bounds__(256) k_merge(long n, const float* hi, const float* lo, double* dst) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    dst[t] = (double)hi[t] + (double)lo[t];
}

// ---- point-wise entry points (device arithmetic checks) --------------------
template <class Real>
__global__ void k_flux_points(int n, const double* rec, double dx, double gam, double invPr, int K, double* out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double* x = rec + (long)t * 52;
    Gas<Real> gas;
    gas.gam = (Real)gam;
    gas.invPr = (Real)invPr;
    gas.K = (Real)K;
    gas.kp3 = (Real)(K + 3);
    gas.ikp3 = (Real)(1.0 / (K + 3));
    Real v[10][5];
    for (int a = 0; a < 10; a++)
        for (int q = 0; q < 5; q++) v[a][q] = (Real)x[a * 5 + q];
    auto slopes = [&](int side, Real Dx[5], Real Dy[5], Real Dz[5]) {
        for (int q = 0; q < 5; q++) {
            Dx[q] = v[4 + side][q];
            Dy[q] = v[6 + side][q];
            Dz[q] = v[8 + side][q];
        }
    };
    Real F[5];
    bgk_flux(gas, v[0], v[1], v[2], v[3], slopes, (Real)x[50], (Real)x[51], (Real)dx, F);
    for (int q = 0; q < 5; q++) out[(long)t * 5 + q] = (double)F[q];
}

template <class Real, int RECON>
__global__ void k_weno_points(int n, const double* rec, double gam, double R, int energy_recon, double* out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double* x = rec + (long)t * 30;
    const Real eps = Real(1e-10);
    Real wl[5], wr[5];
    const int nlin = energy_recon == 1 ? 5 : 4;
    for (int q = 0; q < nlin; q++) {
        Real c6[6];
        for (int s = 0; s < 6; s++) c6[s] = (Real)x[s * 5 + q];
        weno_line<RECON>(c6, eps, wl[q], wr[q]);
    }
    if (energy_recon == 0) {
        Real Tc[6];
        for (int s = 0; s < 6; s++) {
            const double* W = x + s * 5;
            Tc[s] = (Real)((gam - 1) * (W[4] / W[0] - 0.5 * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]) / (W[0] * W[0])) / R);
        }
        Real Tl, Tr;
        weno_line<RECON>(Tc, eps, Tl, Tr);
        const Real cR = Real(R / (gam - 1.0));
        wl[4] = wl[0] * cR * Tl + Real(0.5) * (wl[1] * wl[1] + wl[2] * wl[2] + wl[3] * wl[3]) * M<Real>::rcp(wl[0]);
        wr[4] = wr[0] * cR * Tr + Real(0.5) * (wr[1] * wr[1] + wr[2] * wr[2] + wr[3] * wr[3]) * M<Real>::rcp(wr[0]);
    }
    for (int q = 0; q < 5; q++) {
        out[(long)t * 10 + q] = (double)wl[q];
        out[(long)t * 10 + 5 + q] = (double)wr[q];
    }
}

// ===========================================================================
// launchers
// ===========================================================================
static inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Box of interfaces for sweep `dir`: extended (WENO) or core (flux) set.
static Box3 sweep_box(const Geo& g, int an, bool extended) {
    int P[3] = {g.P0, g.P1, g.P2};
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = extended ? 2 : 3;
        hi[a] = extended ? P[a] - 3 : P[a] - 4;
    }
    lo[an] = 2;
    hi[an] = P[an] - 4;
    Box3 b;
    b.o0 = lo[0]; b.o1 = lo[1]; b.o2 = lo[2];
    b.n0 = hi[0] - lo[0] + 1; b.n1 = hi[1] - lo[1] + 1; b.n2 = hi[2] - lo[2] + 1;
    return b;
}

static Box3 interior_box(const Geo& g) {
    Box3 b;
    b.o0 = b.o1 = b.o2 = 3;
    b.n0 = g.ny; b.n1 = g.nx; b.n2 = g.nz;
    return b;
}

// axis of the sweep direction: X -> 1, Y -> 0, Z -> 2
static inline int normal_axis(int dir) { return dir == 0 ? 1 : (dir == 1 ? 0 : 2); }

static inline long axis_stride(const Geo& g, int a) { return a == 0 ? g.s0 : (a == 1 ? g.s1 : 1); }

// tangential axes paired with the y / z slots (update.cpp:217-245, 291-318, 380-407)
static inline void tangential_axes(int dir, int zbug, int& t1, int& t2) {
    if (dir == 0) { t1 = 0; t2 = 2; }
    else if (dir == 1) { t1 = 1; t2 = 2; }
    else if (zbug) { t1 = 1; t2 = 0; }
    else { t1 = 0; t2 = 1; }
}

template <class Real, class St>
static void launch_sweep(Ctx* h, int dir, St st, const Real* Tw, const Real* tau, Planes10<Real> wi,
                         Planes5<Real> F, const StepPar& par, hipEvent_t* ev) {
    const Geo& g = h->geo;
    const int an = normal_axis(dir);
    const long sn = axis_stride(g, an);
    int t1, t2;
    tangential_axes(dir, h->par.zbug_compat, t1, t2);
    const long st1 = axis_stride(g, t1), st2 = axis_stride(g, t2);
    Box3 be = sweep_box(g, an, true);
    Box3 bc = sweep_box(g, an, false);
    if (ev) (void)hipEventRecord(ev[0], h->stream);
    if (h->par.recon == GKM_RECON_JS)
        hipLaunchKernelGGL((k_recon<Real, St, 1>), dim3(blocks_for(be.count(), 256)), dim3(256), 0, h->stream, g, be,
                           sn, st, Tw, wi, par);
    else
        hipLaunchKernelGGL((k_recon<Real, St, 0>), dim3(blocks_for(be.count(), 256)), dim3(256), 0, h->stream, g, be,
                           sn, st, Tw, wi, par);
    if (ev) (void)hipEventRecord(ev[1], h->stream);
    const int bs = h->flux_block;
    if (dir == 0)
        hipLaunchKernelGGL((k_flux<Real, St, 0>), dim3(blocks_for(bc.count(), bs)), dim3(bs), 0, h->stream, g, bc, sn,
                           st1, st2, st, tau, wi, F, par);
    else if (dir == 1)
        hipLaunchKernelGGL((k_flux<Real, St, 1>), dim3(blocks_for(bc.count(), bs)), dim3(bs), 0, h->stream, g, bc, sn,
                           st1, st2, st, tau, wi, F, par);
    else
        hipLaunchKernelGGL((k_flux<Real, St, 2>), dim3(blocks_for(bc.count(), bs)), dim3(bs), 0, h->stream, g, bc, sn,
                           st1, st2, st, tau, wi, F, par);
    if (ev) (void)hipEventRecord(ev[2], h->stream);
}

void launch_ghost(Ctx* h, const StepPar* par) {
    const Geo& g = h->geo;
    long n = (long)g.P0 * g.P1 * g.P2;
    StepPar p0 = par ? *par : StepPar{};
    if (h->par.precision == GKM_PREC_FP64) {
        Planes10<double> f{};
        StateD st;
        for (int q = 0; q < 5; q++) f.p[q] = st.w[q] = h->Wd[q];
        hipLaunchKernelGGL((k_ghost_prim<5, double, double, StateD>), dim3(blocks_for(n, 256)), dim3(256), 0,
                           h->stream, g, f, st, par ? h->Td : nullptr, h->taud, p0);
    } else {
        Planes10<float> f{};
        StateM st;
        for (int q = 0; q < 5; q++) {
            f.p[q] = st.hi[q] = h->Whi[q];
            f.p[5 + q] = st.lo[q] = h->Wlo[q];
        }
        hipLaunchKernelGGL((k_ghost_prim<10, float, float, StateM>), dim3(blocks_for(n, 256)), dim3(256), 0,
                           h->stream, g, f, st, par ? h->Tf : nullptr, h->tauf, p0);
    }
}

StepPar step_par(const Ctx* h, double dt) {
    StepPar p;
    p.gam = h->par.gamma;
    p.R = h->par.R;
    p.invPr = 1.0 / h->par.Pr;
    p.mu0 = h->par.mu0;
    p.T0 = h->par.T0;
    p.visc_exp = h->par.visc_exp;
    p.dt = dt;
    p.dx = h->grid.dx;
    p.K = h->par.K;
    p.visc_law = h->par.visc_law;
    p.energy_recon = h->par.energy_recon;
    return p;
}

void launch_step(Ctx* h, double dt, bool timed, int step_index) {
    const Geo& g = h->geo;
    const StepPar par = step_par(h, dt);
    hipEvent_t* ev = timed ? h->ev_step + step_index * 12 : nullptr;
    if (ev) (void)hipEventRecord(ev[0], h->stream);
    launch_ghost(h, &par);
    if (ev) (void)hipEventRecord(ev[1], h->stream);
    Box3 bi = interior_box(g);
    if (h->par.precision == GKM_PREC_FP64) {
        StateD st;
        for (int q = 0; q < 5; q++) st.w[q] = h->Wd[q];
        Planes10<double> wi;
        for (int q = 0; q < 10; q++) wi.p[q] = h->WId[q];
        for (int dir = 0; dir < 3; dir++) {
            Planes5<double> F;
            for (int q = 0; q < 5; q++) F.p[q] = h->Fd[5 * dir + q];
            launch_sweep<double, StateD>(h, dir, st, h->Td, h->taud, wi, F, par, ev ? ev + 1 + 3 * dir : nullptr);
        }
        Planes15<double> F;
        for (int q = 0; q < 15; q++) F.p[q] = h->Fd[q];
        hipLaunchKernelGGL(k_update_d, dim3(blocks_for(bi.count(), 256)), dim3(256), 0, h->stream, g, bi, st, F,
                           h->grid.dx);
    } else {
        StateM st;
        for (int q = 0; q < 5; q++) {
            st.hi[q] = h->Whi[q];
            st.lo[q] = h->Wlo[q];
        }
        Planes10<float> wi;
        for (int q = 0; q < 10; q++) wi.p[q] = h->WIf[q];
        for (int dir = 0; dir < 3; dir++) {
            Planes5<float> F;
            for (int q = 0; q < 5; q++) F.p[q] = h->Ff[5 * dir + q];
            launch_sweep<float, StateM>(h, dir, st, h->Tf, h->tauf, wi, F, par, ev ? ev + 1 + 3 * dir : nullptr);
        }
        Planes15<float> F;
        for (int q = 0; q < 15; q++) F.p[q] = h->Ff[q];
        hipLaunchKernelGGL(k_update_m, dim3(blocks_for(bi.count(), 256)), dim3(256), 0, h->stream, g, bi, st, F,
                           (float)(1.0 / h->grid.dx));
    }
    if (ev) (void)hipEventRecord(ev[11], h->stream);
}

void launch_split(Ctx* h, int q, const double* dsrc) {
    long n = h->ncell;
    hipLaunchKernelGGL(k_split, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, n, dsrc, h->Whi[q], h->Wlo[q]);
}

void launch_merge(Ctx* h, int q, double* ddst) {
    long n = h->ncell;
    hipLaunchKernelGGL(k_merge, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, n, h->Whi[q], h->Wlo[q], ddst);
}

void launch_flux_points(hipStream_t s, int prec, int n, const double* rec, double dx, double gam, double invPr, int K,
                        double* out) {
    if (prec == GKM_PREC_FP64)
        hipLaunchKernelGGL(k_flux_points<double>, dim3(blocks_for(n, 64)), dim3(64), 0, s, n, rec, dx, gam, invPr, K,
                           out);
    else
        hipLaunchKernelGGL(k_flux_points<float>, dim3(blocks_for(n, 64)), dim3(64), 0, s, n, rec, dx, gam, invPr, K,
                           out);
}

void launch_weno_points(hipStream_t s, int prec, int recon, int n, const double* rec, double gam, double R,
                        int energy_recon, double* out) {
    dim3 gr(blocks_for(n, 64)), bl(64);
    if (prec == GKM_PREC_FP64) {
        if (recon == GKM_RECON_JS)
            hipLaunchKernelGGL((k_weno_points<double, 1>), gr, bl, 0, s, n, rec, gam, R, energy_recon, out);
        else
            hipLaunchKernelGGL((k_weno_points<double, 0>), gr, bl, 0, s, n, rec, gam, R, energy_recon, out);
    } else {
        if (recon == GKM_RECON_JS)
            hipLaunchKernelGGL((k_weno_points<float, 1>), gr, bl, 0, s, n, rec, gam, R, energy_recon, out);
        else
            hipLaunchKernelGGL((k_weno_points<float, 0>), gr, bl, 0, s, n, rec, gam, R, energy_recon, out);
    }
}

}  // namespace gkm
