// gkm_step_dev.h — device code of one GKM-DNS time step on gfx950: the
// kernels and block functions that gkm_step.hip launches (and the
// translation units that compile some of them with other scheduler
// settings, gkm_step_ilp.hip and gkm_step_mc.hip).
#pragma once
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <utility>

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
//   tau = mu(T)/p, p = rho R T                 the left-cell collision time
//         (update.cpp:262-264, 882-891: p/(R rho) there is this same T)
// so that the sweeps read them instead of recomputing them per interface.
// The temperature recovery is FP64 in both forms, with one reciprocal of rho
// per cell.  mu(T) and tau are formed in the arithmetic type of the flux that
// reads them: the power law as exp2(e log2(T/T0)) (FP32: the hardware
// v_log_f32 / v_exp_f32), Sutherland's law as written.
// ---------------------------------------------------------------------------
template <class Real>
__device__ __forceinline__ Real collision_time(const StepPar& par, double T, double P) {
    if (par.visc_law == 1) {
        const Real x = (Real)(T * par.ud[kUIT0]);
        Real mu;
        if constexpr (sizeof(Real) == 8)
            mu = par.ud[kUMu0] * exp2(par.ud[kUVexp] * log2(x));
        else
            mu = par.uf[kUMu0] * __builtin_amdgcn_exp2f(par.uf[kUVexp] * __builtin_amdgcn_logf(x));
        return mu / (Real)P;
    }
    const double x = T * par.ud[kUIT0];
    return (Real)(par.ud[kUMu0] * x * sqrt(x) * (par.T0 + 110.4) / ((T + 110.4) * P));
}

template <class Real>
__device__ __forceinline__ void cell_prims(const StepPar& par, double r0, double m1, double m2, double m3, double E,
                                           Real& Tw, Real& tau) {
    const double ri = 1.0 / r0;
    const double T = (par.gam - 1) * (E * ri - 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) * (ri * ri)) * par.ud[kUIR];
    Tw = (Real)T;
    tau = collision_time<Real>(par, T, r0 * par.R * T);
}

template <int NPL, class T, class Real, class St>
__global__ void __launch_bounds__(256) k_ghost_prim(Geo g, Box3 all, Planes10<T> f, St st, Real* Tw, Real* tau,
                                                    const StepPar* __restrict__ pars) {
    const unsigned tl = blockIdx.x * blockDim.x + threadIdx.x;
    if (tl >= all.count()) return;
    const unsigned base = blockIdx.y * g.ncell;
    const unsigned t = base + tl;
    const unsigned r = fdiv(tl, all.f2);
    const int k = (int)(tl - r * (unsigned)all.n2);
    const int j = (int)fdiv(r, all.f1);
    const int i = (int)(r - (unsigned)j * (unsigned)all.n1);
    const bool gj = j < 3 || j >= g.ny + 3, gi = i < 3 || i >= g.nx + 3, gk = k < 3 || k >= g.nz + 3;
    unsigned src = t;
    if (gj || gi || gk) {
        const int js = j < 3 ? j + g.ny : (j >= g.ny + 3 ? j - g.ny : j);
        const int is = i < 3 ? i + g.nx : (i >= g.nx + 3 ? i - g.nx : i);
        const int ks = k < 3 ? k + g.nz : (k >= g.nz + 3 ? k - g.nz : k);
        src = base + (unsigned)js * g.s0 + (unsigned)is * g.s1 + (unsigned)ks;
#pragma unroll
        for (int q = 0; q < NPL; q++) put(f.p[q], t, ld(f.p[q], src));
    }
    if (!Tw) return;
    const StepPar& par = pars[blockIdx.y];
    Real Tc, tc;
    cell_prims<Real>(par, st.get(0, src), st.get(1, src), st.get(2, src), st.get(3, src), st.get(4, src), Tc, tc);
    put(Tw, t, Tc);
    put(tau, t, tc);
}

// ---------------------------------------------------------------------------
// Slab-exchange ghost fill (GKM_HALO_PACK).  The 3-plane y faces are
// contiguous in this layout (y is the slowest axis): plane j starts at j*s0.
// pack:   send[box][plane][side][e] <- face (side 0: j = 3..5, side 1: j = ny..ny+2)
// (exchange: a ring neighbour's send buffer becomes this box's receive buffer;
//  on one GPU with periodic ends it is a copy of its own)
// unpack: ghosts j = 0..2 <- recv side 1, j = ny+3..ny+5 <- recv side 0
// then the x and z ghosts of every plane, y ghosts included, take their
// periodic image, which leaves edges and corners as the reference's y-then-x-
// then-z exchange order does (update.cpp:894-1054).
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256) k_halo_pack(Geo g, Planes10<T> f, T* __restrict__ send, int npl) {
    const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned face = 3u * g.s0;
    if (e >= face) return;
    const unsigned q = blockIdx.y >> 1, side = blockIdx.y & 1u, box = blockIdx.z;
    const unsigned src = box * g.ncell + (side ? (unsigned)g.ny : 3u) * g.s0 + e;
    put(send, ((box * npl + q) * 2u + side) * face + e, ld(f.p[q], src));
}

template <class T>
__global__ void __launch_bounds__(256) k_halo_unpack(Geo g, Planes10<T> f, const T* __restrict__ recv, int npl) {
    const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned face = 3u * g.s0;
    if (e >= face) return;
    const unsigned q = blockIdx.y >> 1, side = blockIdx.y & 1u, box = blockIdx.z;
    // the low face (side 0) lands in the high ghosts, the high face in the low ghosts
    const unsigned dst = box * g.ncell + (side ? 0u : (unsigned)(g.ny + 3)) * g.s0 + e;
    put(f.p[q], dst, ld(recv, ((box * npl + q) * 2u + side) * face + e));
}

template <int NPL, class T>
__global__ void __launch_bounds__(256) k_ghost_xz(Geo g, Box3 all, Planes10<T> f) {
    const unsigned tl = blockIdx.x * blockDim.x + threadIdx.x;
    if (tl >= all.count()) return;
    const unsigned base = blockIdx.y * g.ncell;
    const unsigned r = fdiv(tl, all.f2);
    const int k = (int)(tl - r * (unsigned)all.n2);
    const int j = (int)fdiv(r, all.f1);
    const int i = (int)(r - (unsigned)j * (unsigned)all.n1);
    const bool gi = i < 3 || i >= g.nx + 3, gk = k < 3 || k >= g.nz + 3;
    if (!(gi || gk)) return;
    const int is = i < 3 ? i + g.nx : (i >= g.nx + 3 ? i - g.nx : i);
    const int ks = k < 3 ? k + g.nz : (k >= g.nz + 3 ? k - g.nz : k);
    const unsigned src = base + (unsigned)j * g.s0 + (unsigned)is * g.s1 + (unsigned)ks;
#pragma unroll
    for (int q = 0; q < NPL; q++) put(f.p[q], base + tl, ld(f.p[q], src));
}

template <class Real, class St>
__global__ void __launch_bounds__(256) k_prim(Geo g, St st, Real* Tw, Real* tau, const StepPar* __restrict__ pars) {
    const unsigned tl = blockIdx.x * blockDim.x + threadIdx.x;
    if (tl >= g.ncell) return;
    const unsigned t = blockIdx.y * g.ncell + tl;
    const StepPar& par = pars[blockIdx.y];
    Real Tc, tc;
    cell_prims<Real>(par, st.get(0, t), st.get(1, t), st.get(2, t), st.get(3, t), st.get(4, t), Tc, tc);
    put(Tw, t, Tc);
    put(tau, t, tc);
}

// The per-cell primitives of the ghost shell only (GKM_HALO_PACK, after the
// exchange; the Z sweep wrote the interior's): t runs over the 6 P1 P2 cells
// of the y ghost layers, then the 6 ny P2 of the x ghost layers inside them,
// then the 6 ny nx of the z ghost layers inside both (each ghost cell once).
template <class Real, class St>
__global__ void __launch_bounds__(256) k_prim_shell(Geo g, St st, Real* Tw, Real* tau,
                                                    const StepPar* __restrict__ pars) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned P0 = (unsigned)g.P0, P1 = (unsigned)g.P1, P2 = (unsigned)g.P2;
    const unsigned ny = (unsigned)g.ny, nx = (unsigned)g.nx;
    const unsigned n1 = 6u * P1 * P2, n2 = 6u * ny * P2, n3 = 6u * ny * nx;
    if (t >= n1 + n2 + n3) return;
    unsigned j, i, k;
    if (t < n1) {
        const unsigned layer = t / (P1 * P2), r = t - layer * (P1 * P2);
        j = layer < 3u ? layer : P0 - 6u + layer;
        i = r / P2;
        k = r - i * P2;
    } else if (t < n1 + n2) {
        const unsigned u = t - n1, row = u / (6u * P2), v = u - row * (6u * P2), il = v / P2;
        j = 3u + row;
        i = il < 3u ? il : P1 - 6u + il;
        k = v - il * P2;
    } else {
        const unsigned u = t - n1 - n2, row = u / (6u * nx), w = u - row * (6u * nx), kl = w % 6u;
        j = 3u + row;
        i = 3u + w / 6u;
        k = kl < 3u ? kl : P2 - 6u + kl;
    }
    const unsigned c = blockIdx.y * g.ncell + j * g.s0 + i * g.s1 + k;
    const StepPar& par = pars[blockIdx.y];
    Real Tc, tc;
    cell_prims<Real>(par, st.get(0, c), st.get(1, c), st.get(2, c), st.get(3, c), st.get(4, c), Tc, tc);
    put(Tw, c, Tc);
    put(tau, c, tc);
}

// ---------------------------------------------------------------------------
// WENO reconstruction of one sweep on the extended interface set (normal
// index [2, P-4], tangential [2, P-3]): each interface once.  The reference
// evaluates the line of every transverse neighbour again inside each
// interface (5 lines per interface, update.cpp:247-251); the flux kernel
// reads the neighbours' values from here instead, which are the same numbers.
//
// Temperature mode (update.cpp:560-621): rho, rho u, rho v, rho w by WENO, the
// temperature (from the per-cell plane) by WENO, rho E rebuilt from both.
// ---------------------------------------------------------------------------
// (the reconstruction kernels fuse multiplies into adds only within a statement,
// as gkm_device.h's WENO code does, so that every kernel forms the same bits)
#pragma clang fp contract(on)
template <int RECON, int ENERGY, class Real>
__device__ __forceinline__ void recon_one(const Real (&v)[6][6], Real eps, const StepPar& par, Real wl[5],
                                          Real wr[5]) {
    constexpr int nlin = ENERGY == 1 ? 5 : 4;
#pragma unroll
    for (int q = 0; q < nlin; q++) weno_line<RECON>(v[q], eps, wl[q], wr[q]);
    if (ENERGY == 0) {
        Real Tl, Tr;
        weno_line<RECON>(v[5], eps, Tl, Tr);
        const Real cR = upar<Real>(par, kUCR);
        wl[4] = wl[0] * cR * Tl + Real(0.5) * (wl[1] * wl[1] + wl[2] * wl[2] + wl[3] * wl[3]) * M<Real>::rcp(wl[0]);
        wr[4] = wr[0] * cR * Tr + Real(0.5) * (wr[1] * wr[1] + wr[2] * wr[2] + wr[3] * wr[3]) * M<Real>::rcp(wr[0]);
    }
}

// recon_one for a thread walking consecutive interfaces: smoothness weights of
// the left value carried from the previous interface's right value
template <int RECON, int ENERGY, class Real>
__device__ __forceinline__ void recon_one_carry(const Real (&v)[6][6], Real eps, const StepPar& par,
                                                Carry<Real> (&cw)[6], Real wl[5], Real wr[5]) {
    constexpr int nlin = ENERGY == 1 ? 5 : 4;
#pragma unroll
    for (int q = 0; q < nlin; q++) weno_line_carry<RECON>(v[q], eps, cw[q], wl[q], wr[q]);
    if (ENERGY == 0) {
        Real Tl, Tr;
        weno_line_carry<RECON>(v[5], eps, cw[5], Tl, Tr);
        const Real cR = upar<Real>(par, kUCR);
        wl[4] = wl[0] * cR * Tl + Real(0.5) * (wl[1] * wl[1] + wl[2] * wl[2] + wl[3] * wl[3]) * M<Real>::rcp(wl[0]);
        wr[4] = wr[0] * cR * Tr + Real(0.5) * (wr[1] * wr[1] + wr[2] * wr[2] + wr[3] * wr[3]) * M<Real>::rcp(wr[0]);
    }
}

// Unit-stride sweep (Z) as the persistent group kernel runs it: one interface
// per thread, the six-cell stencils from memory (neighbouring threads share
// them through the cache).  The per-stage Z sweep is k_recon_zc below, which
// gives the same bits.
// (the block functions below take the block's place -- bx along the work, box
// -- as arguments, so that the persistent group kernel can run them too)
template <class Real, class St, int RECON, int ENERGY>
__device__ __forceinline__ void recon_z_block(const Geo& g, const Box3& b, unsigned sn, const St& st,
                                              const Real* __restrict__ Tw, const Planes10<Real>& wi,
                                              const StepPar& par, unsigned bx, unsigned box) {
    const unsigned t = bx * blockDim.x + threadIdx.x;
    if (t >= b.count()) return;
    const unsigned c = box * g.ncell + box_index(g, b, t);
    Real v[6][6];
    constexpr int nlin = ENERGY == 1 ? 5 : 4;
#pragma unroll
    for (int q = 0; q < nlin; q++)
#pragma unroll
        for (int s = 0; s < 6; s++) v[q][s] = (Real)st.cell(q, c + s * sn - 2 * sn);
    if (ENERGY == 0)
#pragma unroll
        for (int s = 0; s < 6; s++) v[5][s] = ld(Tw, c + s * sn - 2 * sn);
    Real wl[5], wr[5];
    recon_one<RECON, ENERGY>(v, Real(1e-10), par, wl, wr);
#pragma unroll
    for (int q = 0; q < 5; q++) {
        put(wi.p[q], c, wl[q]);
        put(wi.p[5 + q], c, wr[q]);
    }
}

// The Z sweep with the smoothness weights carried across lanes and an LDS
// row tile.
//
// Positions: each (j, i) line of the extended set has NP = nI + 1 positions
// p = 0 .. nI for the padded interfaces k = p + 1; p = 0 (k = 1) is a
// virtual position whose only use is its right-side weights, which are the
// left-side ones of the line's first interface (k = 2).  Positions are
// numbered line by line, k fastest, and wavefront w takes the 64 positions
// from 63 w, so consecutive wavefronts share one.  Every lane forms the
// right-side weights of its interface (weno_right); lane l takes its
// left-side ones from lane l - 1 (the previous interface of its line, or the
// line's virtual position) and forms both values (weno_line_lr).  Lane 0,
// whose position the previous wavefront's lane 63 writes, and the virtual
// positions write nothing, so every interface is written exactly once, with
// the bits of recon_z_block (the single-interface form forms every weight
// with the same expression).
//
// Tile: a block's 253 positions lie on consecutive z lines of one j row,
// which are contiguous in memory, so the cells its stencils read form one
// range [c_first - 2, c_last + 3].  The block loads that range once
// (coalesced, each cell once) and reads the six-cell stencils from LDS.  A
// line end skips 4 cells and 252 positions cross at most ceil(252 / NP) of
// them, so the range is at most span_max = 258 + 4 ceil(252 / NP) cells; a
// block whose range is longer (it spans two j rows) loads from memory.
struct ZPos {
    unsigned count;  // positions of one box: lines * NP
    FastDiv fNP;     // NP = nI + 1 positions per line
    FastDiv fn1;     // lines per j row (the extended i range)
};

constexpr unsigned kZPosBlock = 4u * 63u;  // new positions per 256-thread block

__device__ __forceinline__ unsigned zpos_cell(const Geo& g, const ZPos& zp, unsigned q, unsigned& p) {
    const unsigned line = fdiv(q, zp.fNP);
    p = q - line * zp.fNP.d;
    const unsigned j = fdiv(line, zp.fn1);
    const unsigned i = line - j * zp.fn1.d;
    return (j + 2) * g.s0 + (i + 2) * g.s1 + (p + 1);
}

template <class T> __device__ __forceinline__ Carry<T> carry_from_prev_lane(const Carry<T>& c, bool symoo) {
    Carry<T> o;
#pragma unroll
    for (int k = 0; k < 3; k++) o.r[k] = __shfl_up(c.r[k], 1u, 64);
    if (symoo) {
#pragma unroll
        for (int k = 0; k < 3; k++) o.is[k] = __shfl_up(c.is[k], 1u, 64);
    }
    return o;
}

template <class Real, class St, int RECON, int ENERGY>
__global__ void __launch_bounds__(256) k_recon_zc(Geo g, ZPos zp, St st, const Real* __restrict__ Tw,
                                                  Planes10<Real> wi, const StepPar* __restrict__ pars,
                                                  unsigned span_max) {
    extern __shared__ unsigned char zt_raw[];
    Real* tile = reinterpret_cast<Real*>(zt_raw);  // [6][span_max]: fields 0..4 (or 0..3 and T in slot 5)
    constexpr int nlin = ENERGY == 1 ? 5 : 4;
    const unsigned box0 = blockIdx.y * g.ncell;
    const unsigned lane = threadIdx.x & 63u;
    const unsigned q0 = blockIdx.x * kZPosBlock;
    const unsigned q = q0 + (threadIdx.x >> 6) * 63u + lane;
    const bool valid = q < zp.count;
    unsigned pf, pl, p;
    const unsigned lo = zpos_cell(g, zp, q0, pf) - 2u;
    const unsigned span = zpos_cell(g, zp, min(q0 + kZPosBlock, zp.count - 1u), pl) + 3u - lo + 1u;
    const bool tiled = span <= span_max;  // block-uniform
    if (tiled) {
        for (unsigned z = threadIdx.x; z < span; z += blockDim.x) {
            const unsigned cc = box0 + lo + z;
#pragma unroll
            for (int f = 0; f < nlin; f++) tile[f * span_max + z] = (Real)st.cell(f, cc);
            if (ENERGY == 0) tile[5 * span_max + z] = ld(Tw, cc);
        }
        __syncthreads();
    }
    // lanes past the end take the last position (in bounds, never written) so
    // that the whole wavefront reaches the lane exchange
    const unsigned cl = zpos_cell(g, zp, valid ? q : zp.count - 1u, p);
    Real v[6][6];
    if (tiled) {
        const unsigned o = cl - lo - 2u;  // the stencil's first cell in the tile
#pragma unroll
        for (int f = 0; f < nlin; f++)
#pragma unroll
            for (int s = 0; s < 6; s++) v[f][s] = tile[f * span_max + o + s];
        if (ENERGY == 0)
#pragma unroll
            for (int s = 0; s < 6; s++) v[5][s] = tile[5 * span_max + o + s];
    } else {
        const unsigned c0 = box0 + cl - 2u;
#pragma unroll
        for (int f = 0; f < nlin; f++)
#pragma unroll
            for (int s = 0; s < 6; s++) v[f][s] = (Real)st.cell(f, c0 + s);
        if (ENERGY == 0)
#pragma unroll
            for (int s = 0; s < 6; s++) v[5][s] = ld(Tw, c0 + s);
    }
    const StepPar& par = pars[blockIdx.y];
    const Real eps = Real(1e-10);
    Real wl[5], wr[5], Tl = Real(0), Tr = Real(0);
#pragma unroll
    for (int f = 0; f < 6; f++) {
        if (f >= nlin && !(ENERGY == 0 && f == 5)) continue;
        Carry<Real> rw;
        weno_right<RECON>(v[f], eps, rw);
        const Carry<Real> lw = carry_from_prev_lane(rw, RECON != 1);
        Real a, b;
        weno_line_lr<RECON>(v[f], eps, lw, rw, a, b);
        if (f < 5) {
            wl[f] = a;
            wr[f] = b;
        } else {
            Tl = a;
            Tr = b;
        }
    }
    if (ENERGY == 0) {
        const Real cR = upar<Real>(par, kUCR);
        wl[4] = wl[0] * cR * Tl + Real(0.5) * (wl[1] * wl[1] + wl[2] * wl[2] + wl[3] * wl[3]) * M<Real>::rcp(wl[0]);
        wr[4] = wr[0] * cR * Tr + Real(0.5) * (wr[1] * wr[1] + wr[2] * wr[2] + wr[3] * wr[3]) * M<Real>::rcp(wr[0]);
    }
    if (!valid || lane == 0u || p == 0u) return;
    const unsigned c = box0 + cl;
#pragma unroll
    for (int f = 0; f < 5; f++) {
        put(wi.p[f], c, wl[f]);
        put(wi.p[5 + f], c, wr[f]);
    }
}

// Strided sweeps (X, Y): a thread walks CHUNK consecutive interfaces along the
// sweep normal with a six-cell register window, so each cell is loaded once
// per chunk instead of six times, and carries the smoothness weights of each
// interface's right value to the next interface's left value; threads run
// along z (unit stride) for coalescing.  Box axes: n0 = chunks, n1 = the other
// tangential axis, n2 = z.
template <class Real, class St, int RECON, int ENERGY, int CHUNK>
__device__ __forceinline__ void recon_win_block(const Geo& g, const Box3& b, unsigned sn, unsigned sa, int n_normal,
                                                const St& st, const Real* __restrict__ Tw,
                                                const Planes10<Real>& wi, const StepPar& par, unsigned bx,
                                                unsigned box) {
    const unsigned t = bx * blockDim.x + threadIdx.x;
    if (t >= b.count()) return;
    const unsigned r = fdiv(t, b.f2);
    const unsigned k = t - r * (unsigned)b.n2;
    const unsigned chunk = fdiv(r, b.f1);
    const unsigned a = r - chunk * (unsigned)b.n1;
    const int first = (int)chunk * CHUNK;                 // normal index offset from 2
    const int count = min(CHUNK, n_normal - first);
    unsigned c = box * g.ncell + (unsigned)(first + 2) * sn + (a + 2) * sa + (k + 2);
    constexpr bool tmode = ENERGY == 0;
    constexpr int nlin = tmode ? 4 : 5;
    Real v[6][6];
#pragma unroll
    for (int q = 0; q < 6; q++)
#pragma unroll
        for (int s = 0; s < 6; s++) v[q][s] = Real(0);
#pragma unroll
    for (int q = 0; q < nlin; q++)
#pragma unroll
        for (int s = 0; s < 5; s++) v[q][s + 1] = (Real)st.cell(q, c + s * sn - 2 * sn);
    if (tmode)
#pragma unroll
        for (int s = 0; s < 5; s++) v[5][s + 1] = ld(Tw, c + s * sn - 2 * sn);
    // left-side weights of the first interface (its cells LLL..RR are v[.][1..5])
    Carry<Real> cw[6];
#pragma unroll
    for (int q = 0; q < 6; q++)
        if (q < nlin || (tmode && q == 5)) weno_pset<RECON>(&v[q][1], Real(1e-10), cw[q]);
    // the cell entering the window next, loaded one iteration ahead so its
    // latency hides behind the previous interface's arithmetic
    Real nx[6];
#pragma unroll
    for (int q = 0; q < nlin; q++) nx[q] = (Real)st.cell(q, c + 3 * sn);
    if (tmode) nx[5] = ld(Tw, c + 3 * sn);
#pragma unroll 1
    for (int it = 0; it < count; it++) {
        // slide: drop the oldest cell, take the prefetched one, fetch the next
#pragma unroll
        for (int q = 0; q < 6; q++)
#pragma unroll
            for (int s = 0; s < 5; s++) v[q][s] = v[q][s + 1];
#pragma unroll
        for (int q = 0; q < nlin; q++) v[q][5] = nx[q];
        if (tmode) v[5][5] = nx[5];
        if (it + 1 < count) {
#pragma unroll
            for (int q = 0; q < nlin; q++) nx[q] = (Real)st.cell(q, c + 4 * sn);
            if (tmode) nx[5] = ld(Tw, c + 4 * sn);
        }
        Real wl[5], wr[5];
        recon_one_carry<RECON, ENERGY>(v, Real(1e-10), par, cw, wl, wr);
#pragma unroll
        for (int q = 0; q < 5; q++) {
            put(wi.p[q], c, wl[q]);
            put(wi.p[5 + q], c, wr[q]);
        }
        c += sn;
    }
}

template <class Real, class St, int RECON, int ENERGY, int CHUNK>
__global__ void __launch_bounds__(256) k_recon_win(Geo g, Box3 b, unsigned sn, unsigned sa, int n_normal, St st,
                                                   const Real* __restrict__ Tw, Planes10<Real> wi,
                                                   const StepPar* __restrict__ pars) {
    recon_win_block<Real, St, RECON, ENERGY, CHUNK>(g, b, sn, sa, n_normal, st, Tw, wi, pars[blockIdx.y], blockIdx.x,
                                                    blockIdx.y);
}

#pragma clang fp contract(fast)

// slot -> field map of the rotated flux frame (update.cpp:340-352, 429-441)
template <int DIR> __device__ __forceinline__ int slot_field(int s) {
    if (DIR == 1) return s == 1 ? 2 : (s == 2 ? 1 : s);
    if (DIR == 2) return s == 1 ? 3 : (s == 3 ? 1 : s);
    return s;
}

// ---------------------------------------------------------------------------
// BGK flux of one sweep on the core interface set (normal [2, P-4],
// tangential [3, P-4]).  flux_at gives the flux of interface c in the rotated
// slot order; k_flux writes the five flux planes of its direction (X, Y), and
// the Z sweep's k_flux_update below also applies the update.  st1 / st2 are
// the strides of the tangential axes paired with the y / z slots.
// FP64: at least 2 waves per SIMD (256 registers); the unconstrained build
// needs 260 and runs at 1 wave, about 1.5x slower.  FP32 fits 3 waves without
// a bound; forcing 4 spills.
// ---------------------------------------------------------------------------

// The transverse slope differences of interface c into per-thread LDS slots,
// slot k of thread t at lds[k * blockDim.x + t]: Dy left (k = s), Dy right
// (5 + s), Dz left (10 + s), Dz right (15 + s), in the rotated slot order.  The
// FP64 flux stages them ahead of its arithmetic, so their forty loads are all
// in flight at once and overlap the other wave's work instead of stalling
// this one in the middle of the flux (the values are the same differences).
template <class Real, int DIR>
__device__ __forceinline__ void flux_stage_slopes(unsigned c, unsigned st1, unsigned st2, const Planes10<Real>& wi,
                                                  Real* lds) {
    const unsigned t = threadIdx.x, b = blockDim.x;
#pragma unroll
    for (int s = 0; s < 5; s++) {
        const int q = slot_field<DIR>(s);
        lds[s * b + t] = ld(wi.p[q], c + st1) - ld(wi.p[q], c - st1);
        lds[(5 + s) * b + t] = ld(wi.p[5 + q], c + st1) - ld(wi.p[5 + q], c - st1);
        lds[(10 + s) * b + t] = ld(wi.p[q], c + st2) - ld(wi.p[q], c - st2);
        lds[(15 + s) * b + t] = ld(wi.p[5 + q], c + st2) - ld(wi.p[5 + q], c - st2);
    }
}

// staged: the transverse differences come from flux_stage_slopes' slots
template <class Real, class St, int DIR>
__device__ __forceinline__ void flux_at(unsigned c, unsigned sn, unsigned st1, unsigned st2, const St& st,
                                        const Real* __restrict__ tau, const Planes10<Real>& wi,
                                        const StepPar& par, Real Fs[5], const Real* staged = nullptr) {
    Gas<Real> gas;
    gas.gam = upar<Real>(par, kUGam);
    gas.invPr = upar<Real>(par, kUInvPr);
    gas.K = upar<Real>(par, kUK);
    gas.kp3 = upar<Real>(par, kUKp3);
    gas.ikp3 = upar<Real>(par, kUIkp3);
    const Real idx = upar<Real>(par, kUIdx);
    const Real idx2 = 2 * idx, hidx = Real(0.5) * idx;

    // collision time mu/p of the left cell (update.cpp:262-264), from the plane
    const Real tau0 = ld(tau, c);
    Real WL[5], WR[5], Wl[5], Wr[5];
#pragma unroll
    for (int s = 0; s < 5; s++) {
        const int q = slot_field<DIR>(s);
        WL[s] = (Real)st.cell(q, c);
        WR[s] = (Real)st.cell(q, c + sn);
        Wl[s] = ld(wi.p[q], c);
        Wr[s] = ld(wi.p[5 + q], c);
    }
    auto slopes = [&](int side, Real Dx[5], Real Dy[5], Real Dz[5]) {
#pragma unroll
        for (int s = 0; s < 5; s++) {
            const int q = slot_field<DIR>(s);
            const int o = side == 0 ? q : 5 + q;
            // differences; the flux scales them by 2/dx (normal: half-cell
            // distance) and 1/(2 dx) (tangential: central difference)
            Dx[s] = side == 0 ? Wl[s] - WL[s] : WR[s] - Wr[s];
            if (staged) {
                const unsigned t = threadIdx.x, b = blockDim.x;
                Dy[s] = staged[(5 * side + s) * b + t];
                Dz[s] = staged[(10 + 5 * side + s) * b + t];
            } else {
                Dy[s] = ld(wi.p[o], c + st1) - ld(wi.p[o], c - st1);
                Dz[s] = ld(wi.p[o], c + st2) - ld(wi.p[o], c - st2);
            }
        }
    };
    bgk_flux(gas, WL, WR, Wl, Wr, slopes, idx2, hidx, tau0, upar<Real>(par, kUDt), idx, Fs);
}

template <class Real, class St, int DIR>
__device__ __forceinline__ void flux_block(const Geo& g, const Box3& b, unsigned sn, unsigned st1, unsigned st2,
                                           const St& st, const Real* __restrict__ tau, const Planes10<Real>& wi,
                                           const Planes5<Real>& F, const StepPar& par, unsigned bx, unsigned box,
                                           Real* stage_lds = nullptr) {
    const unsigned t = bx * blockDim.x + threadIdx.x;
    if (t >= b.count()) return;
    const unsigned c = box * g.ncell + box_index(g, b, t);
    if (stage_lds) flux_stage_slopes<Real, DIR>(c, st1, st2, wi, stage_lds);
    Real Fs[5];
    flux_at<Real, St, DIR>(c, sn, st1, st2, st, tau, wi, par, Fs, stage_lds);
#pragma unroll
    for (int s = 0; s < 5; s++) put(F.p[slot_field<DIR>(s)], c, Fs[s]);
}

template <class Real, class St, int DIR>
__global__ void __launch_bounds__(256, sizeof(Real) == 8 ? 2 : 1)
    k_flux(Geo g, Box3 b, unsigned sn, unsigned st1, unsigned st2, St st, const Real* __restrict__ tau,
           Planes10<Real> wi, Planes5<Real> F, const StepPar* __restrict__ pars) {
    // plain block order: the XCD-contiguous numbering halves this kernel's
    // fetches but measured 4-5 % slower (the kernel is VALU-bound).  FP64
    // stages the transverse slope differences in LDS (flux_stage_slopes).
    extern __shared__ unsigned char flux_lds_raw[];
    Real* stage = sizeof(Real) == 8 ? reinterpret_cast<Real*>(flux_lds_raw) : nullptr;
    flux_block<Real, St, DIR>(g, b, sn, st1, st2, st, tau, wi, F, pars[blockIdx.y], blockIdx.x, blockIdx.y, stage);
}

// ---------------------------------------------------------------------------
// Sweep of the GKM tree's smooth-flow scheme (GKM/CUDA GKM/src/CPUupdate.cpp:
// 29-192) on the core interface set: W0 = (W_i + W_i+1)/2, DWx = (W_i+1 -
// W_i)/dx, DWy / DWz = half the difference of the neighbour rows' interface
// averages along the tangential axes paired with the y / z slots, mu(T) from
// the interface average (:62-65).  The two cells of the interface are read
// in FP64 (T, p and tau are recovered in FP64 in both forms); the tangential
// neighbours only enter slope differences and are read as cell values.
// ---------------------------------------------------------------------------
// the GKM tree's smooth-flow flux at interface c (rotated slot order in Fs)
template <class Real, class St, int DIR>
__device__ __forceinline__ void gks_at(unsigned c, unsigned sn, unsigned st1, unsigned st2, const St& st,
                                       const StepPar& par, Real Fs[5]) {
    Gas<Real> gas;
    gas.gam = upar<Real>(par, kUGam);
    gas.invPr = upar<Real>(par, kUInvPr);
    gas.K = upar<Real>(par, kUK);
    gas.kp3 = upar<Real>(par, kUKp3);
    gas.ikp3 = upar<Real>(par, kUIkp3);
    const double idx = par.ud[kUIdx];
    const Real hidx = Real(0.5 * idx);
    double w0[5];  // interface average, field order
    Real W0[5], Dx[5], Dy[5], Dz[5];
#pragma unroll
    for (int s = 0; s < 5; s++) {
        const int q = slot_field<DIR>(s);
        const double wl = st.get(q, c), wr = st.get(q, c + sn);
        w0[q] = 0.5 * (wr + wl);
        W0[s] = (Real)w0[q];
        Dx[s] = (Real)((wr - wl) * idx);
        const Real n1 = Real(0.5) * ((Real)st.cell(q, c + st1 + sn) + (Real)st.cell(q, c + st1));
        const Real s1 = Real(0.5) * ((Real)st.cell(q, c - st1 + sn) + (Real)st.cell(q, c - st1));
        const Real n2 = Real(0.5) * ((Real)st.cell(q, c + st2 + sn) + (Real)st.cell(q, c + st2));
        const Real s2 = Real(0.5) * ((Real)st.cell(q, c - st2 + sn) + (Real)st.cell(q, c - st2));
        Dy[s] = (n1 - s1) * hidx;
        Dz[s] = (n2 - s2) * hidx;
    }
    const double r = w0[0], u = w0[1] / r, v = w0[2] / r, w = w0[3] / r;
    const double P = (r * (par.gam - 1)) * (w0[4] / r - 0.5 * (u * u + v * v + w * w));
    const Real tau = collision_time<Real>(par, P / (par.R * r), P);
    gks_flux(gas, W0, Dx, Dy, Dz, tau, upar<Real>(par, kUDt), Fs);
}

template <class Real, class St, int DIR>
__global__ void __launch_bounds__(256) k_gks(Geo g, Box3 b, unsigned sn, unsigned st1, unsigned st2, St st,
                                             Planes5<Real> F, const StepPar* __restrict__ pars) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.count()) return;
    const StepPar& par = pars[blockIdx.y];
    const unsigned c = blockIdx.y * g.ncell + box_index(g, b, t);
    Real Fs[5];
    gks_at<Real, St, DIR>(c, sn, st1, st2, st, par, Fs);
#pragma unroll
    for (int s = 0; s < 5; s++) put(F.p[slot_field<DIR>(s)], c, Fs[s]);
}

// ---------------------------------------------------------------------------
// Conservative update of one interior cell (update.cpp:455-469): the flux
// differences are summed in the reference's order.
// ---------------------------------------------------------------------------
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

// The update of interior cell c into the other buffer, in two halves around
// the z flux.  update_stage loads what does not depend on the z flux -- the x
// and y flux differences (summed in the reference's order as far as they go)
// and the old state -- and parks it in LDS, slot s of lane t at
// lds[s * blockDim.x + t]: issued before the flux, these loads overlap the
// flux's own input loads instead of stalling the wave after it, and they hold
// no registers across the flux.  update_finish adds the two z fluxes of the
// cell's right (Fz) and left (Fm) faces and stores the new value; the sum is
// the same expression, term for term, as in one pass.
// new value of a cell in the state's own representation
template <class St> struct CellVal;
template <> struct CellVal<StateD> {
    double w[5];
    __device__ __forceinline__ double get(int q) const { return w[q]; }
};
template <> struct CellVal<StateM> {
    float hi[5], lo[5];
    __device__ __forceinline__ double get(int q) const { return (double)hi[q] + (double)lo[q]; }
};
__device__ __forceinline__ void store_cell(const StateD& N, unsigned c, const CellVal<StateD>& v) {
#pragma unroll
    for (int q = 0; q < 5; q++) put(N.w[q], c, v.w[q]);
}
__device__ __forceinline__ void store_cell(const StateM& N, unsigned c, const CellVal<StateM>& v) {
#pragma unroll
    for (int q = 0; q < 5; q++) {
        put(N.hi[q], c, v.hi[q]);
        put(N.lo[q], c, v.lo[q]);
    }
}

template <class St> struct UpdSlots;
template <> struct UpdSlots<StateD> { static constexpr int n = 10; };  // d[5], w[5]
template <> struct UpdSlots<StateM> { static constexpr int n = 15; };  // R[5], hi[5], lo[5]

__device__ __forceinline__ void update_stage(const Geo& g, unsigned c, const StateD& S, const Planes15<double>& F,
                                             double* lds) {
    const unsigned t = threadIdx.x, b = blockDim.x;
#pragma unroll
    for (int q = 0; q < 5; q++) {
        lds[q * b + t] = ld(F.p[q], c) - ld(F.p[q], c - g.s1) + ld(F.p[5 + q], c) - ld(F.p[5 + q], c - g.s0);
        lds[(5 + q) * b + t] = ld(S.w[q], c);
    }
}
__device__ __forceinline__ void update_finish(unsigned c, const StateD& N, const double Fz[5], const double Fm[5],
                                              const double* lds, const StepPar& par, CellVal<StateD>& v) {
    const unsigned t = threadIdx.x, b = blockDim.x;
    const double idx = par.ud[kUIdx];  // 1 / dx
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const double d = lds[q * b + t] + Fz[q] - Fm[q];
        v.w[q] = lds[(5 + q) * b + t] - idx * d;
    }
    store_cell(N, c, v);
}
__device__ __forceinline__ void update_stage(const Geo& g, unsigned c, const StateM& S, const Planes15<float>& F,
                                             float* lds) {
    const unsigned t = threadIdx.x, b = blockDim.x;
#pragma unroll
    for (int q = 0; q < 5; q++) {
        lds[q * b + t] = ld(F.p[q], c) - ld(F.p[q], c - g.s1) + ld(F.p[5 + q], c) - ld(F.p[5 + q], c - g.s0);
        lds[(5 + q) * b + t] = ld(S.hi[q], c);
        lds[(10 + q) * b + t] = ld(S.lo[q], c);
    }
}
__device__ __forceinline__ void update_finish(unsigned c, const StateM& N, const float Fz[5], const float Fm[5],
                                              const float* lds, const StepPar& par, CellVal<StateM>& v) {
    const unsigned t = threadIdx.x, b = blockDim.x;
    const float idx = par.uf[kUIdx];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const float R = lds[q * b + t] + Fz[q] - Fm[q];
        float hi = lds[(5 + q) * b + t], lo = lds[(10 + q) * b + t];
        two_sum_add(hi, lo, -idx * R);
        v.hi[q] = hi;
        v.lo[q] = lo;
    }
    store_cell(N, c, v);
}

// The next step's ghost cells and per-cell primitives, from the lane that
// wrote interior cell c: the ghosts whose periodic image the cell is -- padded
// J + ny when J < 6, J - ny when J >= ny, the same in x and z, and their
// combinations for edges and corners (every ghost cell has exactly one
// interior image, the cell k_ghost_prim copies it from) -- take its new value,
// and, for the WENO-BGK scheme, the cell and those ghosts take its Tw and tau
// (cell_prims of the new FP64 state, as k_ghost_prim forms them).  Tw and tau
// go to the planes the next step reads.  imask (image_mask) says which images
// the cell has.
__device__ __forceinline__ unsigned image_mask(const Geo& g, unsigned J, unsigned I, unsigned K) {
    return (J < 6u ? 1u : 0u) | (J >= (unsigned)g.ny ? 2u : 0u) | (I < 6u ? 4u : 0u) |
           (I >= (unsigned)g.nx ? 8u : 0u) | (K < 6u ? 16u : 0u) | (K >= (unsigned)g.nz ? 32u : 0u);
}

// A new cell value is physical when rho and T are positive and finite; T is
// formed as cell_prims forms it (a non-finite momentum or energy makes T
// non-finite or NaN, so this covers every field).  The fused Z kernel raises
// the box's flag on any other value; the host reads it at its next wait.
template <class St>
__device__ __forceinline__ bool physical(const StepPar& par, const CellVal<St>& v) {
    const double r0 = v.get(0);
    const double m1 = v.get(1), m2 = v.get(2), m3 = v.get(3);
    const double ri = 1.0 / r0;
    const double T = (par.gam - 1) * (v.get(4) * ri - 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) * (ri * ri)) * par.ud[kUIR];
    return r0 > 0.0 && r0 < __builtin_huge_val() && T > 0.0 && T < __builtin_huge_val();
}

// Returns, for the WENO-BGK scheme, whether the cell is physical (rho and the
// Tw just formed positive and finite), and true for the GKS scheme.
template <class Real, class St, int SCHEME>
__device__ __forceinline__ bool ghost_images(const Geo& g, unsigned imask, unsigned c, const St& N,
                                             const CellVal<St>& v, Real* Tw, Real* tau, const StepPar& par) {
    constexpr bool prim = SCHEME != GKM_SCHEME_GKS;
    Real Tc = Real(0), tc = Real(0);
    bool ok = true;
    if constexpr (prim) {
        cell_prims<Real>(par, v.get(0), v.get(1), v.get(2), v.get(3), v.get(4), Tc, tc);
        put(Tw, c, Tc);
        put(tau, c, tc);
        const double r0 = v.get(0);
        ok = r0 > 0.0 && r0 < __builtin_huge_val() && Tc > Real(0) && Tc < Real(__builtin_huge_val());
    }
    if (!imask) return ok;
    const int n[3] = {g.ny * (int)g.s0, g.nx * (int)g.s1, g.nz};
    int off[3][3];  // per axis: own position, low image (+n), high image (-n); 0 = none
#pragma unroll
    for (int a = 0; a < 3; a++) {
        off[a][0] = 0;
        off[a][1] = (imask >> (2 * a)) & 1u ? n[a] : 0;
        off[a][2] = (imask >> (2 * a + 1)) & 1u ? -n[a] : 0;
    }
    for (int x = 0; x < 3; x++) {
        if (x && !off[0][x]) continue;
        for (int y = 0; y < 3; y++) {
            if (y && !off[1][y]) continue;
            for (int z = 0; z < 3; z++) {
                if ((z && !off[2][z]) || !(x || y || z)) continue;
                const unsigned ci = c + (unsigned)(off[0][x] + off[1][y] + off[2][z]);
                store_cell(N, ci, v);
                if constexpr (prim) {
                    put(Tw, ci, Tc);
                    put(tau, ci, tc);
                }
            }
        }
    }
    return ok;
}

// ---------------------------------------------------------------------------
// Z sweep fused with the conservative update.  Interfaces are numbered along
// the z lines of the interior (j, i) columns, interface k+1/2 for k in
// [2, P2-4] (nI = nz+1 per line), flattened with k fastest.  Wavefront w
// takes the 64 consecutive interfaces starting at 63 w, so consecutive
// wavefronts share one: lane l computes the flux at its interface, takes its
// left neighbour's from lane l-1 (__shfl_up), and for l >= 1 and k >= 3
// updates cell k from the x, y fluxes of the earlier sweeps and the two z
// fluxes (update_stage before the flux, update_finish after it).  The state
// is read from the current buffer and the
// new interior written to the other one, so no lane's read of a neighbour
// cell can see an updated value.  The flux is the WENO-BGK one (flux_at, from
// the interface planes) or the GKS one (gks_at, from the cells).
// ---------------------------------------------------------------------------
// Lines are taken in strips of W columns of i, (j, i) within a strip, so a
// line's y neighbours are W lines away in launch order instead of nx; with the
// XCD-contiguous block numbering (xcd_block) both neighbours of a line are read
// while they are still in that XCD's L2.
struct ZLines {
    unsigned count;   // interfaces of one box: lines * nI
    FastDiv fnI;      // nI = nz + 1 interfaces per line
    FastDiv fstrip;   // W * ny lines per full strip
    FastDiv fW, fWl;  // strip width W and the last strip's width (nx - nfull W, or W)
    unsigned nfull;   // strips of width W
};


// line l -> interior column (j, i)
__device__ __forceinline__ void zline_column(const ZLines& zl, unsigned l, unsigned& j, unsigned& i) {
    const unsigned s = fdiv(l, zl.fstrip);
    const unsigned r = l - s * zl.fstrip.d;
    const FastDiv& fw = s < zl.nfull ? zl.fW : zl.fWl;
    j = fdiv(r, fw);
    i = s * zl.fW.d + (r - j * fw.d);
}

// vb: the block's number along the box's wavefronts; lds: the block's dynamic
// LDS (UpdSlots<St>::n values, then one image mask, per thread)
template <class Real, class St, int SCHEME, bool GHOST>
__device__ __forceinline__ void flux_update_block(const Geo& g, const ZLines& zl, unsigned st1, unsigned st2,
                                                  const St& st, const St& nx, const Real* __restrict__ tau,
                                                  const Planes10<Real>& wi, const Planes15<Real>& F,
                                                  const StepPar& par, Real* __restrict__ Tw_nx,
                                                  Real* __restrict__ tau_nx, unsigned* __restrict__ bad,
                                                  unsigned vb, unsigned box, unsigned char* upd_lds_raw) {
    const unsigned gt = vb * blockDim.x + threadIdx.x;
    const unsigned lane = threadIdx.x & 63u;
    const unsigned e = (gt >> 6) * 63u + lane;
    const bool valid = e < zl.count;
    unsigned c = 0, imask = 0;
    bool upd = false;  // lane 0's cell belongs to the previous wavefront; k = 2 is a ghost
    if (valid) {
        const unsigned line = fdiv(e, zl.fnI);
        const unsigned kk = e - line * (zl.fnI.d);
        upd = lane != 0 && kk != 0;
        unsigned j, i;
        zline_column(zl, line, j, i);
        c = box * g.ncell + (j + 3) * g.s0 + (i + 3) * g.s1 + (kk + 2);
        if constexpr (GHOST) imask = image_mask(g, j + 3, i + 3, kk + 2);
    }
    // the flags as a wave-uniform mask (scalar registers), taken before the
    // flux so that no per-lane index value stays live in a vector register
    // across it
    const unsigned long long upd_mask = __ballot(upd);
    // UpdSlots<St>::n * blockDim.x values, then blockDim.x image masks
    Real* upd_lds = reinterpret_cast<Real*>(upd_lds_raw);
    unsigned* img_lds = reinterpret_cast<unsigned*>(upd_lds + UpdSlots<St>::n * blockDim.x);
    // the image mask waits in LDS across the flux, like the update's operands
    if constexpr (GHOST) img_lds[threadIdx.x] = imask;
    // Where the staging goes was measured (256^3): FP64, in a block of its own
    // ahead of the flux (fused kernel 2.05 -> 1.91 ms; the compiler issues it
    // first and waits for it before the flux's loads); FP32, in the flux's block
    // (0.875 -> 0.83 ms; its loads are scheduled among the flux's own).  Every
    // valid lane stages in the second form: c - s0 and c - s1 are in the box
    // for k = 2 and for lane 0 too, and their slots are never read.
    constexpr bool kStageFirst = sizeof(Real) == 8;
    if (kStageFirst && upd) update_stage(g, c, st, F, upd_lds);
    Real Fz[5] = {Real(0), Real(0), Real(0), Real(0), Real(0)};
    if (valid) {
        if (!kStageFirst) update_stage(g, c, st, F, upd_lds);
        Real Fs[5];
        if constexpr (SCHEME == GKM_SCHEME_GKS)
            gks_at<Real, St, 2>(c, 1u, st1, st2, st, par, Fs);
        else
            flux_at<Real, St, 2>(c, 1u, st1, st2, st, tau, wi, par, Fs);
#pragma unroll
        for (int s = 0; s < 5; s++) Fz[slot_field<2>(s)] = Fs[s];
    }
    Real Fm[5];
#pragma unroll
    for (int q = 0; q < 5; q++) Fm[q] = __shfl_up(Fz[q], 1u, 64);
    if (!((upd_mask >> __lane_id()) & 1ull)) return;
    CellVal<St> v;
    update_finish(c, nx, Fz, Fm, upd_lds, par, v);
    bool ok;
    if constexpr (SCHEME != GKM_SCHEME_GKS) {
        // the cell's Tw and tau for the next step, and (GHOST) its ghost images
        ok = ghost_images<Real, St, SCHEME>(g, GHOST ? img_lds[threadIdx.x] : 0u, c, nx, v, Tw_nx, tau_nx, par);
    } else {
        if constexpr (GHOST) ghost_images<Real, St, SCHEME>(g, img_lds[threadIdx.x], c, nx, v, Tw_nx, tau_nx, par);
        ok = physical<St>(par, v);
    }
    if (!ok) bad[box] = 1u;
}

// Occupancy pinned at exactly 2 (FP64) / 3 (FP32) waves per SIMD: with only a
// lower bound the scheduler aims for more waves than the registers allow and
// orders the flux's loads less well (256^3: FP64 1.75 -> 1.58 ms, FP32
// 0.887 -> 0.86 ms).  The X/Y flux kernels measured slower with the same pin.
template <class Real, class St, int SCHEME, bool GHOST>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, 256),
                               amdgpu_waves_per_eu(sizeof(Real) == 8 ? 2 : 3, sizeof(Real) == 8 ? 2 : 3)))
    k_flux_update(Geo g, ZLines zl, unsigned st1, unsigned st2, St st, St nx, const Real* __restrict__ tau,
                  Planes10<Real> wi, Planes15<Real> F, const StepPar* __restrict__ pars, Real* __restrict__ Tw_nx,
                  Real* __restrict__ tau_nx, unsigned* __restrict__ bad) {
    extern __shared__ unsigned char upd_lds_raw[];
    // (staging the transverse slope differences here too, as k_flux does in
    // FP64, measured 1.56 -> 1.74 ms: it brings back a 36-byte spill)
    flux_update_block<Real, St, SCHEME, GHOST>(g, zl, st1, st2, st, nx, tau, wi, F, pars[blockIdx.y], Tw_nx, tau_nx,
                                               bad, xcd_block(), blockIdx.y, upd_lds_raw);
}

// ---------------------------------------------------------------------------
// Persistent group step (gkm_group_step_persistent).  One cooperative launch
// advances every box of every member handle (one size class each) nsteps
// steps.  The grid holds as many blocks of 256 threads as fit on the chip at
// once; every block runs the six phases of a step -- WENO X, flux X, WENO Y,
// flux Y, WENO Z, fused Z flux + update + next ghosts -- over the phase's
// blocks of every class and box in turn (the same block functions as the
// per-stage kernels), with a grid barrier after each phase: a phase reads
// what the one before wrote anywhere in the box (interface planes, X/Y
// fluxes, the new state).  The state and tau buffers alternate each step as
// the handles' do.  Every block reaches every barrier: the phase loops are
// block-uniform and nothing returns early from the kernel.
// ---------------------------------------------------------------------------
template <class Real, class St> struct PClass {
    Geo g;
    int nbox, chunk;
    unsigned nblk[6];                // blocks of 256 threads per box, per phase
    Box3 rb[2];                      // WENO X, Y: chunks x other tangential x z
    unsigned rsn[2], rsa[2];
    int rnn[2];
    Box3 zb;                         // WENO Z: the extended interface set
    Box3 fb[2];                      // flux X, Y: the core interface set
    unsigned fsn[2], fst1[2], fst2[2];
    ZLines zl;                       // fused Z kernel
    unsigned zst1, zst2;
    St st[2];                        // state buffers; st[0] current at launch
    Real* tau[2];
    Real* Tw;
    Planes10<Real> wi;
    Planes15<Real> F;
    const StepPar* pars;
    unsigned* bad;
};

constexpr int kPersistThreads = 256;

template <class Real, class St, int RECON, int ENERGY>
__device__ __forceinline__ void persist_phase_block(const PClass<Real, St>& C, int ph, int par, unsigned bx,
                                                    unsigned box, unsigned char* lds) {
    const St& cur = C.st[par];
    const St& nxt = C.st[par ^ 1];
    const StepPar& sp = C.pars[box];
    switch (ph) {
    case 0:
    case 2: {
        const int d = ph >> 1;
        if (C.chunk == 4)
            recon_win_block<Real, St, RECON, ENERGY, 4>(C.g, C.rb[d], C.rsn[d], C.rsa[d], C.rnn[d], cur, C.Tw, C.wi,
                                                        sp, bx, box);
        else if (C.chunk == 16)
            recon_win_block<Real, St, RECON, ENERGY, 16>(C.g, C.rb[d], C.rsn[d], C.rsa[d], C.rnn[d], cur, C.Tw, C.wi,
                                                         sp, bx, box);
        else
            recon_win_block<Real, St, RECON, ENERGY, 8>(C.g, C.rb[d], C.rsn[d], C.rsa[d], C.rnn[d], cur, C.Tw, C.wi,
                                                        sp, bx, box);
        break;
    }
    case 1:
    case 3: {
        const int d = ph >> 1;
        Planes5<Real> f;
#pragma unroll
        for (int q = 0; q < 5; q++) f.p[q] = C.F.p[5 * d + q];
        // as k_flux: FP64 stages the transverse slope differences (the same code, so the same bits)
        Real* stage = sizeof(Real) == 8 ? reinterpret_cast<Real*>(lds) : nullptr;
        if (d == 0)
            flux_block<Real, St, 0>(C.g, C.fb[0], C.fsn[0], C.fst1[0], C.fst2[0], cur, C.tau[par], C.wi, f, sp, bx,
                                    box, stage);
        else
            flux_block<Real, St, 1>(C.g, C.fb[1], C.fsn[1], C.fst1[1], C.fst2[1], cur, C.tau[par], C.wi, f, sp, bx,
                                    box, stage);
        break;
    }
    case 4:
        recon_z_block<Real, St, RECON, ENERGY>(C.g, C.zb, 1u, cur, C.Tw, C.wi, sp, bx, box);
        break;
    default:
        flux_update_block<Real, St, GKM_SCHEME_WENO_BGK, true>(C.g, C.zl, C.zst1, C.zst2, cur, nxt, C.tau[par], C.wi,
                                                               C.F, sp, C.Tw, C.tau[par ^ 1], C.bad, bx, box, lds);
        break;
    }
}

template <class Real, class St, int RECON, int ENERGY>
__global__ void __launch_bounds__(kPersistThreads, sizeof(Real) == 8 ? 2 : 1)
    k_group_persist(const PClass<Real, St>* __restrict__ cls, int ncls, int nsteps) {
    extern __shared__ unsigned char persist_lds[];
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const unsigned G = gridDim.x, p = blockIdx.x;
    for (int step = 0; step < nsteps; step++) {
        const int par = step & 1;
        for (int ph = 0; ph < 6; ph++) {
            unsigned V = 0;  // the phase's blocks over every class and box
            for (int k = 0; k < ncls; k++) V += cls[k].nblk[ph] * (unsigned)cls[k].nbox;
            // the fused Z phase takes consecutive blocks on one XCD (its lines'
            // neighbours then share that XCD's L2, as xcd_block arranges for
            // the per-stage kernel); the others take blocks round-robin
            unsigned v0 = p, v1 = V, dv = G;
            if (ph == 5) {
                const unsigned x = p & 7u, q = V >> 3, rem = V & 7u;
                v0 = x * q + min(x, rem) + (p >> 3);
                v1 = x * q + min(x, rem) + q + (x < rem ? 1u : 0u);
                dv = G >> 3;
            }
            for (unsigned v = v0; v < v1; v += dv) {
                unsigned w = v;
                int k = 0;
                while (w >= cls[k].nblk[ph] * (unsigned)cls[k].nbox) {
                    w -= cls[k].nblk[ph] * (unsigned)cls[k].nbox;
                    k++;
                }
                const unsigned nb = cls[k].nblk[ph];
                const unsigned box = w / nb, bx = w - box * nb;
                persist_phase_block<Real, St, RECON, ENERGY>(cls[k], ph, par, bx, box, persist_lds);
            }
            // the grid barrier also makes the phase's stores visible to every
            // thread of the grid (cooperative groups' sync); fences of our own
            // on both sides of it measured 12 % slower (6.10 vs 5.43 ms per
            // step on the 96-box ensemble) with the same bits
            if (step + 1 < nsteps || ph < 5) grid.sync();
        }
    }
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
    bgk_flux(gas, v[0], v[1], v[2], v[3], slopes, Real(1), Real(1), (Real)x[50], (Real)x[51], (Real)(1.0 / dx),
             F);
    for (int q = 0; q < 5; q++) out[(long)t * 5 + q] = (double)F[q];
}

// GKS flux of single records: W0[5] DWx[5] DWy[5] DWz[5] mu dt (slot order)
template <class Real>
__global__ void k_gks_points(int n, const double* rec, double gam, double invPr, int K, double* out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double* x = rec + (size_t)t * 22;
    Gas<Real> gas;
    gas.gam = (Real)gam;
    gas.invPr = (Real)invPr;
    gas.K = (Real)K;
    gas.kp3 = (Real)(K + 3);
    gas.ikp3 = (Real)(1.0 / (K + 3));
    Real W0[5], Dx[5], Dy[5], Dz[5], F[5];
    for (int q = 0; q < 5; q++) {
        W0[q] = (Real)x[q];
        Dx[q] = (Real)x[5 + q];
        Dy[q] = (Real)x[10 + q];
        Dz[q] = (Real)x[15 + q];
    }
    // tau = mu/p of the interface average (GKMflux.cpp:47), in FP64
    const double r = x[0], u = x[1] / r, v = x[2] / r, w = x[3] / r;
    const double P = (r * (gam - 1)) * (x[4] / r - 0.5 * (u * u + v * v + w * w));
    gks_flux(gas, W0, Dx, Dy, Dz, (Real)(x[20] / P), (Real)x[21], F);
    for (int q = 0; q < 5; q++) out[(size_t)t * 5 + q] = (double)F[q];
}

template <class Real, int RECON>
__global__ void k_weno_points(int n, const double* rec, double gam, double R, int energy_recon, double* out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double* x = rec + (long)t * 30;
    StepPar par{};
    par.gam = gam;
    par.R = R;
    par.K = 0;
    par.dx = 1.0;
    derive_uniforms(par);
    Real v[6][6];
    for (int q = 0; q < 5; q++)
        for (int s = 0; s < 6; s++) v[q][s] = (Real)x[s * 5 + q];
    for (int s = 0; s < 6; s++) {
        const double* W = x + s * 5;
        v[5][s] = (Real)((gam - 1) * (W[4] / W[0] - 0.5 * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]) / (W[0] * W[0])) / R);
    }
    Real wl[5], wr[5];
    if (energy_recon == 1)
        recon_one<RECON, 1>(v, Real(1e-10), par, wl, wr);
    else
        recon_one<RECON, 0>(v, Real(1e-10), par, wl, wr);
    for (int q = 0; q < 5; q++) {
        out[(long)t * 10 + q] = (double)wl[q];
        out[(long)t * 10 + 5 + q] = (double)wr[q];
    }
}


// ---------------------------------------------------------------------------
// The windowed X/Y reconstruction (k_recon_win) is launched from translation
// units of their own, compiled with another setting of the compiler's
// instruction scheduler: the kernel is the one above, so the arithmetic and
// the bits are the same and only the instruction order differs.
//   gkm_step_ilp.hip (scheduled for instruction-level parallelism): FP32;
//   gkm_step_mc.hip (scheduled for long memory clauses): FP64.
// ---------------------------------------------------------------------------
template <class Real, class St, int RECON, int ENERGY, int CHUNK>
void launch_recon_win_k(dim3 grid, hipStream_t s, Geo g, Box3 b, unsigned sn, unsigned sa, int n_normal, St st,
                        const Real* Tw, Planes10<Real> wi, const StepPar* pars);

}  // namespace gkm
