// gkm_step.hip — kernels of one GKM-DNS time step on gfx950 and their
// launchers: periodic ghost fill fused with the per-cell primitives, WENO
// reconstruction per sweep, BGK flux per sweep, conservative update, state
// split/merge for the mixed form.
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

// ---- split / merge of the mixed state --------------------------------------
__global__ void __launch_bounds__(256) k_split(unsigned n, const double* src, float* hi, float* lo) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double v = ld(src, t);
    const float h = (float)v;
    put(hi, t, h);
    put(lo, t, (float)(v - (double)h));
}

__global__ void __launch_bounds__(256) k_merge(unsigned n, const float* hi, const float* lo, double* dst) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    put(dst, t, (double)ld(hi, t) + (double)ld(lo, t));
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

// ===========================================================================
// launchers
// ===========================================================================
static inline unsigned blocks_for(unsigned n, int bs) { return (n + bs - 1) / bs; }

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
    return make_box(lo[0], lo[1], lo[2], hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1);
}

static Box3 padded_box(const Geo& g) { return make_box(0, 0, 0, g.P0, g.P1, g.P2); }

// axis of the sweep direction: X -> 1, Y -> 0, Z -> 2
static inline int normal_axis(int dir) { return dir == 0 ? 1 : (dir == 1 ? 0 : 2); }

static inline unsigned axis_stride(const Geo& g, int a) { return a == 0 ? g.s0 : (a == 1 ? g.s1 : 1u); }

// z lines of the fused Z kernel, in strips of `strip` columns of i
static ZLines make_zlines(const Geo& g, unsigned strip) {
    ZLines zl;
    const unsigned nI = (unsigned)g.nz + 1;
    zl.count = (unsigned)g.ny * (unsigned)g.nx * nI;
    zl.fnI = make_fastdiv(nI);
    const unsigned W = std::min(strip, (unsigned)g.nx);
    zl.nfull = (unsigned)g.nx / W;
    const unsigned wl = (unsigned)g.nx - zl.nfull * W;
    zl.fstrip = make_fastdiv(W * (unsigned)g.ny);
    zl.fW = make_fastdiv(W);
    zl.fWl = make_fastdiv(wl ? wl : W);
    return zl;
}

// tangential axes paired with the y / z slots (update.cpp:217-245, 291-318, 380-407)
static inline void tangential_axes(int dir, int zbug, int& t1, int& t2) {
    if (dir == 0) { t1 = 0; t2 = 2; }
    else if (dir == 1) { t1 = 1; t2 = 2; }
    else if (zbug) { t1 = 1; t2 = 0; }
    else { t1 = 0; t2 = 1; }
}

template <class Real, class St, int RECON, int ENERGY, int CHUNK>
static void launch_recon_win(Ctx* h, const Box3& b, unsigned sn, unsigned sa, int n_normal, St st, const Real* Tw,
                             Planes10<Real> wi) {
    hipLaunchKernelGGL((k_recon_win<Real, St, RECON, ENERGY, CHUNK>), dim3(blocks_for(b.count(), 256), h->nbox),
                       dim3(256), 0, h->stream, h->geo, b, sn, sa, n_normal, st, Tw, wi, h->d_pars);
}

template <class Real, class St, int RECON, int ENERGY>
static void launch_recon(Ctx* h, int dir, St st, const Real* Tw, Planes10<Real> wi) {
    const Geo& g = h->geo;
    const int an = normal_axis(dir);
    const unsigned sn = axis_stride(g, an);
    if (dir == 2) {
        Box3 be = sweep_box(g, an, true);
        const unsigned NP = (unsigned)be.n2 + 1u;  // the virtual position and the nI interfaces of a line
        ZPos zp;
        zp.count = (unsigned)be.n0 * (unsigned)be.n1 * NP;
        zp.fNP = make_fastdiv(NP);
        zp.fn1 = make_fastdiv((unsigned)be.n1);
        const unsigned waves = (zp.count - 1u) / 63u + 1u;  // wavefront w covers [63 w, 63 w + 63]
        const unsigned span_max = 258u + 4u * ((kZPosBlock + NP - 1u) / NP);
        const size_t lds = 6 * sizeof(Real) * (size_t)span_max;
        hipLaunchKernelGGL((k_recon_zc<Real, St, RECON, ENERGY>), dim3((waves + 3u) / 4u, h->nbox), dim3(256), lds,
                           h->stream, g, zp, st, Tw, wi, h->d_pars, span_max);
        return;
    }
    // X (normal = axis 1) walks i with threads over (j, k); Y (normal = axis 0) walks j over (i, k)
    const int P[3] = {g.P0, g.P1, g.P2};
    const int oa = an == 1 ? 0 : 1;  // the other tangential axis
    const int n_normal = P[an] - 5;  // [2, P-4]
    const int chunk = h->recon_chunk;
    const int chunks = (n_normal + chunk - 1) / chunk;
    Box3 b = make_box(0, 0, 0, chunks, P[oa] - 4, P[2] - 4);
    const unsigned sa = axis_stride(g, oa);
    if (chunk == 4)
        launch_recon_win<Real, St, RECON, ENERGY, 4>(h, b, sn, sa, n_normal, st, Tw, wi);
    else if (chunk == 16)
        launch_recon_win<Real, St, RECON, ENERGY, 16>(h, b, sn, sa, n_normal, st, Tw, wi);
    else
        launch_recon_win<Real, St, RECON, ENERGY, 8>(h, b, sn, sa, n_normal, st, Tw, wi);
}

template <class Real, class St>
static void launch_recon_any(Ctx* h, int dir, St st, const Real* Tw, Planes10<Real> wi) {
    const bool js = h->par.recon == GKM_RECON_JS, e = h->par.energy_recon == GKM_ENERGY_E;
    if (js && e) launch_recon<Real, St, 1, 1>(h, dir, st, Tw, wi);
    else if (js) launch_recon<Real, St, 1, 0>(h, dir, st, Tw, wi);
    else if (e) launch_recon<Real, St, 0, 1>(h, dir, st, Tw, wi);
    else launch_recon<Real, St, 0, 0>(h, dir, st, Tw, wi);
}

// Records event off + k of a timed step (base: the step's kEvPerStep events)
// when the handle's timing mode asks for it; nothing when base is null.
struct EvRec {
    Ctx* h;
    hipEvent_t* base;
    int off;
    void operator()(int k) const {
        if (base && (h->ev_mask >> (off + k) & 1u)) (void)hipEventRecord(base[off + k], h->stream);
    }
};

template <class Real, class St>
static void launch_sweep(Ctx* h, int dir, St st, const Real* Tw, const Real* tau, Planes10<Real> wi,
                         Planes5<Real> F, EvRec ev) {
    const Geo& g = h->geo;
    const int an = normal_axis(dir);
    const unsigned sn = axis_stride(g, an);
    int t1, t2;
    tangential_axes(dir, h->par.zbug_compat, t1, t2);
    const unsigned st1 = axis_stride(g, t1), st2 = axis_stride(g, t2);
    Box3 bc = sweep_box(g, an, false);
    const int bs = h->flux_block;
    const dim3 grid(blocks_for(bc.count(), bs), h->nbox);
    ev(0);
    if (h->par.flux_scheme == GKM_SCHEME_GKS) {  // no reconstruction: one kernel per sweep
        ev(1);
        if (dir == 0)
            hipLaunchKernelGGL((k_gks<Real, St, 0>), grid, dim3(bs), 0, h->stream, g, bc, sn, st1, st2, st, F,
                               h->d_pars);
        else if (dir == 1)
            hipLaunchKernelGGL((k_gks<Real, St, 1>), grid, dim3(bs), 0, h->stream, g, bc, sn, st1, st2, st, F,
                               h->d_pars);
        else
            hipLaunchKernelGGL((k_gks<Real, St, 2>), grid, dim3(bs), 0, h->stream, g, bc, sn, st1, st2, st, F,
                               h->d_pars);
        ev(2);
        return;
    }
    // WENO-BGK: the X and Y sweeps (the Z sweep's flux is fused with the update)
    launch_recon_any<Real, St>(h, dir, st, Tw, wi);
    ev(1);
    const size_t lds = sizeof(Real) == 8 ? 20 * sizeof(Real) * (size_t)bs : 0;  // flux_stage_slopes
    if (dir == 0)
        hipLaunchKernelGGL((k_flux<Real, St, 0>), grid, dim3(bs), lds, h->stream, g, bc, sn, st1, st2, st, tau, wi, F,
                           h->d_pars);
    else
        hipLaunchKernelGGL((k_flux<Real, St, 1>), grid, dim3(bs), lds, h->stream, g, bc, sn, st1, st2, st, tau, wi, F,
                           h->d_pars);
    ev(2);
}

// Z sweep: reconstruction (WENO-BGK), then the flux fused with the update of
// every interior cell into the other state buffer
// ghost: the kernel also writes the next state's ghost cells and its Tw and
// tau planes (Tw_nx, tau_nx), so the next step needs no ghost fill
template <class Real, class St>
static void launch_sweep_z_update(Ctx* h, St st, St nx, const Real* Tw, const Real* tau, Planes10<Real> wi,
                                  Planes15<Real> F, EvRec ev, bool ghost, Real* Tw_nx, Real* tau_nx) {
    const Geo& g = h->geo;
    int t1, t2;
    tangential_axes(2, h->par.zbug_compat, t1, t2);
    const unsigned st1 = axis_stride(g, t1), st2 = axis_stride(g, t2);
    const ZLines zl = make_zlines(g, (unsigned)h->z_strip);
    const unsigned waves = (zl.count - 1 + 62) / 63;  // wavefront w covers [63 w, 63 w + 63]
    const int bs = h->zflux_block;
    const dim3 grid(blocks_for(waves * 64u, bs), h->nbox);
    const size_t lds = (sizeof(Real) * UpdSlots<St>::n + sizeof(unsigned)) * (size_t)bs;
    ev(0);
    auto fused = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, grid, dim3(bs), lds, h->stream, g, zl, st1, st2, st, nx, tau, wi, F, h->d_pars,
                           Tw_nx, tau_nx, h->d_bad);
    };
    if (h->par.flux_scheme == GKM_SCHEME_GKS) {  // no reconstruction
        ev(1);
        if (ghost)
            fused(k_flux_update<Real, St, GKM_SCHEME_GKS, true>);
        else
            fused(k_flux_update<Real, St, GKM_SCHEME_GKS, false>);
    } else {
        launch_recon_any<Real, St>(h, 2, st, Tw, wi);
        ev(1);
        if (ghost)
            fused(k_flux_update<Real, St, GKM_SCHEME_WENO_BGK, true>);
        else
            fused(k_flux_update<Real, St, GKM_SCHEME_WENO_BGK, false>);
    }
    ev(2);
}

// after a step: the buffers swap, and the new current state's ghosts and
// primitives are valid when the step's Z sweep wrote them (periodic fill)
void swap_state(Ctx* h) {
    for (int q = 0; q < 5; q++) {
        std::swap(h->Wd[q], h->Wd_nx[q]);
        std::swap(h->Whi[q], h->Whi_nx[q]);
        std::swap(h->Wlo[q], h->Wlo_nx[q]);
    }
    std::swap(h->taud, h->taud_nx);
    std::swap(h->tauf, h->tauf_nx);
    h->flip ^= 1;
    const bool pack = h->halo_mode == GKM_HALO_PACK;
    h->derived_valid = !pack;
    // in the slab exchange the Z sweep wrote the interior's Tw and tau
    h->prims_valid = pack && h->par.flux_scheme != GKM_SCHEME_GKS;
}

void launch_ghost(Ctx* h, bool prim) {
    const Geo& g = h->geo;
    Box3 all = padded_box(g);
    const dim3 grid(blocks_for(all.count(), 256), h->nbox);
    const bool pack = h->halo_mode == GKM_HALO_PACK;
    if (h->par.precision == GKM_PREC_FP64) {
        Planes10<double> f{};
        StateD st;
        for (int q = 0; q < 5; q++) f.p[q] = st.w[q] = h->Wd[q];
        if (pack) {
            const unsigned face = 3u * g.s0;
            const dim3 pg(blocks_for(face, 256), 2 * 5, h->nbox);
            hipLaunchKernelGGL(k_halo_pack<double>, pg, dim3(256), 0, h->stream, g, f, (double*)h->halo_send, 5);
            (void)hipMemcpyAsync(h->halo_recv, h->halo_send, sizeof(double) * 2 * face * 5 * h->nbox,
                                 hipMemcpyDeviceToDevice, h->stream);
            hipLaunchKernelGGL(k_halo_unpack<double>, pg, dim3(256), 0, h->stream, g, f, (const double*)h->halo_recv,
                               5);
            hipLaunchKernelGGL((k_ghost_xz<5, double>), grid, dim3(256), 0, h->stream, g, all, f);
        } else {
            hipLaunchKernelGGL((k_ghost_prim<5, double, double, StateD>), grid, dim3(256), 0, h->stream, g, all, f,
                               st, prim ? h->Td : nullptr, h->taud, h->d_pars);
        }
    } else {
        Planes10<float> f{};
        StateM st;
        for (int q = 0; q < 5; q++) {
            f.p[q] = st.hi[q] = h->Whi[q];
            f.p[5 + q] = st.lo[q] = h->Wlo[q];
        }
        if (pack) {
            const unsigned face = 3u * g.s0;
            const dim3 pg(blocks_for(face, 256), 2 * 10, h->nbox);
            hipLaunchKernelGGL(k_halo_pack<float>, pg, dim3(256), 0, h->stream, g, f, (float*)h->halo_send, 10);
            (void)hipMemcpyAsync(h->halo_recv, h->halo_send, sizeof(float) * 2 * face * 10 * h->nbox,
                                 hipMemcpyDeviceToDevice, h->stream);
            hipLaunchKernelGGL(k_halo_unpack<float>, pg, dim3(256), 0, h->stream, g, f, (const float*)h->halo_recv,
                               10);
            hipLaunchKernelGGL((k_ghost_xz<10, float>), grid, dim3(256), 0, h->stream, g, all, f);
        } else {
            hipLaunchKernelGGL((k_ghost_prim<10, float, float, StateM>), grid, dim3(256), 0, h->stream, g, all, f, st,
                               prim ? h->Tf : nullptr, h->tauf, h->d_pars);
        }
    }
}

// per-cell primitives when the ghost fill did not compute them (GKM_HALO_PACK):
// every padded cell, or (shell) the ghost shell only
static void launch_prim(Ctx* h, bool shell = false) {
    const Geo& g = h->geo;
    const unsigned nshell =
        6u * (unsigned)(g.P1 * g.P2) + 6u * (unsigned)(g.ny * g.P2) + 6u * (unsigned)(g.ny * g.nx);
    const dim3 grid(blocks_for(shell ? nshell : g.ncell, 256), h->nbox);
    if (h->par.precision == GKM_PREC_FP64) {
        StateD st;
        for (int q = 0; q < 5; q++) st.w[q] = h->Wd[q];
        if (shell)
            hipLaunchKernelGGL((k_prim_shell<double, StateD>), grid, dim3(256), 0, h->stream, g, st, h->Td, h->taud,
                               h->d_pars);
        else
            hipLaunchKernelGGL((k_prim<double, StateD>), grid, dim3(256), 0, h->stream, g, st, h->Td, h->taud,
                               h->d_pars);
    } else {
        StateM st;
        for (int q = 0; q < 5; q++) {
            st.hi[q] = h->Whi[q];
            st.lo[q] = h->Wlo[q];
        }
        if (shell)
            hipLaunchKernelGGL((k_prim_shell<float, StateM>), grid, dim3(256), 0, h->stream, g, st, h->Tf, h->tauf,
                               h->d_pars);
        else
            hipLaunchKernelGGL((k_prim<float, StateM>), grid, dim3(256), 0, h->stream, g, st, h->Tf, h->tauf,
                               h->d_pars);
    }
}

StepPar step_par(const gkm_grid& grid, const gkm_params& pp, double dt) {
    StepPar p;
    std::memset(&p, 0, sizeof p);
    p.gam = pp.gamma;
    p.R = pp.R;
    p.invPr = 1.0 / pp.Pr;
    p.mu0 = pp.mu0;
    p.T0 = pp.T0;
    p.visc_exp = pp.visc_exp;
    p.dt = dt;
    p.dx = grid.dx;
    p.K = pp.K;
    p.visc_law = pp.visc_law;
    p.energy_recon = pp.energy_recon;
    derive_uniforms(p);
    return p;
}

void prepare_step(Ctx* h) {
    const bool prim = h->par.flux_scheme != GKM_SCHEME_GKS;
    if (h->halo_mode != GKM_HALO_PACK) {
        if (!h->derived_valid) {
            launch_ghost(h, prim);
            h->derived_valid = true;
        }
    } else if (prim && !h->prims_valid) {
        // the interior's Tw and tau (the step recomputes the ghost shell's after its exchange)
        launch_prim(h);
        h->prims_valid = true;
    }
}

void ensure_ghosts(Ctx* h) {
    if (!h->derived_valid) launch_ghost(h, false);
}

void launch_step(Ctx* h, hipEvent_t* evs) {
    const EvRec ev{h, evs, 0};
    // the GKS scheme reads no per-cell primitive planes; both schemes fuse the
    // update into their Z sweep, which writes the other state buffer.  With
    // the periodic fill the Z sweep also writes that buffer's ghosts and
    // primitives, and the ghost fill runs only when they are not valid (after
    // set_state and the like); with the slab exchange (GKM_HALO_PACK) every
    // step starts with the exchange, and the Z sweep writes the interior's
    // primitives, so after the exchange only the ghost shell's are formed.
    // (prepare_step is also called before a captured step, outside the graph,
    // so that the graph holds the same work on every replay.)
    const bool prim = h->par.flux_scheme != GKM_SCHEME_GKS;
    const bool pack = h->halo_mode == GKM_HALO_PACK;
    ev(0);
    prepare_step(h);
    if (pack) launch_ghost(h, prim);
    ev(12);
    if (prim && pack) launch_prim(h, true);
    ev(1);
    if (h->par.precision == GKM_PREC_FP64) {
        StateD st, nx;
        for (int q = 0; q < 5; q++) {
            st.w[q] = h->Wd[q];
            nx.w[q] = h->Wd_nx[q];
        }
        Planes10<double> wi;
        for (int q = 0; q < 10; q++) wi.p[q] = h->WId[q];
        Planes15<double> F;
        for (int q = 0; q < 15; q++) F.p[q] = h->Fd[q];
        for (int dir = 0; dir < 2; dir++) {  // X and Y; Z is fused with the update
            Planes5<double> Fd;
            for (int q = 0; q < 5; q++) Fd.p[q] = h->Fd[5 * dir + q];
            launch_sweep<double, StateD>(h, dir, st, h->Td, h->taud, wi, Fd, EvRec{h, evs, 1 + 3 * dir});
        }
        launch_sweep_z_update<double, StateD>(h, st, nx, h->Td, h->taud, wi, F, EvRec{h, evs, 7}, !pack, h->Td,
                                              h->taud_nx);
    } else {
        StateM st, nx;
        for (int q = 0; q < 5; q++) {
            st.hi[q] = h->Whi[q];
            st.lo[q] = h->Wlo[q];
            nx.hi[q] = h->Whi_nx[q];
            nx.lo[q] = h->Wlo_nx[q];
        }
        Planes10<float> wi;
        for (int q = 0; q < 10; q++) wi.p[q] = h->WIf[q];
        Planes15<float> F;
        for (int q = 0; q < 15; q++) F.p[q] = h->Ff[q];
        for (int dir = 0; dir < 2; dir++) {
            Planes5<float> Ff;
            for (int q = 0; q < 5; q++) Ff.p[q] = h->Ff[5 * dir + q];
            launch_sweep<float, StateM>(h, dir, st, h->Tf, h->tauf, wi, Ff, EvRec{h, evs, 1 + 3 * dir});
        }
        launch_sweep_z_update<float, StateM>(h, st, nx, h->Tf, h->tauf, wi, F, EvRec{h, evs, 7}, !pack, h->Tf,
                                             h->tauf_nx);
    }
    ev(11);
    swap_state(h);
}

// ---- persistent group step: class table and launch -------------------------
template <class Real, class St>
static void fill_pclass(Ctx* h, const St& cur, const St& nxt, Real* tau0, Real* tau1, Real* Tw, Real* const wi[10],
                        Real* const F[15], PClass<Real, St>& C) {
    std::memset(&C, 0, sizeof C);
    const Geo& g = h->geo;
    C.g = g;
    C.nbox = h->nbox;
    C.chunk = h->recon_chunk;
    const int P[3] = {g.P0, g.P1, g.P2};
    for (int d = 0; d < 2; d++) {  // WENO X (normal axis 1), Y (axis 0), as launch_recon forms them
        const int an = normal_axis(d), oa = an == 1 ? 0 : 1;
        const int n_normal = P[an] - 5;
        const int chunks = (n_normal + C.chunk - 1) / C.chunk;
        C.rb[d] = make_box(0, 0, 0, chunks, P[oa] - 4, P[2] - 4);
        C.rsn[d] = axis_stride(g, an);
        C.rsa[d] = axis_stride(g, oa);
        C.rnn[d] = n_normal;
        C.nblk[2 * d] = blocks_for(C.rb[d].count(), kPersistThreads);
        int t1, t2;  // flux X, Y, as launch_sweep forms it
        tangential_axes(d, h->par.zbug_compat, t1, t2);
        C.fb[d] = sweep_box(g, an, false);
        C.fsn[d] = axis_stride(g, an);
        C.fst1[d] = axis_stride(g, t1);
        C.fst2[d] = axis_stride(g, t2);
        C.nblk[2 * d + 1] = blocks_for(C.fb[d].count(), kPersistThreads);
    }
    C.zb = sweep_box(g, 2, true);
    C.nblk[4] = blocks_for(C.zb.count(), kPersistThreads);
    C.zl = make_zlines(g, (unsigned)h->z_strip);
    int t1, t2;
    tangential_axes(2, h->par.zbug_compat, t1, t2);
    C.zst1 = axis_stride(g, t1);
    C.zst2 = axis_stride(g, t2);
    C.nblk[5] = blocks_for((C.zl.count - 1 + 62) / 63 * 64u, kPersistThreads);
    C.st[0] = cur;
    C.st[1] = nxt;
    C.tau[0] = tau0;
    C.tau[1] = tau1;
    C.Tw = Tw;
    for (int q = 0; q < 10; q++) C.wi.p[q] = wi[q];
    for (int q = 0; q < 15; q++) C.F.p[q] = F[q];
    C.pars = h->d_pars;
    C.bad = h->d_bad;
}

size_t persist_class_bytes(int precision) {
    return precision == GKM_PREC_FP64 ? sizeof(PClass<double, StateD>) : sizeof(PClass<float, StateM>);
}

void persist_fill_class(Ctx* h, void* dst) {
    if (h->par.precision == GKM_PREC_FP64) {
        StateD cur, nxt;
        for (int q = 0; q < 5; q++) {
            cur.w[q] = h->Wd[q];
            nxt.w[q] = h->Wd_nx[q];
        }
        fill_pclass<double, StateD>(h, cur, nxt, h->taud, h->taud_nx, h->Td, h->WId, h->Fd,
                                    *(PClass<double, StateD>*)dst);
    } else {
        StateM cur, nxt;
        for (int q = 0; q < 5; q++) {
            cur.hi[q] = h->Whi[q];
            cur.lo[q] = h->Wlo[q];
            nxt.hi[q] = h->Whi_nx[q];
            nxt.lo[q] = h->Wlo_nx[q];
        }
        fill_pclass<float, StateM>(h, cur, nxt, h->tauf, h->tauf_nx, h->Tf, h->WIf, h->Ff,
                                   *(PClass<float, StateM>*)dst);
    }
}

template <class Real, class St, int RECON, int ENERGY>
static hipError_t launch_persist_t(const void* table, int ncls, int nsteps, hipStream_t s, int* grid_out) {
    const void* fn = reinterpret_cast<const void*>(&k_group_persist<Real, St, RECON, ENERGY>);
    // the fused Z phase's update slots and image masks, or the FP64 flux
    // phases' 20 staged slope differences per thread, whichever is larger
    const size_t lds = std::max((sizeof(Real) * UpdSlots<St>::n + sizeof(unsigned)),
                                sizeof(Real) == 8 ? 20 * sizeof(Real) : (size_t)0) *
                       kPersistThreads;
    int dev = 0, coop = 0, ncu = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_group_persist<Real, St, RECON, ENERGY>,
                                                         kPersistThreads, lds);
    if (e != hipSuccess) return e;
    if (!coop) return hipErrorNotSupported;
    // every block resident at once (the barriers wait for all of them); a
    // multiple of 8, one share per XCD
    const int G = per_cu * ncu / 8 * 8;
    if (G < 8) return hipErrorInvalidConfiguration;
    if (grid_out) *grid_out = G;
    const PClass<Real, St>* cls = static_cast<const PClass<Real, St>*>(table);
    void* args[] = {(void*)&cls, (void*)&ncls, (void*)&nsteps};
    return hipLaunchCooperativeKernel(fn, dim3(G), dim3(kPersistThreads), args, (unsigned)lds, s);
}

hipError_t launch_persist(const gkm_params& par, const void* table, int ncls, int nsteps, hipStream_t s,
                          int* grid_out) {
    const bool js = par.recon == GKM_RECON_JS, en = par.energy_recon == GKM_ENERGY_E;
    if (par.precision == GKM_PREC_FP64) {
        if (js && en) return launch_persist_t<double, StateD, 1, 1>(table, ncls, nsteps, s, grid_out);
        if (js) return launch_persist_t<double, StateD, 1, 0>(table, ncls, nsteps, s, grid_out);
        if (en) return launch_persist_t<double, StateD, 0, 1>(table, ncls, nsteps, s, grid_out);
        return launch_persist_t<double, StateD, 0, 0>(table, ncls, nsteps, s, grid_out);
    }
    if (js && en) return launch_persist_t<float, StateM, 1, 1>(table, ncls, nsteps, s, grid_out);
    if (js) return launch_persist_t<float, StateM, 1, 0>(table, ncls, nsteps, s, grid_out);
    if (en) return launch_persist_t<float, StateM, 0, 1>(table, ncls, nsteps, s, grid_out);
    return launch_persist_t<float, StateM, 0, 0>(table, ncls, nsteps, s, grid_out);
}

void launch_split(Ctx* h, int box, int q, const double* dsrc) {
    unsigned n = h->geo.ncell;
    const size_t off = (size_t)box * n;
    hipLaunchKernelGGL(k_split, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, n, dsrc, h->Whi[q] + off,
                       h->Wlo[q] + off);
}

void launch_merge(Ctx* h, int box, int q, double* ddst) {
    unsigned n = h->geo.ncell;
    const size_t off = (size_t)box * n;
    hipLaunchKernelGGL(k_merge, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, n, h->Whi[q] + off,
                       h->Wlo[q] + off, ddst);
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

void launch_gks_points(hipStream_t s, int prec, int n, const double* rec, double gam, double invPr, int K,
                       double* out) {
    if (prec == GKM_PREC_FP64)
        hipLaunchKernelGGL(k_gks_points<double>, dim3(blocks_for(n, 64)), dim3(64), 0, s, n, rec, gam, invPr, K, out);
    else
        hipLaunchKernelGGL(k_gks_points<float>, dim3(blocks_for(n, 64)), dim3(64), 0, s, n, rec, gam, invPr, K, out);
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
