// gkm_stats.hip — FP64 reductions over the interior: CFL signal speed and the
// turbulence statistics (replaces timestep, weno/src/update.cpp:1057-1079, the
// sums of PrintStats, update.cpp:638-784, and the GKM on-device TKE / rms
// reductions, GKM/CUDA GKM/src/GPUupdate.cu:425-483,1742-1783).
//
// Each pass: a grid-stride kernel accumulates per-thread FP64 values, reduces
// them across the 64-wide wavefront with shuffles and across the block in LDS,
// and writes one partial per block; a one-block kernel then reduces the
// partials in a fixed order, so the result does not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "gkm_internal.h"

namespace gkm {

constexpr int kRedThreads = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    return v;
}

// Reduce NQ per-thread values over the block; thread 0 writes the partials.
// Slots listed in MAXMASK are max-reduced, the others summed.
// Blocks of NW wavefronts.
template <int NQ, unsigned MAXMASK, int NW = kRedThreads / 64>
__device__ __forceinline__ void block_reduce_store(double (&acc)[NQ], double* part) {
    __shared__ double lds[NQ][NW];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        double v = ((MAXMASK >> q) & 1u) ? wave_max(acc[q]) : wave_sum(acc[q]);
        if (lane == 0) lds[q][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        const int q = threadIdx.x;
        double v = lds[q][0];
        for (int w = 1; w < NW; w++) v = ((MAXMASK >> q) & 1u) ? fmax(v, lds[q][w]) : v + lds[q][w];
        part[(long)blockIdx.x * NQ + q] = v;
    }
}

template <class St>
__device__ __forceinline__ void vel(const St& st, long c, double u[3]) {
    const double r = st.get(0, c);
    u[0] = st.get(1, c) / r;
    u[1] = st.get(2, c) / r;
    u[2] = st.get(3, c) / r;
}

struct StatPar {
    double gam, R, mu0, T0, visc_exp, dx;
    double iR, iT0;  // 1/R, 1/T0
    int visc_law;
};

__device__ __forceinline__ double visc(const StatPar& p, double T) {
    if (p.visc_law == 1) return p.mu0 * pow(T / p.T0, p.visc_exp);
    return p.mu0 * pow(T / p.T0, 1.5) * (p.T0 + 110.4) / (T + 110.4);
}

// pass A: slot layout
//  0 rho  1 p  2 T  3 1/rho  4 c  5 |u|^2  6 rho|u|^2  7..11 W0..W4
// 12..14 (du/dx)^2,3,4  15..17 (dv/dy)^n  18..20 (dw/dz)^n
// 21 |curl u|^2  22 (div u)^2  23 mu|curl u|^2  24 mu (div u)^2
// 25 mu du_i/dx_j du_i/dx_j  26 S_ij S_ij (PrintStats form)  27 max(|u|+c)
template <class St>
__global__ void __launch_bounds__(kRedThreads) k_stats_a(Geo g, St st, StatPar p, double* part) {
    double acc[kStatA];
#pragma unroll
    for (int q = 0; q < kStatA; q++) acc[q] = 0.0;
    const long n = (long)g.nx * g.ny * g.nz;
    const double i2dx = 1.0 / (2.0 * p.dx);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int k = (int)(t % g.nz);
        const long r = t / g.nz;
        const int i = (int)(r % g.nx);
        const int j = (int)(r / g.nx);
        const long c = (long)(j + 3) * g.s0 + (long)(i + 3) * g.s1 + (k + 3);
        const double rho = st.get(0, c);
        double u[3];
        vel(st, c, u);
        const double q2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        const double P = (rho * (p.gam - 1)) * (st.get(4, c) / rho - 0.5 * q2);
        const double T = P / (p.R * rho);
        const double cs = sqrt(p.gam * P / rho);
        acc[0] += rho; acc[1] += P; acc[2] += T; acc[3] += 1.0 / rho; acc[4] += cs;
        acc[5] += q2; acc[6] += rho * q2;
#pragma unroll
        for (int q = 0; q < 5; q++) acc[7 + q] += st.get(q, c);
        double d[3][3];
        const long sa[3] = {g.s1, g.s0, 1};  // d/dx, d/dy, d/dz
#pragma unroll
        for (int a = 0; a < 3; a++) {
            double up[3], um[3];
            vel(st, c + sa[a], up);
            vel(st, c - sa[a], um);
#pragma unroll
            for (int m = 0; m < 3; m++) d[m][a] = (up[m] - um[m]) * i2dx;
        }
        const double mu = visc(p, T);
        const double dudx = d[0][0], dudy = d[0][1], dudz = d[0][2];
        const double dvdx = d[1][0], dvdy = d[1][1], dvdz = d[1][2];
        const double dwdx = d[2][0], dwdy = d[2][1], dwdz = d[2][2];
        const double ox = dwdy - dvdz, oy = dudz - dwdx, oz = dvdx - dudy;
        const double om2 = ox * ox + oy * oy + oz * oz;
        const double th = dudx + dvdy + dwdz;
        double gg = 0.0;
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) gg += d[a][b] * d[a][b];
        double x2 = dudx * dudx; acc[12] += x2; acc[13] += x2 * dudx; acc[14] += x2 * x2;
        x2 = dvdy * dvdy; acc[15] += x2; acc[16] += x2 * dvdy; acc[17] += x2 * x2;
        x2 = dwdz * dwdz; acc[18] += x2; acc[19] += x2 * dwdz; acc[20] += x2 * x2;
        acc[21] += om2; acc[22] += th * th; acc[23] += mu * om2; acc[24] += mu * th * th; acc[25] += mu * gg;
        acc[26] += dudx * dudx + 0.5 * (dudy + dvdx) * (dudy + dvdx) + 0.5 * (dudz + dwdx) * (dudz + dwdx) +
                   0.5 * (dwdy + dvdz) * (dwdy + dvdz) + dvdy * dvdy + dwdz * dwdz;
        acc[27] = fmax(acc[27], sqrt(q2) + cs);
    }
    block_reduce_store<kStatA, (1u << 27)>(acc, part);
}

// pass B: squared deviations of rho, p, T, 1/rho about the given means
template <class St>
__global__ void __launch_bounds__(kRedThreads) k_stats_b(Geo g, St st, StatPar p, double m0, double m1, double m2,
                                                         double m3, double* part) {
    double acc[kStatB] = {0.0, 0.0, 0.0, 0.0};
    const long n = (long)g.nx * g.ny * g.nz;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int k = (int)(t % g.nz);
        const long r = t / g.nz;
        const int i = (int)(r % g.nx);
        const int j = (int)(r / g.nx);
        const long c = (long)(j + 3) * g.s0 + (long)(i + 3) * g.s1 + (k + 3);
        const double rho = st.get(0, c);
        double u[3];
        vel(st, c, u);
        const double q2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        const double P = (rho * (p.gam - 1)) * (st.get(4, c) / rho - 0.5 * q2);
        const double T = P / (p.R * rho);
        acc[0] += (rho - m0) * (rho - m0);
        acc[1] += (P - m1) * (P - m1);
        acc[2] += (T - m2) * (T - m2);
        acc[3] += (1.0 / rho - m3) * (1.0 / rho - m3);
    }
    block_reduce_store<kStatB, 0u>(acc, part);
}

// final pass: one block (kRedThreads) reduces nb partial rows of NQ slots in
// a fixed order: thread t takes rows t, t + kRedThreads, ... in turn, all
// slots in one sweep over the rows; each slot then goes through the wavefront
// sums and the wavefronts' values in order.  Thread q < NQ writes res[q].
template <int NQ, unsigned MAXMASK>
__device__ __forceinline__ void reduce_rows(int nb, const double* __restrict__ part, double* res) {
    double v[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) v[q] = 0.0;
    for (int b = threadIdx.x; b < nb; b += kRedThreads) {
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const double x = part[(long)b * NQ + q];
            v[q] = ((MAXMASK >> q) & 1u) ? fmax(v[q], x) : v[q] + x;
        }
    }
    __shared__ double lds[NQ][kRedThreads / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const double w = ((MAXMASK >> q) & 1u) ? wave_max(v[q]) : wave_sum(v[q]);
        if (lane == 0) lds[q][wv] = w;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        const int q = threadIdx.x;
        double t = lds[q][0];
        for (int w = 1; w < kRedThreads / 64; w++) t = ((MAXMASK >> q) & 1u) ? fmax(t, lds[q][w]) : t + lds[q][w];
        res[q] = t;
    }
}

template <int NQ, unsigned MAXMASK>
__global__ void __launch_bounds__(kRedThreads) k_reduce_final(int nb, const double* part, double* out) {
    reduce_rows<NQ, MAXMASK>(nb, part, out);
}

static StatPar stat_par(const Ctx* h, int box) {
    const gkm_params& q = h->bpar[box];
    StatPar p;
    p.gam = q.gamma;
    p.R = q.R;
    p.mu0 = q.mu0;
    p.T0 = q.T0;
    p.visc_exp = q.visc_exp;
    p.dx = h->grid.dx;
    p.iR = 1.0 / q.R;
    p.iT0 = 1.0 / q.T0;
    p.visc_law = q.visc_law;
    return p;
}

hipError_t launch_stats_a(Ctx* h, int box, double* host_out) {
    const Geo& g = h->geo;
    const StatPar p = stat_par(h, box);
    const int nb = h->red_blocks;
    const size_t off = (size_t)box * h->ncell;
    if (h->par.precision == GKM_PREC_FP64) {
        StateD st;
        for (int q = 0; q < 5; q++) st.w[q] = h->Wd[q] + off;
        hipLaunchKernelGGL(k_stats_a<StateD>, dim3(nb), dim3(kRedThreads), 0, h->stream, g, st, p, h->red_part);
    } else {
        StateM st;
        for (int q = 0; q < 5; q++) {
            st.hi[q] = h->Whi[q] + off;
            st.lo[q] = h->Wlo[q] + off;
        }
        hipLaunchKernelGGL(k_stats_a<StateM>, dim3(nb), dim3(kRedThreads), 0, h->stream, g, st, p, h->red_part);
    }
    hipLaunchKernelGGL((k_reduce_final<kStatA, (1u << 27)>), dim3(1), dim3(kRedThreads), 0, h->stream, nb,
                       h->red_part, h->red_out);
    hipError_t e = hipMemcpyAsync(host_out, h->red_out, sizeof(double) * kStatA, hipMemcpyDeviceToHost, h->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(h->stream);
}

// pass B into out (device), on stream s
static void launch_stats_b_dev(Ctx* h, int box, const double m[4], hipStream_t s, double* out) {
    const Geo& g = h->geo;
    const StatPar p = stat_par(h, box);
    const int nb = h->red_blocks;
    const size_t off = (size_t)box * h->ncell;
    if (h->par.precision == GKM_PREC_FP64) {
        StateD st;
        for (int q = 0; q < 5; q++) st.w[q] = h->Wd[q] + off;
        hipLaunchKernelGGL(k_stats_b<StateD>, dim3(nb), dim3(kRedThreads), 0, s, g, st, p, m[0], m[1], m[2], m[3],
                           h->red_part);
    } else {
        StateM st;
        for (int q = 0; q < 5; q++) {
            st.hi[q] = h->Whi[q] + off;
            st.lo[q] = h->Wlo[q] + off;
        }
        hipLaunchKernelGGL(k_stats_b<StateM>, dim3(nb), dim3(kRedThreads), 0, s, g, st, p, m[0], m[1], m[2], m[3],
                           h->red_part);
    }
    hipLaunchKernelGGL((k_reduce_final<kStatB, 0u>), dim3(1), dim3(kRedThreads), 0, s, nb, h->red_part, out);
}

hipError_t launch_stats_b(Ctx* h, int box, const double m[4], double* host_out) {
    launch_stats_b_dev(h, box, m, h->stream, h->red_out);
    hipError_t e = hipMemcpyAsync(host_out, h->red_out, sizeof(double) * kStatB, hipMemcpyDeviceToHost, h->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(h->stream);
}

// ---------------------------------------------------------------------------
// PrintStats with spline gradients (weno/src/update.cpp:638-784).
//
// Along every grid line of each axis (through interior nodes of the other
// two), the slopes D of the natural cubic spline through u, v, w =
// (rho u, rho v, rho w)/rho at the line's m padded nodes:
//   2 D_0 + D_1 = (3/h)(f_1 - f_0)
//   D_{i-1} + 4 D_i + D_{i+1} = (3/h)(f_{i+1} - f_{i-1})
//   D_{m-2} + 2 D_{m-1} = (3/h)(f_{m-1} - f_{m-2})
// The matrix is the same for every line of an axis, so the Thomas pivots
// inv_i = 1/(b_i - inv_{i-1}) come precomputed: forward y_i = (r_i - y_{i-1})
// inv_i, backward D_i = y_i - inv_i D_{i+1}.  Only interior nodes' slopes are
// used.
//
// One pass per axis.  z lines, contiguous, are solved whole in LDS when they
// have up to kSplMaxM nodes (boxes up to 335 cells a side); x and y lines are
// streamed, a lane per line and component (k_spl_stream).  Each slope is added
// to the sums where it is made (k_spl_lines_sums for z, which also takes the
// sums that need only the state, then the x and y passes): only mu(T) and the
// three slopes of the S_ij cross terms that a later pass pairs with its own
// (dudz, dvdz, dvdx) cross HBM as planes, and each block writes one row of
// partial sums that k_pstats_final reduces.  Longer z lines go in chunks
// through the gradient planes (k_spl_chunked); then every axis writes its
// three planes ((3 A + comp) * ncell; A = 0 y, 1 x, 2 z) and k_pstats forms
// the sums from the nine planes and the state.
// ---------------------------------------------------------------------------
constexpr int kSpLines = 64, kSpThreads = 192, kSpChunk = 16;

// density, |u|^2, pressure and temperature of cell c
template <class St>
__device__ __forceinline__ void cell_pt(const St& st, unsigned c, const StatPar& p, double& rho, double& q2,
                                        double& P, double& T) {
    rho = st.get(0, c);
    double u[3];
    vel(st, c, u);
    q2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    P = (rho * (p.gam - 1)) * (st.get(4, c) / rho - 0.5 * q2);
    T = P / (p.R * rho);
}

// PrintStats sums over the interior (update.cpp:658-707):
//  0 rho|u|^2  1 mu du_i/dx_j du_i/dx_j  2 S_ij S_ij  3..5 (du/dx)^2,3,4
//  6..8 (dv/dy)^n  9..11 (dw/dz)^n  12 p  13 T  14 1/rho
// and, for the rms of p, T and 1/rho (update.cpp:743-770 takes them about the
// means in a second sweep), the sums of x - K (15..17) and (x - K)^2 (18..20)
// about a shift K, the values of one fixed cell: sum (x - mean)^2 = S2 - S1^2/N
// in exact arithmetic, and with K inside the range of the data the rounding
// stays at that of the two-sweep form, so one sweep over the state does.
// dX[m] = d(u_m)/dx, dY[m] = d(u_m)/dy, dZ[m] = d(u_m)/dz
constexpr int kPSum = kPStat + 6;
template <class St>
__device__ __forceinline__ void pstat_add(const St& st, unsigned c, const StatPar& p, const double K[3],
                                          const double dX[3], const double dY[3], const double dZ[3],
                                          double (&acc)[kPSum]) {
    double rho, q2, P, T;
    cell_pt(st, c, p, rho, q2, P, T);
    const double dudx = dX[0], dudy = dY[0], dudz = dZ[0];
    const double dvdx = dX[1], dvdy = dY[1], dvdz = dZ[1];
    const double dwdx = dX[2], dwdy = dY[2], dwdz = dZ[2];
    const double mu = visc(p, T);
    acc[0] += rho * q2;
    acc[1] += mu * (dudx * dudx + dudy * dudy + dudz * dudz + dvdx * dvdx + dvdy * dvdy + dvdz * dvdz +
                    dwdx * dwdx + dwdy * dwdy + dwdz * dwdz);
    acc[2] += dudx * dudx + 0.5 * (dudy + dvdx) * (dudy + dvdx) + 0.5 * (dudz + dwdx) * (dudz + dwdx) +
              0.5 * (dwdy + dvdz) * (dwdy + dvdz) + dvdy * dvdy + dwdz * dwdz;
    double x2 = dudx * dudx; acc[3] += x2; acc[4] += x2 * dudx; acc[5] += x2 * x2;
    x2 = dvdy * dvdy; acc[6] += x2; acc[7] += x2 * dvdy; acc[8] += x2 * x2;
    x2 = dwdz * dwdz; acc[9] += x2; acc[10] += x2 * dwdz; acc[11] += x2 * x2;
    const double iv = 1.0 / rho;
    acc[12] += P; acc[13] += T; acc[14] += iv;
    const double dp = P - K[0], dT = T - K[1], dv = iv - K[2];
    acc[15] += dp; acc[16] += dT; acc[17] += dv;
    acc[18] += dp * dp; acc[19] += dT * dT; acc[20] += dv * dv;
}

// component q of the velocity at c
template <class St> __device__ __forceinline__ double velq(const St& st, unsigned c, int q) {
    return st.get(1 + q, c) / st.get(0, c);
}

// z lines run along the unit stride, so a lane per line would touch a
// different cache line on every lane.  The block walks its 64 lines in chunks
// of 16 nodes staged through LDS: all 192 threads load and store 16-node row
// segments (coalesced), and lane l of wavefront q runs the recursion of
// component q of line l out of the tile.
template <class St>
__device__ __forceinline__ void spline_z_block(const Geo& g, const St& st, const double* __restrict__ inv,
                                               double r3h, double* __restrict__ grad, int b,
                                               double (&tile)[3][kSpLines][kSpChunk + 1]) {
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int m = g.P2;
    const int nline = g.ny * g.nx;
    const int line0 = b * kSpLines;
    double* G[3] = {grad + (size_t)6 * g.ncell, grad + (size_t)7 * g.ncell, grad + (size_t)8 * g.ncell};
    auto line_base = [&](int id) -> unsigned {
        const unsigned j = (unsigned)(id / g.nx) + 3, i = (unsigned)(id % g.nx) + 3;
        return j * g.s0 + i * g.s1;
    };
    const bool mine = line0 + lane < nline;
    // cooperative pass over the block's lines x chunk nodes [n0, n0 + kSpChunk):
    // element e covers line e / kSpChunk, node e % kSpChunk
    auto coop = [&](int n0, auto&& body) {
        for (int e = threadIdx.x; e < kSpLines * kSpChunk; e += kSpThreads) {
            const int l = e / kSpChunk, t = e % kSpChunk;
            const int n = n0 + t;
            if (line0 + l < nline && n < m) body(l, t, line_base(line0 + l) + (unsigned)n);
        }
    };
    double f0 = mine ? velq(st, line_base(line0 + lane), q) : 0.0, fm = f0, y = 0.0;
    const int nch = (m + kSpChunk - 1) / kSpChunk;
    // forward: the tile holds f at nodes n0+1 .. n0+16; lane l walks nodes
    // n0 .. n0+15 of its line and leaves y_n where it read f_{n+1}
    for (int ch = 0; ch < nch; ch++) {
        const int n0 = ch * kSpChunk;
        coop(n0 + 1, [&](int l, int t, unsigned c) {
            const double r = st.get(0, c);
            tile[0][l][t] = st.get(1, c) / r;
            tile[1][l][t] = st.get(2, c) / r;
            tile[2][l][t] = st.get(3, c) / r;
        });
        __syncthreads();
        for (int t = 0; t < kSpChunk; t++) {
            const int n = n0 + t;
            if (n >= m) break;
            const double fp = tile[q][lane][t];  // node n+1 (not used at n = m-1)
            const double hi = n == m - 1 ? f0 : fp;
            const double lo = n == 0 ? f0 : fm;
            y = (r3h * (hi - lo) - y) * inv[n];
            tile[q][lane][t] = y;
            fm = f0;
            f0 = fp;
        }
        __syncthreads();
        coop(n0, [&](int l, int t, unsigned c) {
            put(G[0], c, tile[0][l][t]);
            put(G[1], c, tile[1][l][t]);
            put(G[2], c, tile[2][l][t]);
        });
        __syncthreads();
    }
    // backward from D_{m-1} = y_{m-1}: D_n = y_n - inv_n D_{n+1}
    double d = y;
    for (int ch = nch - 1; ch >= 0; ch--) {
        const int n0 = ch * kSpChunk;
        coop(n0, [&](int l, int t, unsigned c) {
            tile[0][l][t] = ld(G[0], c);
            tile[1][l][t] = ld(G[1], c);
            tile[2][l][t] = ld(G[2], c);
        });
        __syncthreads();
        for (int t = kSpChunk - 1; t >= 0; t--) {
            const int n = n0 + t;
            if (n >= m - 1) continue;  // node m-1 keeps y_{m-1}
            d = tile[q][lane][t] - inv[n] * d;
            tile[q][lane][t] = d;
        }
        __syncthreads();
        coop(n0, [&](int l, int t, unsigned c) {
            put(G[0], c, tile[0][l][t]);
            put(G[1], c, tile[1][l][t]);
            put(G[2], c, tile[2][l][t]);
        });
        __syncthreads();
    }
}

// chunked pass over the z lines: blocks of 64 lines x 3 components
template <class St>
__global__ void __launch_bounds__(kSpThreads) k_spl_chunked(Geo g, St st, const double* __restrict__ inv,
                                                            double r3h, double* __restrict__ grad) {
    __shared__ double tile[3][kSpLines][kSpChunk + 1];
    spline_z_block(g, st, inv, r3h, grad, (int)blockIdx.x, tile);
}

// ---- whole z lines in LDS ----------------------------------------------------
// For z lines of up to kSplMaxM nodes (boxes up to 335 cells a side) a block
// holds a group of kSplL whole lines, all three components, in LDS:
//   load:  every thread loads state and forms u, v, w at the group's nodes
//   walk:  lane (component c, line l), 24 lanes, runs the forward and the
//          backward recursion in place (f -> y -> D), out of LDS
//   out:   the interior slopes go to the axis's planes.
// Nothing but the state and the slopes crosses HBM (the chunked kernel above
// also stores and reloads y), and every load and store runs along the lines,
// which are contiguous.
constexpr int kSplL = 8, kSplThreads = 192;
constexpr int kSplLoadLines = 4;  // lines whose state words are loaded together
constexpr int kSplMaxM = 341;  // 3 kSplL m doubles within 64 KB

// line t of axis A (lines over the two other axes' interiors, the faster
// inner: z for the x and y lines)
template <int A> struct SplAxis {
    unsigned n_in, n_out, stride;
    int m;
    __device__ SplAxis(const Geo& g)
        : n_in(A == 2 ? g.nx : g.nz), n_out(A == 0 ? g.nx : g.ny),
          stride(A == 0 ? g.s0 : (A == 1 ? g.s1 : 1u)), m(A == 0 ? g.P0 : (A == 1 ? g.P1 : g.P2)) {}
    __device__ unsigned base(const Geo& g, unsigned t) const {
        const unsigned a = t % n_in + 3, b = t / n_in + 3;
        return A == 0 ? b * g.s1 + a : (A == 1 ? b * g.s0 + a : b * g.s0 + a * g.s1);
    }
};

// Elements (line l, node n) of a group as the 192 threads share them: thread
// t takes node t + 192 k of all 8 lines.
//
// spl_load: the velocities of the group's nodes into the tile.  A batch's
// loads (one node of kSplLoadLines lines) are all issued before its divisions
// and LDS writes, with no branch around an element: a missing line (the last
// group's) repeats the group's last line, whose slot is never read.
template <class St, int A>
__device__ __forceinline__ void spl_load(const Geo& g, const St& st, const SplAxis<A>& ax, unsigned t0,
                                         unsigned nl, double* tile) {
    const int m = ax.m, t = threadIdx.x;
    // the batch's raw words (the mixed state's hi and lo floats) are loaded
    // into registers first and only then combined: combining each value as
    // it is loaded had the compiler wait for every pair of loads in turn
    using Raw = typename std::conditional<std::is_same<St, StateM>::value, float, double>::type;
    auto raw = [&](int q, unsigned c, Raw& h, Raw& l) {
        if constexpr (std::is_same<St, StateM>::value) {
            h = ld(st.hi[q], c);
            l = ld(st.lo[q], c);
        } else {
            h = ld(st.w[q], c);
            l = Raw(0);
        }
    };
    auto put_f = [&](int l, int n, const double (&v)[4]) {
        // one FP64 division instead of three (PrintStats 1.111 -> 1.102 ms at 256^3)
        const double iv = 1.0 / v[0];
        tile[(0 * m + n) * kSplL + l] = v[1] * iv;
        tile[(1 * m + n) * kSplL + l] = v[2] * iv;
        tile[(2 * m + n) * kSplL + l] = v[3] * iv;
    };
    {
        static_assert(A == 2, "x and y lines go through k_spl_stream");
        static_assert(kSplL % kSplLoadLines == 0, "whole batches");
        unsigned b[kSplL];
#pragma unroll
        for (int l = 0; l < kSplL; l++) b[l] = ax.base(g, min(t0 + l, nl - 1));
        for (int n = t; n < m; n += kSplThreads) {
#pragma unroll
            for (int l0 = 0; l0 < kSplL; l0 += kSplLoadLines) {
                Raw h[kSplLoadLines][4], lo[kSplLoadLines][4];
#pragma unroll
                for (int l = 0; l < kSplLoadLines; l++)
#pragma unroll
                    for (int q = 0; q < 4; q++) raw(q, b[l0 + l] + (unsigned)n, h[l][q], lo[l][q]);
#pragma unroll
                for (int l = 0; l < kSplLoadLines; l++) {
                    double v[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) v[q] = (double)h[l][q] + (double)lo[l][q];
                    put_f(l0 + l, n, v);
                }
            }
        }
    }
}

// The group's velocities into tile[(c m + n) kSplL + l] (c = component),
// then the first 48 lanes solve the lines in place, two lanes a line and
// component, while the wavefronts past the first run side().
//
// The solve is the twisted factorization of the line's system: the lane of
// the upper half eliminates rows 0 .. k-1 downward (the Thomas forward sweep,
// y_i = (r_i - y_{i-1}) inv_i), the lane of the lower half rows m-1 .. k+1
// upward (z_i = (r_i - z_{i+1}) inv'_i), k = m/2; row k then gives
//   D_k = (r_k - y_{k-1} - z_{k+1}) / (4 - inv_{k-1} - inv'_{k+1}),
// and each lane substitutes back outward from it (D_i = y_i - inv_i D_{i+1},
// D_i = z_i - inv'_i D_{i-1}).  The system reads the same from either end, so
// inv'_i = inv_{m-1-i}: both lanes run the same recursions in their own
// local node order (j = i upward, j = m-1-i downward, where r changes sign),
// and the serial chain is half the whole-line walk's.  Neither lane writes a
// node the other reads: the upper one reads f of nodes 0 .. k and writes
// 0 .. k, the lower one reads k .. m-1 and writes k+1 .. m-1.  The values
// differ from the whole-line walk's by rounding only.
template <class St, int A, class Side>
__device__ __forceinline__ void spl_group_solve(const Geo& g, const St& st, const SplAxis<A>& ax, unsigned t0,
                                                unsigned nl, const double* __restrict__ inv, double r3h,
                                                double* tile, Side&& side) {
    static_assert(6 * kSplL <= 64, "the solving lanes fit one wavefront");
    const int m = ax.m;
    spl_load(g, st, ax, t0, nl, tile);
    __syncthreads();
    // the wavefront's number as a scalar: the branch is taken per wavefront
    if (__builtin_amdgcn_readfirstlane(threadIdx.x / 64u) != 0u) {
        side();  // the other wavefronts' work, beside the solve (which reads and writes only the tile)
    } else if (threadIdx.x < 6 * kSplL) {
        const int lane = threadIdx.x, hf = lane / (3 * kSplL), cl = lane % (3 * kSplL);
        const int c = cl / kSplL, l = cl % kSplL;
        const int k = m / 2;          // the twist row
        const int nc = m - 1 - k;     // rows both halves eliminate (the upper one one more when m is even)
        const int st = hf ? -kSplL : kSplL;
        double* T = tile + (size_t)c * m * kSplL + l + (hf ? (size_t)(m - 1) * kSplL : 0);  // local j at T[j st]
        const double r3 = hf ? -r3h : r3h;
        constexpr int U = 8;  // LDS reads of U nodes ahead of the dependent steps
        // elimination: local row 0 (one-sided), then the central rows
        double fm = T[0], f0 = T[st];
        double y = (r3 * (f0 - fm)) * inv[0];
        T[0] = y;
        int j = 1;
        for (; j + U <= nc; j += U) {  // f of local nodes j+1 .. j+U, all <= nc
            double fn[U];
#pragma unroll
            for (int u = 0; u < U; u++) fn[u] = T[(j + 1 + u) * st];
#pragma unroll
            for (int u = 0; u < U; u++) {
                y = (r3 * (fn[u] - fm) - y) * inv[j + u];
                T[(j + u) * st] = y;
                fm = f0;
                f0 = fn[u];
            }
        }
        for (; j < nc; j++) {
            const double fp = T[(j + 1) * st];
            y = (r3 * (fp - fm) - y) * inv[j];
            T[j * st] = y;
            fm = f0;
            f0 = fp;
        }
        if (k > nc && hf == 0) {  // the upper half's row k-1 (m even)
            const double fp = T[(nc + 1) * st];
            y = (r3 * (fp - fm) - y) * inv[nc];
            T[nc * st] = y;
            fm = f0;
        }
        // each lane now holds f and its elimination's value at its last row,
        // k-1 (upper) or k+1 (lower): row k from both, the same bits in both lanes
        const int partner = hf ? lane - 3 * kSplL : lane + 3 * kSplL;
        const double fo = __shfl(fm, partner, 64), yo = __shfl(y, partner, 64);
        const double fa = hf ? fo : fm, ya = hf ? yo : y;  // row k-1
        const double fb = hf ? fm : fo, zb = hf ? y : yo;  // row k+1
        double d = (r3h * (fb - fa) - ya - zb) * (1.0 / (4.0 - inv[k - 1] - inv[m - 2 - k]));
        if (hf == 0) T[k * st] = d;
        // back-substitution outward from row k down to local node 3
        if (k > nc && hf == 0) {
            d = T[nc * st] - inv[nc] * d;
            T[nc * st] = d;
        }
        int n = nc - 1;
        for (; n - (U - 1) >= 3; n -= U) {
            double yv[U];
#pragma unroll
            for (int u = 0; u < U; u++) yv[u] = T[(n - u) * st];
#pragma unroll
            for (int u = 0; u < U; u++) {
                d = yv[u] - inv[n - u] * d;
                T[(n - u) * st] = d;
            }
        }
        for (; n >= 3; n--) {
            d = T[n * st] - inv[n] * d;
            T[n * st] = d;
        }
    }
    __syncthreads();
}

// x (A = 1) and y (A = 0) lines by one streaming sweep per line and velocity
// component, with no LDS tile.  The forward recursion runs node by node as in
// the whole-line walk (the same expressions, so the same y).  The backward one
// is unrolled: D_i = y_i - inv_i D_{i+1} with D_{m-1} = y_{m-1} gives
//   D_i = sum_j W[i][j] y_{i+j},  W[i][j] = (-1)^j inv_i ... inv_{i+j-1}
// (j up to m-1-i), weights made on the host.  Past the first nodes the pivots
// are 2 - sqrt(3) = 0.268, so W[i][j] falls like 0.268^j, and the sum is cut
// at j < kSplRing = 32 (0.268^32 = 5e-19 of the leading term, below FP64
// rounding); near the line's end it is whole.  So D_i is ready once y_{i+31}
// is: a lane keeps the last 32 y in registers and writes D_i then.  Lanes run
// along z, so every load and store is a full row; wavefront q of a block
// takes velocity component q of the block's 64 lines.  The values differ from
// the two-walk ones by rounding only.
constexpr int kSplAhead = 4;
constexpr int kSplStreamThreads = 192;  // 64 lines x 3 velocity components
template <class St> struct SplRaw;  // raw words of rho and rho u_q at one node
template <> struct SplRaw<StateM> {
    float h0, l0, hq, lq;
    __device__ __forceinline__ void load(const StateM& st, int q, unsigned c) {
        h0 = ld(st.hi[0], c);
        l0 = ld(st.lo[0], c);
        hq = ld(st.hi[1 + q], c);
        lq = ld(st.lo[1 + q], c);
    }
    __device__ __forceinline__ double f() const { return ((double)hq + (double)lq) / ((double)h0 + (double)l0); }
};
template <> struct SplRaw<StateD> {
    double w0, wq;
    __device__ __forceinline__ void load(const StateD& st, int q, unsigned c) {
        w0 = ld(st.w[0], c);
        wq = ld(st.w[1 + q], c);
    }
    __device__ __forceinline__ double f() const { return wq / w0; }
};

// The slope planes the fused passes exchange (planes of h->spl_grad): mu(T)
// and dudz, dvdz from the z pass, dvdx from the x pass.
enum : int { kPlMu, kPlDudz, kPlDvdz, kPlDvdx };

// What one slope d = d(u_q)/d(axis A) adds to the PrintStats sums (slot
// layout of pstat_add), given mu and the partner slope of its S_ij cross term:
//   mu sum_ij (du_i/dx_j)^2   each slope once, mu d^2 (slot 1)
//   S_ij S_ij                 (du/dx)^2 + (dv/dy)^2 + (dw/dz)^2 whole, the
//                             off-diagonal squares halved, and the three
//                             cross products dudy dvdx, dudz dwdx, dwdy dvdz
//                             added by the pass that holds the second factor
//                             (y pass: u with dvdx, w with dvdz; x pass: w
//                             with dudz) (slot 2)
//   (du/dx)^n, (dv/dy)^n, (dw/dz)^n, n = 2, 3, 4   (slots 3..11)
struct SlopeSums {
    double mu_g2 = 0.0, sij = 0.0, d2 = 0.0, d3 = 0.0, d4 = 0.0;
    __device__ __forceinline__ void add(double d, double mu, bool diag, double partner) {
        const double d2v = d * d;
        mu_g2 += mu * d2v;
        sij += diag ? d2v : 0.5 * d2v + d * partner;
        if (diag) {
            d2 += d2v;
            d3 += d2v * d;
            d4 += d2v * d2v;
        }
    }
    // into the kPSum slots; diag_slot = 3, 6 or 9 for the diagonal slope's moments
    __device__ __forceinline__ void put(double (&acc)[kPSum], int diag_slot) const {
        acc[1] += mu_g2;
        acc[2] += sij;
        if (diag_slot == 3) { acc[3] += d2; acc[4] += d3; acc[5] += d4; }
        if (diag_slot == 6) { acc[6] += d2; acc[7] += d3; acc[8] += d4; }
        if (diag_slot == 9) { acc[9] += d2; acc[10] += d3; acc[11] += d4; }
    }
};

// SUMS = false: the slopes go to the axis's three planes for k_pstats.
// SUMS = true (the fused PrintStats): each slope is added to the sums where it
// is made, and only the planes a later pass needs are written (x pass:
// dvdx); each block writes one row of kPSum partial sums to `part`.
template <class St, int A, bool SUMS>
__global__ void __launch_bounds__(kSplStreamThreads) k_spl_stream(Geo g, St st, const double* __restrict__ inv,
                                                   const double* __restrict__ wt, double r3h,
                                                   double* __restrict__ grad, double* __restrict__ part) {
    static_assert(A != 2, "z lines go through k_spl_lines_sums or k_spl_chunked");
    static_assert(kSplRing % kSplAhead == 0, "queue slots fixed within the unrolled ring");
    const SplAxis<A> ax(g);
    const unsigned nl = ax.n_in * ax.n_out;
    const unsigned t = blockIdx.x * 64u + (threadIdx.x & 63u);
    const int q = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));  // this wavefront's component
    SlopeSums sums;
    if (!SUMS && t >= nl) return;
    // with SUMS every thread reaches the block reduction below
    if (t < nl) {
    const int m = ax.m;
    const unsigned b = ax.base(g, t), s = ax.stride;
    double* G = grad + (size_t)(3 * A + q) * g.ncell;
    // raw words of nodes n+1 .. n+kSplAhead, node k in slot k % kSplAhead
    SplRaw<St> Q[kSplAhead];
    SplRaw<St> r0;
    r0.load(st, q, b);
#pragma unroll
    for (int k = 1; k <= kSplAhead; k++) Q[k % kSplAhead].load(st, q, b + (unsigned)min(k, m - 1) * s);
    double fm = r0.f(), f0 = fm, fp = fm, y = 0.0;  // f_{n-1}, f_n, f_{n+1}; y_{-1} = 0
    double ring[kSplRing];  // y_k in slot k % kSplRing
#pragma unroll
    for (int j = 0; j < kSplRing; j++) ring[j] = 0.0;
    for (int n0 = 0; n0 <= m + kSplRing - 5; n0 += kSplRing) {
#pragma unroll
        for (int u = 0; u < kSplRing; u++) {
            const int n = n0 + u;
            if (n < m) {
                if (n + 1 < m) {
                    fp = Q[(u + 1) % kSplAhead].f();
                    const int k = n + 1 + kSplAhead;
                    if (k < m) Q[(u + 1) % kSplAhead].load(st, q, b + (unsigned)k * s);
                }
                // one-sided differences at the two ends, central inside
                const double hi = n == m - 1 ? f0 : fp;
                const double lo = n == 0 ? f0 : fm;
                y = (r3h * (hi - lo) - y) * inv[n];
                ring[u] = y;
                fm = f0;
                f0 = fp;
            }
            const int i = n - (kSplRing - 1);  // the node whose window is now complete
            if (i >= 3 && i <= m - 4) {
                const double* w = wt + (size_t)i * kSplRing;
                double d = 0.0;
#pragma unroll
                for (int j = kSplRing - 1; j >= 0; j--) d = fma(w[j], ring[(u + 1 + j) % kSplRing], d);
                const unsigned c = b + (unsigned)i * s;
                if constexpr (!SUMS) {
                    put(G, c, d);
                } else {
                    const double mu = ld(grad + (size_t)kPlMu * g.ncell, c);
                    if constexpr (A == 1) {  // x: d = d(u_q)/dx
                        if (q == 1) put(grad + (size_t)kPlDvdx * g.ncell, c, d);
                        sums.add(d, mu, q == 0, q == 2 ? ld(grad + (size_t)kPlDudz * g.ncell, c) : 0.0);
                    } else {  // y: d = d(u_q)/dy
                        const double pt = q == 0   ? ld(grad + (size_t)kPlDvdx * g.ncell, c)
                                          : q == 2 ? ld(grad + (size_t)kPlDvdz * g.ncell, c)
                                                   : 0.0;
                        sums.add(d, mu, q == 1, pt);
                    }
                }
            }
        }
    }
    }
    if constexpr (SUMS) {
        double acc[kPSum];
#pragma unroll
        for (int k = 0; k < kPSum; k++) acc[k] = 0.0;
        sums.put(acc, A == 1 ? (q == 0 ? 3 : -1) : (q == 1 ? 6 : -1));
        block_reduce_store<kPSum, 0u, kSplStreamThreads / 64>(acc, part);
    }
}

// The five state words of a cell, loaded before they are combined (the
// mixed state's hi and lo floats, or the FP64 values).
template <class St> struct Raw5;
template <> struct Raw5<StateM> {
    float h[5], l[5];
    __device__ __forceinline__ void load(const StateM& st, unsigned c) {
#pragma unroll
        for (int q = 0; q < 5; q++) {
            h[q] = ld(st.hi[q], c);
            l[q] = ld(st.lo[q], c);
        }
    }
    __device__ __forceinline__ double get(int q) const { return (double)h[q] + (double)l[q]; }
};
template <> struct Raw5<StateD> {
    double w[5];
    __device__ __forceinline__ void load(const StateD& st, unsigned c) {
#pragma unroll
        for (int q = 0; q < 5; q++) w[q] = ld(st.w[q], c);
    }
    __device__ __forceinline__ double get(int q) const { return w[q]; }
};

// rho, |u|^2, p, T and 1/rho of a cell with one reciprocal of rho (the
// fused pass's form of cell_pt), and mu(T) with the power law as
// exp2(e log2(T/T0)) (the device library's pow is ~220 instructions)
template <class R>
__device__ __forceinline__ void cell_fast(const R& r, const StatPar& p, double& rho, double& q2, double& P,
                                          double& T, double& iv, double& mu) {
    rho = r.get(0);
    iv = 1.0 / rho;
    const double u = r.get(1) * iv, v = r.get(2) * iv, w = r.get(3) * iv;
    q2 = u * u + v * v + w * w;
    P = (rho * (p.gam - 1)) * (r.get(4) * iv - 0.5 * q2);
    T = P * p.iR * iv;
    const double x = T * p.iT0;
    mu = p.visc_law == 1 ? p.mu0 * exp2(p.visc_exp * log2(x))
                         : p.mu0 * x * sqrt(x) * (p.T0 + 110.4) / (T + 110.4);
}

// The fused z pass.  While the first wavefront solves a group's z lines, the
// other two take the sums that need only the state (rho|u|^2, p, T, 1/rho
// and their shifted sums) at the group's interior nodes and write mu(T) to
// its plane; then every thread adds the z slopes' terms, with mu read back
// from the plane, and writes dudz, dvdz to their planes for the x and y
// passes.  One row of kPSum partial sums per block.
constexpr int kSplBatch = 4;
template <class St>
__global__ void __launch_bounds__(kSplThreads) __attribute__((amdgpu_waves_per_eu(3, 3)))
    k_spl_lines_sums(Geo g, St st, const double* __restrict__ inv, double r3h, StatPar p, int gpb,
                     double* __restrict__ grad, double* __restrict__ part) {
    static_assert(kSplThreads > 64 && kSplThreads % 64 == 0, "a solving wavefront and side wavefronts");
    extern __shared__ double spl_tile[];
    const SplAxis<2> ax(g);
    const unsigned nl = ax.n_in * ax.n_out;
    const int m = ax.m, ni = m - 6;
    double acc[kPSum];
#pragma unroll
    for (int k = 0; k < kPSum; k++) acc[k] = 0.0;
    double K[3];  // the deviation sums' shift: p, T, 1/rho of the first interior cell
    {
        double rho, q2;
        cell_pt(st, 3 * g.s0 + 3 * g.s1 + 3, p, rho, q2, K[0], K[1]);
        K[2] = 1.0 / rho;
    }
    double* const mu_pl = grad + (size_t)kPlMu * g.ncell;
    SlopeSums sums;
    // gpb consecutive groups per block, the blocks of an XCD on neighbouring groups
    for (int grp = 0; grp < gpb; grp++) {
        const unsigned t0 = (xcd_block() * (unsigned)gpb + (unsigned)grp) * kSplL;
        if (t0 >= nl) break;  // block-uniform
        if (grp > 0) __syncthreads();  // the previous group's slopes have been read
        // the group's interior nodes, element e = (line e / ni, node 3 + e % ni),
        // consecutive threads on consecutive nodes
        const int ne = (int)min((unsigned)kSplL, nl - t0) * ni;
        auto elem = [&](int e) { return ax.base(g, t0 + (unsigned)(e / ni)) + (unsigned)(3 + e % ni); };
        // beside the solve: the state sums and mu(T), the state words of
        // kSplBatch elements loaded before any is combined
        auto state_sums = [&]() {
            constexpr int nt = kSplThreads - 64;
            for (int e0 = threadIdx.x - 64; e0 < ne; e0 += kSplBatch * nt) {
                Raw5<St> raw[kSplBatch];
                unsigned cc[kSplBatch];
#pragma unroll
                for (int k = 0; k < kSplBatch; k++) {
                    cc[k] = elem(min(e0 + k * nt, ne - 1));  // past the end: repeat the last, unused
                    raw[k].load(st, cc[k]);
                }
#pragma unroll
                for (int k = 0; k < kSplBatch; k++) {
                    if (e0 + k * nt >= ne) continue;
                    double rho, q2, P, T, iv, mu;
                    cell_fast(raw[k], p, rho, q2, P, T, iv, mu);
                    put(mu_pl, cc[k], mu);
                    acc[0] += rho * q2;
                    acc[12] += P; acc[13] += T; acc[14] += iv;
                    const double dp = P - K[0], dT = T - K[1], dv = iv - K[2];
                    acc[15] += dp; acc[16] += dT; acc[17] += dv;
                    acc[18] += dp * dp; acc[19] += dT * dT; acc[20] += dv * dv;
                }
            }
        };
        // its closing barrier also orders the mu stores before the loads below
        spl_group_solve(g, st, ax, t0, nl, inv, r3h, spl_tile, state_sums);
        for (int e0 = threadIdx.x; e0 < ne; e0 += kSplBatch * kSplThreads) {
            double mu[kSplBatch];
            unsigned cc[kSplBatch];
#pragma unroll
            for (int k = 0; k < kSplBatch; k++) {
                cc[k] = elem(min(e0 + k * kSplThreads, ne - 1));
                mu[k] = ld(mu_pl, cc[k]);
            }
#pragma unroll
            for (int k = 0; k < kSplBatch; k++) {
                const int e = e0 + k * kSplThreads;
                if (e >= ne) continue;
                const int l = e / ni, n = 3 + e % ni;
                const double dz[3] = {spl_tile[(0 * m + n) * kSplL + l], spl_tile[(1 * m + n) * kSplL + l],
                                      spl_tile[(2 * m + n) * kSplL + l]};
                put(grad + (size_t)kPlDudz * g.ncell, cc[k], dz[0]);
                put(grad + (size_t)kPlDvdz * g.ncell, cc[k], dz[1]);
                sums.add(dz[0], mu[k], false, 0.0);  // dudz: its cross term dudz dwdx is added by the x pass
                sums.add(dz[1], mu[k], false, 0.0);  // dvdz: dwdy dvdz by the y pass
                sums.add(dz[2], mu[k], true, 0.0);
            }
        }
    }
    sums.put(acc, 9);
    block_reduce_store<kPSum, 0u, kSplThreads / 64>(acc, part);
}

// The sums over the interior from the nine slope planes and the state,
// grid-stride over the interior cells (z fastest); one partial row per block.
template <class St>
__global__ void __launch_bounds__(kRedThreads) k_pstats(Geo g, St st, const double* __restrict__ grad, StatPar p,
                                                        FastDiv fnz, FastDiv fnx, double* part) {
    double acc[kPSum];
#pragma unroll
    for (int q = 0; q < kPSum; q++) acc[q] = 0.0;
    double K[3];  // the deviation sums' shift: p, T, 1/rho of the first interior cell
    {
        double rho, q2;
        cell_pt(st, 3 * g.s0 + 3 * g.s1 + 3, p, rho, q2, K[0], K[1]);
        K[2] = 1.0 / rho;
    }
    const unsigned n = (unsigned)g.nx * g.ny * g.nz;
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const unsigned r = fdiv(t, fnz), k = t - r * fnz.d;
        const unsigned j = fdiv(r, fnx), i = r - j * fnx.d;
        const unsigned c = (j + 3) * g.s0 + (i + 3) * g.s1 + (k + 3);
        double dX[3], dY[3], dZ[3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            dY[q] = ld(grad + (size_t)q * g.ncell, c);
            dX[q] = ld(grad + (size_t)(3 + q) * g.ncell, c);
            dZ[q] = ld(grad + (size_t)(6 + q) * g.ncell, c);
        }
        pstat_add(st, c, p, K, dX, dY, dZ, acc);
    }
    block_reduce_store<kPSum, 0u>(acc, part);
}

// The PrintStats row from nb partial rows: the kPStat sums, then the kStatB
// slots of pass B (rho: 0, not reported; p, T, 1/rho: the sums of squared
// deviations about the means, S2 - S1 (S1 / N) from the shifted sums).  The
// rows are read as one flat array by kFinalThreads threads (a multiple of
// kPSum): thread t sums elements t, t + kFinalThreads, ..., all of slot
// t % kPSum, so consecutive threads read consecutive words and kFinalUnroll
// loads of each thread are in flight together; the kFinalThreads / kPSum
// values of a slot are then added in thread order.
constexpr int kFinalRowsPerPass = 48, kFinalThreads = kPSum * kFinalRowsPerPass, kFinalUnroll = 4;
__global__ void __launch_bounds__(1024) k_pstats_final(int nb, const double* part, double N, double* out) {
    static_assert(kFinalThreads <= 1024, "one block");
    __shared__ double sl[kPSum][kFinalRowsPerPass];
    __shared__ double r[kPSum];
    const int t = threadIdx.x;
    if (t < kFinalThreads) {
        const long n = (long)nb * kPSum;
        double v = 0.0;
        long e = t;
        for (; e + (kFinalUnroll - 1) * (long)kFinalThreads < n; e += kFinalUnroll * (long)kFinalThreads) {
            double x[kFinalUnroll];
#pragma unroll
            for (int u = 0; u < kFinalUnroll; u++) x[u] = part[e + u * (long)kFinalThreads];
#pragma unroll
            for (int u = 0; u < kFinalUnroll; u++) v += x[u];
        }
        for (; e < n; e += kFinalThreads) v += part[e];
        sl[t % kPSum][t / kPSum] = v;
    }
    __syncthreads();
    if (t < kPSum) {
        double v = sl[t][0];
        for (int k = 1; k < kFinalRowsPerPass; k++) v += sl[t][k];
        r[t] = v;
    }
    __syncthreads();
    if (t < kPStat) {
        out[t] = r[t];
    } else if (t < kPStat + kStatB) {
        const int i = t - kPStat;  // 0 rho, 1 p, 2 T, 3 1/rho
        const double s1 = i ? r[kPStat - 1 + i] : 0.0, s2 = i ? r[kPStat + 2 + i] : 0.0;
        out[t] = fmax(0.0, s2 - s1 * (s1 / N));
    }
}

// The slopes of the three axes into their planes (whole lines in LDS where
// they fit, m <= kSplMaxM, else in chunks), then the sums; returns the number
// of partial rows.
// Rows of partial sums the fused passes write: one per kSplGroups z-line
// groups, one per 64 x lines, one per 64 y lines.
constexpr int kSplGroups = 4;
static void fused_rows(const Geo& g, unsigned rows[3]) {
    const unsigned groups = ((unsigned)g.ny * g.nx + kSplL - 1) / kSplL;
    rows[0] = (groups + kSplGroups - 1) / kSplGroups;
    rows[1] = ((unsigned)g.ny * g.nz + 63) / 64;
    rows[2] = ((unsigned)g.nx * g.nz + 63) / 64;
}

size_t print_stats_part_doubles(const Geo& g) {
    unsigned r[3];
    fused_rows(g, r);
    return (size_t)(r[0] + r[1] + r[2]) * kPSum;
}

// Returns the number of partial rows in *part_out.
template <class St>
static int launch_print_stats_st(Ctx* h, const St& st, const StatPar& p, hipStream_t s, const double** part_out) {
    const Geo& g = h->geo;
    const double r3h = 3.0 / h->grid.dx;
    const double* inv[3] = {h->spl_inv, h->spl_inv + g.P0, h->spl_inv + g.P0 + g.P1};
    const int m[3] = {g.P0, g.P1, g.P2};
    const unsigned nl[3] = {(unsigned)g.nx * g.nz, (unsigned)g.ny * g.nz, (unsigned)g.ny * g.nx};
    auto tile_bytes = [](int mm) { return sizeof(double) * 3 * (size_t)mm * kSplL; };
    const unsigned sb[2] = {(nl[0] + 63) / 64, (nl[1] + 63) / 64};
    const dim3 sth(kSplStreamThreads);
    const double* wt_x = h->spl_wt + (size_t)g.P0 * kSplRing;
    if (m[2] <= kSplMaxM) {
        // fused: the sums are taken where the slopes are made (z, then x, then y)
        unsigned rows[3];
        fused_rows(g, rows);
        double* part = h->spl_part;
        hipLaunchKernelGGL(k_spl_lines_sums<St>, dim3(rows[0]), dim3(kSplThreads), tile_bytes(m[2]), s, g, st,
                           inv[2], r3h, p, kSplGroups, h->spl_grad, part);
        hipLaunchKernelGGL((k_spl_stream<St, 1, true>), dim3(sb[1]), sth, 0, s, g, st, inv[1], wt_x, r3h,
                           h->spl_grad, part + (size_t)rows[0] * kPSum);
        hipLaunchKernelGGL((k_spl_stream<St, 0, true>), dim3(sb[0]), sth, 0, s, g, st, inv[0], h->spl_wt, r3h,
                           h->spl_grad, part + (size_t)(rows[0] + rows[1]) * kPSum);
        *part_out = part;
        return (int)(rows[0] + rows[1] + rows[2]);
    }
    // z lines too long for LDS: the nine slope planes, then the sums from them
    hipLaunchKernelGGL((k_spl_chunked<St>), dim3((nl[2] + kSpLines - 1) / kSpLines), dim3(kSpThreads), 0, s, g, st,
                       inv[2], r3h, h->spl_grad);
    hipLaunchKernelGGL((k_spl_stream<St, 1, false>), dim3(sb[1]), sth, 0, s, g, st, inv[1], wt_x, r3h,
                       h->spl_grad, nullptr);
    hipLaunchKernelGGL((k_spl_stream<St, 0, false>), dim3(sb[0]), sth, 0, s, g, st, inv[0], h->spl_wt, r3h,
                       h->spl_grad, nullptr);
    hipLaunchKernelGGL(k_pstats<St>, dim3(h->red_blocks), dim3(kRedThreads), 0, s, g, st, h->spl_grad, p,
                       make_fastdiv((unsigned)g.nz), make_fastdiv((unsigned)g.nx), h->red_part);
    *part_out = h->red_part;
    return h->red_blocks;
}

// The PrintStats sums (kPStat) and the squared deviations of p, T and 1/rho
// about their means (kStatB slots), from one sweep over the state, reduced on
// the device and copied to host_out (kPStat + kStatB doubles) by one copy,
// all on stream s.  Returns without waiting.
hipError_t launch_print_stats_async(Ctx* h, int box, hipStream_t s, double* host_out) {
    const StatPar p = stat_par(h, box);
    const size_t off = (size_t)box * h->ncell;
    int nb;
    const double* part = nullptr;
    if (h->par.precision == GKM_PREC_FP64) {
        StateD st;
        for (int q = 0; q < 5; q++) st.w[q] = h->Wd[q] + off;
        nb = launch_print_stats_st(h, st, p, s, &part);
    } else {
        StateM st;
        for (int q = 0; q < 5; q++) {
            st.hi[q] = h->Whi[q] + off;
            st.lo[q] = h->Wlo[q] + off;
        }
        nb = launch_print_stats_st(h, st, p, s, &part);
    }
    hipLaunchKernelGGL(k_pstats_final, dim3(1), dim3(kFinalThreads), 0, s, nb, part,
                       (double)h->grid.nx * h->grid.ny * h->grid.nz, h->red_out);
    return hipMemcpyAsync(host_out, h->red_out, sizeof(double) * (kPStat + kStatB), hipMemcpyDeviceToHost, s);
}

}  // namespace gkm
