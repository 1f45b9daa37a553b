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
template <int NQ, unsigned MAXMASK>
__device__ __forceinline__ void block_reduce_store(double (&acc)[NQ], double* part) {
    __shared__ double lds[NQ][kRedThreads / 64];
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
        for (int w = 1; w < kRedThreads / 64; w++) v = ((MAXMASK >> q) & 1u) ? fmax(v, lds[q][w]) : v + lds[q][w];
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

// final pass: one block reduces nb partial rows of NQ slots in block order
template <int NQ, unsigned MAXMASK>
__global__ void __launch_bounds__(kRedThreads) k_reduce_final(int nb, const double* part, double* out) {
    for (int q = 0; q < NQ; q++) {
        const bool ismax = (MAXMASK >> q) & 1u;
        double v = ismax ? 0.0 : 0.0;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            const double x = part[(long)b * NQ + q];
            v = ismax ? fmax(v, x) : v + x;
        }
        v = ismax ? wave_max(v) : wave_sum(v);
        __shared__ double lds[kRedThreads / 64];
        if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = lds[0];
            for (int w = 1; w < kRedThreads / 64; w++) s = ismax ? fmax(s, lds[w]) : s + lds[w];
            out[q] = s;
        }
        __syncthreads();
    }
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

hipError_t launch_stats_b(Ctx* h, int box, const double m[4], double* host_out) {
    const Geo& g = h->geo;
    const StatPar p = stat_par(h, box);
    const int nb = h->red_blocks;
    const size_t off = (size_t)box * h->ncell;
    if (h->par.precision == GKM_PREC_FP64) {
        StateD st;
        for (int q = 0; q < 5; q++) st.w[q] = h->Wd[q] + off;
        hipLaunchKernelGGL(k_stats_b<StateD>, dim3(nb), dim3(kRedThreads), 0, h->stream, g, st, p, m[0], m[1], m[2],
                           m[3], h->red_part);
    } else {
        StateM st;
        for (int q = 0; q < 5; q++) {
            st.hi[q] = h->Whi[q] + off;
            st.lo[q] = h->Wlo[q] + off;
        }
        hipLaunchKernelGGL(k_stats_b<StateM>, dim3(nb), dim3(kRedThreads), 0, h->stream, g, st, p, m[0], m[1], m[2],
                           m[3], h->red_part);
    }
    hipLaunchKernelGGL((k_reduce_final<kStatB, 0u>), dim3(1), dim3(kRedThreads), 0, h->stream, nb, h->red_part,
                       h->red_out);
    hipError_t e = hipMemcpyAsync(host_out, h->red_out, sizeof(double) * kStatB, hipMemcpyDeviceToHost, h->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(h->stream);
}

// ---------------------------------------------------------------------------
// PrintStats with spline gradients (weno/src/update.cpp:638-784).
//
// k_spline<A>: one thread per grid line along storage axis A (0 = y, 1 = x,
// 2 = z) through interior nodes.  It walks the whole padded line once forward
// and once back, solving for the slopes of the natural cubic spline through
// u, v, w = (rho u, rho v, rho w)/rho at its m nodes:
//   2 D_0 + D_1 = (3/h)(f_1 - f_0)
//   D_{i-1} + 4 D_i + D_{i+1} = (3/h)(f_{i+1} - f_{i-1})
//   D_{m-2} + 2 D_{m-1} = (3/h)(f_{m-1} - f_{m-2})
// The matrix is the same for every line of the axis, so the Thomas pivots
// inv_i = 1/(b_i - inv_{i-1}) come precomputed: forward y_i = (r_i - y_{i-1})
// inv_i, backward D_i = y_i - inv_i D_{i+1}.  The forward values of nodes
// 3..m-1 go to the gradient plane and are overwritten by the slopes on the
// way back; only interior nodes' slopes are used.
// grad planes: (3 A + comp) * ncell.
// ---------------------------------------------------------------------------
template <class St, int A>
__global__ void __launch_bounds__(256) k_spline(Geo g, St st, const double* __restrict__ inv, double r3h,
                                                double* __restrict__ grad) {
    const int m = A == 0 ? g.P0 : (A == 1 ? g.P1 : g.P2);
    // lines: the two other axes over their interior ranges, the faster one inner
    const int n_in = A == 2 ? g.nx : g.nz;
    const int n_out = A == 0 ? g.nx : g.ny;
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (unsigned)n_in * (unsigned)n_out) return;
    const unsigned a_in = t % (unsigned)n_in + 3, a_out = t / (unsigned)n_in + 3;
    unsigned base, stride;
    if (A == 0) {  // y lines over (x, z)
        base = a_out * g.s1 + a_in;
        stride = g.s0;
    } else if (A == 1) {  // x lines over (y, z)
        base = a_out * g.s0 + a_in;
        stride = g.s1;
    } else {  // z lines over (y, x)
        base = a_out * g.s0 + a_in * g.s1;
        stride = 1;
    }
    double* G0 = grad + (size_t)(3 * A) * g.ncell;
    double* G1 = G0 + g.ncell;
    double* G2 = G1 + g.ncell;
    auto velo = [&](unsigned c, double f[3]) {
        const double r = st.get(0, c);
        f[0] = st.get(1, c) / r;
        f[1] = st.get(2, c) / r;
        f[2] = st.get(3, c) / r;
    };
    double fm[3], f0[3], fp[3], y[3] = {0.0, 0.0, 0.0};
    velo(base, f0);
    velo(base + stride, fp);
    fm[0] = f0[0]; fm[1] = f0[1]; fm[2] = f0[2];
    for (int i = 0; i < m; i++) {
        const unsigned c = base + (unsigned)i * stride;
        const double w = inv[i];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            // one-sided differences at the two ends, central inside
            const double hi = i == m - 1 ? f0[q] : fp[q];
            const double lo = i == 0 ? f0[q] : fm[q];
            y[q] = (r3h * (hi - lo) - y[q]) * w;
        }
        if (i >= 3) {
            put(G0, c, y[0]);
            put(G1, c, y[1]);
            put(G2, c, y[2]);
        }
        if (i + 1 < m) {
#pragma unroll
            for (int q = 0; q < 3; q++) {
                fm[q] = f0[q];
                f0[q] = fp[q];
            }
            if (i + 2 < m) velo(c + 2 * stride, fp);
        }
    }
    // y now holds y_{m-1} = D_{m-1}; back-substitute down to node 3
    double d[3] = {y[0], y[1], y[2]};
    for (int i = m - 2; i >= 3; i--) {
        const unsigned c = base + (unsigned)i * stride;
        const double w = inv[i];
        d[0] = ld(G0, c) - w * d[0];
        d[1] = ld(G1, c) - w * d[1];
        d[2] = ld(G2, c) - w * d[2];
        put(G0, c, d[0]);
        put(G1, c, d[1]);
        put(G2, c, d[2]);
    }
}

// z lines run along the unit stride, so one thread per line would touch a
// different cache line on every lane.  k_spline_z gives a 64-lane block 64
// lines and walks them in chunks of 16 nodes staged through LDS: the block
// loads and stores 16-node row segments (128 B of doubles per line), and
// each lane runs the same Thomas recursion as k_spline on its own line out
// of LDS.
constexpr int kSzLines = 64, kSzChunk = 16;

template <class St>
__global__ void __launch_bounds__(kSzLines) k_spline_z(Geo g, St st, const double* __restrict__ inv, double r3h,
                                                       double* __restrict__ grad) {
    __shared__ double tile[3][kSzLines][kSzChunk + 1];
    const int lane = threadIdx.x;
    const int m = g.P2;
    const int nline = g.ny * g.nx;
    const int line0 = blockIdx.x * kSzLines;
    double* G[3] = {grad + (size_t)6 * g.ncell, grad + (size_t)7 * g.ncell, grad + (size_t)8 * g.ncell};
    auto line_base = [&](int id) -> unsigned {
        const unsigned j = (unsigned)(id / g.nx) + 3, i = (unsigned)(id % g.nx) + 3;
        return j * g.s0 + i * g.s1;
    };
    const bool mine = line0 + lane < nline;
    const unsigned base = mine ? line_base(line0 + lane) : 0u;
    // cooperative pass over the block's lines x chunk nodes [n0, n0 + kSzChunk):
    // element e = r * 64 + lane covers line e / kSzChunk, node e % kSzChunk
    auto coop = [&](int n0, auto&& body) {
#pragma unroll 4
        for (int r = 0; r < kSzLines * kSzChunk / kSzLines; r++) {
            const int e = r * kSzLines + lane;
            const int l = e / kSzChunk, t = e % kSzChunk;
            const int n = n0 + t;
            if (line0 + l < nline && n < m) body(l, t, line_base(line0 + l) + (unsigned)n);
        }
    };
    auto velo = [&](unsigned c, double f[3]) {
        const double r = st.get(0, c);
        f[0] = st.get(1, c) / r;
        f[1] = st.get(2, c) / r;
        f[2] = st.get(3, c) / r;
    };
    double f0[3] = {0.0, 0.0, 0.0}, fm[3], y[3] = {0.0, 0.0, 0.0};
    if (mine) velo(base, f0);
    fm[0] = f0[0]; fm[1] = f0[1]; fm[2] = f0[2];
    const int nch = (m + kSzChunk - 1) / kSzChunk;
    // forward: the tile holds f at nodes n0+1 .. n0+16; lane l walks nodes
    // n0 .. n0+15 of its line and leaves y_n where it read f_{n+1}
    for (int ch = 0; ch < nch; ch++) {
        const int n0 = ch * kSzChunk;
        coop(n0 + 1, [&](int l, int t, unsigned c) {
            double f[3];
            velo(c, f);
            tile[0][l][t] = f[0];
            tile[1][l][t] = f[1];
            tile[2][l][t] = f[2];
        });
        __syncthreads();
        for (int t = 0; t < kSzChunk; t++) {
            const int n = n0 + t;
            if (n >= m) break;
            const double w = inv[n];
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const double fp = tile[q][lane][t];  // node n+1 (not used at n = m-1)
                const double hi = n == m - 1 ? f0[q] : fp;
                const double lo = n == 0 ? f0[q] : fm[q];
                y[q] = (r3h * (hi - lo) - y[q]) * w;
                tile[q][lane][t] = y[q];
                fm[q] = f0[q];
                f0[q] = fp;
            }
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
    double d[3] = {y[0], y[1], y[2]};
    for (int ch = nch - 1; ch >= 0; ch--) {
        const int n0 = ch * kSzChunk;
        coop(n0, [&](int l, int t, unsigned c) {
            tile[0][l][t] = ld(G[0], c);
            tile[1][l][t] = ld(G[1], c);
            tile[2][l][t] = ld(G[2], c);
        });
        __syncthreads();
        for (int t = kSzChunk - 1; t >= 0; t--) {
            const int n = n0 + t;
            if (n >= m - 1) continue;  // node m-1 keeps y_{m-1}
            const double w = inv[n];
#pragma unroll
            for (int q = 0; q < 3; q++) {
                d[q] = tile[q][lane][t] - w * d[q];
                tile[q][lane][t] = d[q];
            }
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

// PrintStats sums over the interior (update.cpp:658-707):
//  0 rho|u|^2  1 mu du_i/dx_j du_i/dx_j  2 S_ij S_ij  3..5 (du/dx)^2,3,4
//  6..8 (dv/dy)^n  9..11 (dw/dz)^n  12 p  13 T  14 1/rho
template <class St>
__global__ void __launch_bounds__(kRedThreads) k_pstats(Geo g, St st, const double* __restrict__ grad, StatPar p,
                                                        double* part) {
    double acc[kPStat];
#pragma unroll
    for (int q = 0; q < kPStat; q++) acc[q] = 0.0;
    const long n = (long)g.nx * g.ny * g.nz;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int k = (int)(t % g.nz);
        const long r = t / g.nz;
        const int i = (int)(r % g.nx);
        const int j = (int)(r / g.nx);
        const unsigned c = (unsigned)(j + 3) * g.s0 + (unsigned)(i + 3) * g.s1 + (unsigned)(k + 3);
        const double rho = st.get(0, c);
        double u[3];
        vel(st, c, u);
        const double q2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        const double P = (rho * (p.gam - 1)) * (st.get(4, c) / rho - 0.5 * q2);
        const double T = P / (p.R * rho);
        auto gr = [&](int a, int comp) { return ld(grad + (size_t)(3 * a + comp) * g.ncell, c); };
        const double dudx = gr(1, 0), dudy = gr(0, 0), dudz = gr(2, 0);
        const double dvdx = gr(1, 1), dvdy = gr(0, 1), dvdz = gr(2, 1);
        const double dwdx = gr(1, 2), dwdy = gr(0, 2), dwdz = gr(2, 2);
        const double mu = visc(p, T);
        acc[0] += rho * q2;
        acc[1] += mu * (dudx * dudx + dudy * dudy + dudz * dudz + dvdx * dvdx + dvdy * dvdy + dvdz * dvdz +
                        dwdx * dwdx + dwdy * dwdy + dwdz * dwdz);
        acc[2] += dudx * dudx + 0.5 * (dudy + dvdx) * (dudy + dvdx) + 0.5 * (dudz + dwdx) * (dudz + dwdx) +
                  0.5 * (dwdy + dvdz) * (dwdy + dvdz) + dvdy * dvdy + dwdz * dwdz;
        double x2 = dudx * dudx; acc[3] += x2; acc[4] += x2 * dudx; acc[5] += x2 * x2;
        x2 = dvdy * dvdy; acc[6] += x2; acc[7] += x2 * dvdy; acc[8] += x2 * x2;
        x2 = dwdz * dwdz; acc[9] += x2; acc[10] += x2 * dwdz; acc[11] += x2 * x2;
        acc[12] += P; acc[13] += T; acc[14] += 1.0 / rho;
    }
    block_reduce_store<kPStat, 0u>(acc, part);
}

template <class St>
static void launch_print_stats_st(Ctx* h, const St& st, const StatPar& p) {
    const Geo& g = h->geo;
    const double r3h = 3.0 / h->grid.dx;
    const double* inv0 = h->spl_inv;
    const double* inv1 = inv0 + g.P0;
    const double* inv2 = inv1 + g.P1;
    const unsigned ly = (unsigned)g.nx * g.nz, lx = (unsigned)g.ny * g.nz, lz = (unsigned)g.ny * g.nx;
    hipLaunchKernelGGL((k_spline<St, 0>), dim3((ly + 255) / 256), dim3(256), 0, h->stream, g, st, inv0, r3h,
                       h->spl_grad);
    hipLaunchKernelGGL((k_spline<St, 1>), dim3((lx + 255) / 256), dim3(256), 0, h->stream, g, st, inv1, r3h,
                       h->spl_grad);
    hipLaunchKernelGGL(k_spline_z<St>, dim3((lz + kSzLines - 1) / kSzLines), dim3(kSzLines), 0, h->stream, g, st,
                       inv2, r3h, h->spl_grad);
    hipLaunchKernelGGL(k_pstats<St>, dim3(h->red_blocks), dim3(kRedThreads), 0, h->stream, g, st, h->spl_grad, p,
                       h->red_part);
}

hipError_t launch_print_stats(Ctx* h, int box, double* host_out) {
    const StatPar p = stat_par(h, box);
    const size_t off = (size_t)box * h->ncell;
    if (h->par.precision == GKM_PREC_FP64) {
        StateD st;
        for (int q = 0; q < 5; q++) st.w[q] = h->Wd[q] + off;
        launch_print_stats_st(h, st, p);
    } else {
        StateM st;
        for (int q = 0; q < 5; q++) {
            st.hi[q] = h->Whi[q] + off;
            st.lo[q] = h->Wlo[q] + off;
        }
        launch_print_stats_st(h, st, p);
    }
    hipLaunchKernelGGL((k_reduce_final<kPStat, 0u>), dim3(1), dim3(kRedThreads), 0, h->stream, h->red_blocks,
                       h->red_part, h->red_out);
    hipError_t e = hipMemcpyAsync(host_out, h->red_out, sizeof(double) * kPStat, hipMemcpyDeviceToHost, h->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(h->stream);
}

}  // namespace gkm
