// gkm_ic.cpp — decaying isotropic turbulence initial condition on the host
// (the init_BS.m recipe, GKM/CUDA GKM/InitialConditions/init_BS.m:24-157),
// for the C ABI of include/gkm_ic.h.  Host code only: the three-dimensional
// transforms are separable one-dimensional DFTs (iterative radix-2 for powers
// of two, Bluestein's chirp-z convolution through a radix-2 transform for other
// lengths), spread over host threads by lines.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/gkm.h"
#include "../../include/gkm_ic.h"

namespace gkm {
int report_error(int code, const char* msg);
}

namespace {

using cplx = std::complex<double>;
constexpr double kPi = 3.14159265358979323846264338327950288;

// iterative radix-2 transform of length m = 2^p, exponent sign -1 (forward)
struct Radix2 {
    int m = 0;
    std::vector<int> rev;
    std::vector<cplx> w;  // exp(-2 pi i t / m), t < m/2
    explicit Radix2(int m_) : m(m_), rev(m_), w(m_ / 2) {
        int p = 0;
        while ((1 << p) < m) p++;
        for (int t = 0; t < m; t++) {
            int r = 0;
            for (int b = 0; b < p; b++)
                if (t & (1 << b)) r |= 1 << (p - 1 - b);
            rev[t] = r;
        }
        for (int t = 0; t < m / 2; t++) w[t] = std::polar(1.0, -2.0 * kPi * t / m);
    }
    void forward(cplx* a) const {
        for (int t = 0; t < m; t++)
            if (t < rev[t]) std::swap(a[t], a[rev[t]]);
        for (int len = 2; len <= m; len <<= 1) {
            const int half = len >> 1, step = m / len;
            for (int s = 0; s < m; s += len)
                for (int t = 0; t < half; t++) {
                    const cplx x = a[s + t], y = a[s + t + half] * w[t * step];
                    a[s + t] = x + y;
                    a[s + t + half] = x - y;
                }
        }
    }
};

// forward DFT of any length n: X_f = sum_t x_t exp(-2 pi i f t / n)
struct Dft {
    int n;
    bool pow2;
    Radix2 r2;
    std::vector<cplx> chirp;  // exp(-i pi t^2 / n), t < n
    std::vector<cplx> bhat;   // transform of the conjugate chirp, wrapped to length m
    static int next_pow2(int x) {
        int m = 1;
        while (m < x) m <<= 1;
        return m;
    }
    explicit Dft(int n_)
        : n(n_), pow2((n_ & (n_ - 1)) == 0), r2(((n_ & (n_ - 1)) == 0) ? n_ : next_pow2(2 * n_ - 1)) {
        if (pow2) return;
        const int m = r2.m;
        chirp.resize(n);
        for (long t = 0; t < n; t++) chirp[t] = std::polar(1.0, -kPi * (double)((t * t) % (2L * n)) / n);
        bhat.assign(m, cplx(0, 0));
        bhat[0] = std::conj(chirp[0]);
        for (int t = 1; t < n; t++) bhat[t] = bhat[m - t] = std::conj(chirp[t]);
        r2.forward(bhat.data());
    }
    // in place on x[0..n), work holds r2.m values
    void forward(cplx* x, std::vector<cplx>& work) const {
        if (pow2) {
            r2.forward(x);
            return;
        }
        const int m = r2.m;
        work.assign(m, cplx(0, 0));
        for (int t = 0; t < n; t++) work[t] = x[t] * chirp[t];
        r2.forward(work.data());
        for (int t = 0; t < m; t++) work[t] = std::conj(work[t] * bhat[t]);  // conj: inverse by the forward transform
        r2.forward(work.data());
        const double s = 1.0 / m;
        for (int t = 0; t < n; t++) x[t] = std::conj(work[t]) * s * chirp[t];
    }
};

// runs f(line) for line in [0, count) on up to 16 host threads
template <class F> void parallel_lines(long count, F&& f) {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const long nt = std::max(1L, std::min((long)hw, count / 64));
    std::vector<std::thread> pool;
    for (long t = 0; t < nt; t++)
        pool.emplace_back([&, t] {
            for (long l = t; l < count; l += nt) f(l);
        });
    for (auto& th : pool) th.join();
}

// 3-D transform of a[j][i][k] (extents n0, n1, n2); inverse = conj-forward-conj, unscaled
void dft3(std::vector<cplx>& a, int n0, int n1, int n2, bool inverse) {
    const long s0 = (long)n1 * n2, s1 = n2;
    if (inverse)
        for (auto& v : a) v = std::conj(v);
    const int ext[3] = {n0, n1, n2};
    const long str[3] = {s0, s1, 1};
    for (int ax = 0; ax < 3; ax++) {
        const int n = ext[ax];
        const Dft dft(n);
        const int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;  // the other two axes
        const long lines = (long)ext[o1] * ext[o2];
        parallel_lines(lines, [&](long l) {
            thread_local std::vector<cplx> line, work;
            line.resize(n);
            const long a1 = l / ext[o2], a2 = l % ext[o2];
            const long base = a1 * str[o1] + a2 * str[o2];
            for (int t = 0; t < n; t++) line[t] = a[base + t * str[ax]];
            dft.forward(line.data(), work);
            for (int t = 0; t < n; t++) a[base + t * str[ax]] = line[t];
        });
    }
    if (inverse)
        for (auto& v : a) v = std::conj(v);
}

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
};

// signed wavenumber of DFT index t of length n (numpy's fftfreq(n, 1/n))
inline double wavenumber(int t, int n) { return (double)(t < (n + 1) / 2 ? t : t - n); }

int ic_fail(const char* msg) { return gkm::report_error(GKM_E_ARG, msg); }

}  // namespace

extern "C" {

int gkm_dhit_velocity(int nx, int ny, int nz, const gkm_ic* ic, double gamma, double* u, double* v, double* w,
                      gkm_ic_info* info) {
    if (!ic || !u || !v || !w) return ic_fail("null argument");
    if (nx < 1 || ny < 1 || nz < 1) return ic_fail("box extents must be positive");
    if (!(ic->k0 > 0) || !(ic->Mt > 0) || !(ic->Re > 0) || !(ic->rho0 > 0) || !(gamma > 1))
        return ic_fail("k0, Mt, Re, rho0 must be positive and gamma > 1");
    const long N = (long)nx * ny * nz;
    SplitMix64 rng{ic->seed};
    std::vector<cplx> f[3];
    for (int c = 0; c < 3; c++) {
        f[c].resize(N);
        for (long t = 0; t < N; t++) f[c][t] = cplx(rng.uniform(), 0.0);
        dft3(f[c], ny, nx, nz, false);
    }
    // solenoidal projection, shell energies (init_BS.m:29-83)
    const int n0 = ny, n1 = nx, n2 = nz;
    auto at = [&](int j, int i, int k) { return ((long)j * n1 + i) * n2 + k; };
    int nshell = 0;
    for (int j = 0; j < n0; j++)
        for (int i = 0; i < n1; i++)
            for (int k = 0; k < n2; k++) {
                const double ky = wavenumber(j, n0), kx = wavenumber(i, n1), kz = wavenumber(k, n2);
                nshell = std::max(nshell, (int)std::lrint(std::sqrt(kx * kx + ky * ky + kz * kz)) + 1);
            }
    std::vector<double> espec(nshell, 0.0);
    std::vector<int> shell(N);
    std::vector<char> nyq(N);
    for (int j = 0; j < n0; j++)
        for (int i = 0; i < n1; i++)
            for (int k = 0; k < n2; k++) {
                const long c = at(j, i, k);
                const double ky = wavenumber(j, n0), kx = wavenumber(i, n1), kz = wavenumber(k, n2);
                const double ksq = std::max(kx * kx + ky * ky + kz * kz, 1e-14);
                cplx& U = f[0][c];
                cplx& V = f[1][c];
                cplx& W = f[2][c];
                const cplx kdotv = kx * U + ky * V + kz * W;
                U -= kdotv * kx / ksq;
                V -= kdotv * ky / ksq;
                W -= kdotv * kz / ksq;
                // Nyquist planes of even extents cannot stay Hermitian: dropped
                nyq[c] = (n0 % 2 == 0 && j == n0 / 2) || (n1 % 2 == 0 && i == n1 / 2) || (n2 % 2 == 0 && k == n2 / 2);
                // |k|^2 is an integer, so |k| is never a half-integer and rint is unambiguous
                shell[c] = (int)std::lrint(std::sqrt(kx * kx + ky * ky + kz * kz));
                if (!nyq[c]) espec[shell[c]] += std::norm(U) + std::norm(V) + std::norm(W);
            }
    std::vector<double> fac(nshell);
    const double k0 = ic->k0, Mt = ic->Mt;
    for (int m = 0; m < nshell; m++) {
        const double e_ex = std::sqrt(2.0 / kPi) * (16.0 / 3.0) * Mt * Mt * std::pow((double)m, 4) *
                            std::exp(-2.0 * m * m / (k0 * k0)) * (double)N * (double)N / std::pow(k0, 5);
        fac[m] = std::sqrt(e_ex / std::max(espec[m], 1e-14));
    }
    double* out[3] = {u, v, w};
    double q2 = 0.0;
    for (int c = 0; c < 3; c++) {
        for (long t = 0; t < N; t++) f[c][t] *= nyq[t] ? 0.0 : fac[shell[t]];
        dft3(f[c], ny, nx, nz, true);
        double part = 0.0;
        for (long t = 0; t < N; t++) {
            const double x = f[c][t].real() / (double)N;
            out[c][t] = x;
            part += x * x;
        }
        q2 += part;
    }
    if (info) {
        info->q2 = q2 / (double)N;
        info->urms = std::sqrt(info->q2 / 3.0);
        info->c0 = std::sqrt(3.0) * info->urms / Mt;
        info->p0 = ic->rho0 * info->c0 * info->c0 / gamma;
        info->mu0 = info->urms * (2.0 / k0) * ic->rho0 / ic->Re;
        info->t0 = (2.0 / k0) / info->urms;
    }
    return GKM_OK;
}

int gkm_init_dhit_box(gkm_t* h, int box, const gkm_ic* ic, gkm_ic_info* info) {
    gkm_grid g;
    gkm_params p;
    int rc = gkm_get_grid(h, &g);
    if (!rc) rc = gkm_get_params_box(h, box, &p);
    if (rc) return rc;
    if (!ic) return ic_fail("null initial condition");
    const long N = (long)g.nx * g.ny * g.nz;
    std::vector<double> u(N), v(N), w(N);
    gkm_ic_info in;
    rc = gkm_dhit_velocity(g.nx, g.ny, g.nz, ic, p.gamma, u.data(), v.data(), w.data(), &in);
    if (rc) return rc;
    // conservative padded state (FlowField.cpp:97-102), ghosts filled below
    const int P0 = g.ny + 6, P1 = g.nx + 6, P2 = g.nz + 6;
    const long np = (long)P0 * P1 * P2;
    std::vector<double> Wv[5];
    for (auto& x : Wv) x.assign(np, 0.0);
    const double d0 = ic->rho0, e0 = in.p0 / (d0 * (p.gamma - 1.0));
    for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++)
            for (int k = 0; k < g.nz; k++) {
                const long s = ((long)j * g.nx + i) * g.nz + k;
                const long c = ((long)(j + 3) * P1 + (i + 3)) * P2 + (k + 3);
                const double vsqr = u[s] * u[s] + v[s] * v[s] + w[s] * w[s];
                Wv[0][c] = d0;
                Wv[1][c] = u[s] * d0;
                Wv[2][c] = v[s] * d0;
                Wv[3][c] = w[s] * d0;
                Wv[4][c] = d0 * (e0 + 0.5 * vsqr);
            }
    p.mu0 = in.mu0;
    p.T0 = in.p0 / (p.R * d0);  // FlowField.cpp:46
    rc = gkm_set_params_box(h, box, &p);
    if (rc) return rc;
    const double* W[5] = {Wv[0].data(), Wv[1].data(), Wv[2].data(), Wv[3].data(), Wv[4].data()};
    rc = gkm_set_state_box(h, box, W);
    if (!rc) rc = gkm_fill_ghosts(h);
    if (!rc && info) *info = in;
    return rc;
}

int gkm_init_dhit(gkm_t* h, const gkm_ic* ic, gkm_ic_info* info) { return gkm_init_dhit_box(h, 0, ic, info); }

}  // extern "C"
