// gkm_io.cpp — host readers and writers of the reference's initial-condition
// and restart files (declared in include/gkm_io.h).  Plain C++ stdio; no
// device code.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/gkm.h"
#include "../../include/gkm_io.h"

namespace gkm {
int report_error(int code, const char* msg);
}

namespace {

int io_fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int io_fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return gkm::report_error(code, buf);
}

struct File {
    FILE* f = nullptr;
    File(const std::string& path, const char* mode) : f(std::fopen(path.c_str(), mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

// "Init%i%i%i" of rank (a, b, c): the three coordinates printed back to back
// (FlowField.cpp:73)
std::string rank_file(const char* dir, int a, int b, int c) {
    char name[64];
    std::snprintf(name, sizeof name, "Init%i%i%i", a, b, c);
    std::string d(dir);
    if (!d.empty() && d.back() != '/') d += '/';
    return d + name;
}

int check_decomp(int nc, const int pd[3]) {
    if (nc < 3) return io_fail(GKM_E_ARG, "nc = %d: need at least 3 cells per edge", nc);
    for (int a = 0; a < 3; a++)
        if (pd[a] < 1 || nc % pd[a] != 0)
            return io_fail(GKM_E_ARG, "proc_dim[%d] = %d does not divide nc = %d", a, pd[a], nc);
    return GKM_OK;
}

// padded whole-box index of interior cell (j, i, k)
inline size_t padded(int nc, int j, int i, int k) {
    const size_t P = (size_t)nc + 6;
    return ((size_t)j + 3) * P * P + ((size_t)i + 3) * P + ((size_t)k + 3);
}

}  // namespace

extern "C" {

int gkm_read_weno_params(const char* path, gkm_weno_params* out) {
    if (!path || !out) return io_fail(GKM_E_ARG, "null argument");
    File in(path, "rb");
    if (!in.f) return io_fail(GKM_E_IO, "cannot open %s", path);
    gkm_weno_params p;
    double v[7];
    if (std::fread(&p.nc, sizeof(int), 1, in.f) != 1 || std::fread(v, sizeof(double), 7, in.f) != 7)
        return io_fail(GKM_E_IO, "%s: short params file (need int + 7 doubles)", path);
    p.gam = v[0];
    p.den0 = v[1];
    p.Mt = v[2];
    p.Re = v[3];
    p.mu0 = v[4];
    p.p0 = v[5];
    p.t0 = v[6];
    *out = p;
    return GKM_OK;
}

int gkm_write_weno_params(const char* path, const gkm_weno_params* p) {
    if (!path || !p) return io_fail(GKM_E_ARG, "null argument");
    File out(path, "wb");
    if (!out.f) return io_fail(GKM_E_IO, "cannot create %s", path);
    const double v[7] = {p->gam, p->den0, p->Mt, p->Re, p->mu0, p->p0, p->t0};
    if (std::fwrite(&p->nc, sizeof(int), 1, out.f) != 1 || std::fwrite(v, sizeof(double), 7, out.f) != 7)
        return io_fail(GKM_E_IO, "write to %s failed", path);
    return GKM_OK;
}

int gkm_read_weno_init(const char* dir, int nc, const int proc_dim[3], double gam, double* const W[5]) {
    if (!dir || !proc_dim || !W) return io_fail(GKM_E_ARG, "null argument");
    for (int q = 0; q < 5; q++)
        if (!W[q]) return io_fail(GKM_E_ARG, "null state plane %d", q);
    int rc = check_decomp(nc, proc_dim);
    if (rc) return rc;
    const int sy = nc / proc_dim[0], sx = nc / proc_dim[1], sz = nc / proc_dim[2];
    const size_t nseg = (size_t)sy * sx * sz;
    // u, v, w, den, p of one segment, in file order
    std::vector<double> buf(5 * nseg);
    for (int a = 0; a < proc_dim[0]; a++)
        for (int b = 0; b < proc_dim[1]; b++)
            for (int c = 0; c < proc_dim[2]; c++) {
                const std::string path = rank_file(dir, a, b, c);
                File in(path, "rb");
                if (!in.f) return io_fail(GKM_E_IO, "cannot open %s", path.c_str());
                if (std::fread(buf.data(), sizeof(double), buf.size(), in.f) != buf.size())
                    return io_fail(GKM_E_IO, "%s: short file (need 5 x %zu doubles)", path.c_str(), nseg);
                const double *u0 = buf.data(), *v0 = u0 + nseg, *w0 = v0 + nseg, *d0 = w0 + nseg,
                             *p0 = d0 + nseg;
                for (int j = 0; j < sy; j++)
                    for (int i = 0; i < sx; i++)
                        for (int k = 0; k < sz; k++) {
                            const size_t J = ((size_t)k * sx + i) * sy + j;
                            const size_t I = padded(nc, a * sy + j, b * sx + i, c * sz + k);
                            const double Vsqr = u0[J] * u0[J] + v0[J] * v0[J] + w0[J] * w0[J];
                            W[0][I] = d0[J];
                            W[1][I] = u0[J] * d0[J];
                            W[2][I] = v0[J] * d0[J];
                            W[3][I] = w0[J] * d0[J];
                            W[4][I] = d0[J] * (p0[J] / (d0[J] * (gam - 1.0)) + 0.5 * Vsqr);
                        }
            }
    return GKM_OK;
}

int gkm_write_weno_init(const char* dir, int nc, const int proc_dim[3], double gam, const double* const W[5]) {
    if (!dir || !proc_dim || !W) return io_fail(GKM_E_ARG, "null argument");
    for (int q = 0; q < 5; q++)
        if (!W[q]) return io_fail(GKM_E_ARG, "null state plane %d", q);
    int rc = check_decomp(nc, proc_dim);
    if (rc) return rc;
    const int sy = nc / proc_dim[0], sx = nc / proc_dim[1], sz = nc / proc_dim[2];
    const size_t nseg = (size_t)sy * sx * sz;
    std::vector<double> buf(5 * nseg);
    for (int a = 0; a < proc_dim[0]; a++)
        for (int b = 0; b < proc_dim[1]; b++)
            for (int c = 0; c < proc_dim[2]; c++) {
                double *u0 = buf.data(), *v0 = u0 + nseg, *w0 = v0 + nseg, *d0 = w0 + nseg, *p0 = d0 + nseg;
                for (int j = 0; j < sy; j++)
                    for (int i = 0; i < sx; i++)
                        for (int k = 0; k < sz; k++) {
                            const size_t J = ((size_t)k * sx + i) * sy + j;
                            const size_t I = padded(nc, a * sy + j, b * sx + i, c * sz + k);
                            // c2p (update.cpp:882-891)
                            const double deni = W[0][I];
                            const double ui = W[1][I] / deni, vi = W[2][I] / deni, wi = W[3][I] / deni;
                            const double Vsqr = ui * ui + vi * vi + wi * wi;
                            u0[J] = ui;
                            v0[J] = vi;
                            w0[J] = wi;
                            d0[J] = deni;
                            p0[J] = (deni * (gam - 1)) * (W[4][I] / deni - 0.5 * Vsqr);
                        }
                const std::string path = rank_file(dir, a, b, c);
                File out(path, "wb");
                if (!out.f) return io_fail(GKM_E_IO, "cannot create %s", path.c_str());
                if (std::fwrite(buf.data(), sizeof(double), buf.size(), out.f) != buf.size())
                    return io_fail(GKM_E_IO, "write to %s failed", path.c_str());
            }
    return GKM_OK;
}

int gkm_read_init_dat(const char* path, gkm_init_dat* hdr, float* u, float* v, float* w) {
    if (!path || !hdr) return io_fail(GKM_E_ARG, "null argument");
    File in(path, "rb");
    if (!in.f) return io_fail(GKM_E_IO, "cannot open %s", path);
    gkm_init_dat h;
    float x[9];
    if (std::fread(&h.n, sizeof(int), 1, in.f) != 1 || std::fread(x, sizeof(float), 9, in.f) != 9)
        return io_fail(GKM_E_IO, "%s: short header (need int + 9 floats)", path);
    if (h.n < 1 || h.n > 4096) return io_fail(GKM_E_IO, "%s: implausible N = %d", path, h.n);
    h.rho0 = x[0];
    h.Mt = x[1];
    h.Re = x[2];
    h.k0 = x[3];
    h.urms = x[4];
    h.mu = x[5];
    h.crms = x[6];
    h.p0 = x[7];
    h.eddy_time = x[8];
    *hdr = h;
    const size_t n3 = (size_t)h.n * h.n * h.n;
    float* dst[3] = {u, v, w};
    for (int q = 0; q < 3; q++) {
        if (!dst[q]) continue;
        if (std::fseek(in.f, (long)(sizeof(int) + 9 * sizeof(float) + q * n3 * sizeof(float)), SEEK_SET) != 0 ||
            std::fread(dst[q], sizeof(float), n3, in.f) != n3)
            return io_fail(GKM_E_IO, "%s: short velocity field %d (need %zu floats)", path, q, n3);
    }
    return GKM_OK;
}

int gkm_write_init_dat(const char* path, const gkm_init_dat* h, const float* u, const float* v, const float* w) {
    if (!path || !h || !u || !v || !w) return io_fail(GKM_E_ARG, "null argument");
    if (h->n < 1) return io_fail(GKM_E_ARG, "N = %d", h->n);
    File out(path, "wb");
    if (!out.f) return io_fail(GKM_E_IO, "cannot create %s", path);
    const float x[9] = {h->rho0, h->Mt, h->Re, h->k0, h->urms, h->mu, h->crms, h->p0, h->eddy_time};
    const size_t n3 = (size_t)h->n * h->n * h->n;
    if (std::fwrite(&h->n, sizeof(int), 1, out.f) != 1 || std::fwrite(x, sizeof(float), 9, out.f) != 9 ||
        std::fwrite(u, sizeof(float), n3, out.f) != n3 || std::fwrite(v, sizeof(float), n3, out.f) != n3 ||
        std::fwrite(w, sizeof(float), n3, out.f) != n3)
        return io_fail(GKM_E_IO, "write to %s failed", path);
    return GKM_OK;
}

int gkm_init_dat_state(const gkm_init_dat* h, const float* u, const float* v, const float* w, double gam,
                       double* const W[5]) {
    if (!h || !u || !v || !w || !W) return io_fail(GKM_E_ARG, "null argument");
    for (int q = 0; q < 5; q++)
        if (!W[q]) return io_fail(GKM_E_ARG, "null state plane %d", q);
    const int n = h->n;
    const double den0 = h->rho0, p0 = h->p0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++)
            for (int k = 0; k < n; k++) {
                const size_t J = ((size_t)k * n + i) * n + j;
                const size_t I = padded(n, j, i, k);
                const double uu = u[J], vv = v[J], ww = w[J];
                const double Vsqr = uu * uu + vv * vv + ww * ww;
                W[0][I] = den0;
                W[1][I] = uu * den0;
                W[2][I] = vv * den0;
                W[3][I] = ww * den0;
                W[4][I] = den0 * (p0 / (den0 * (gam - 1.0)) + .5 * Vsqr);
            }
    return GKM_OK;
}

}  // extern "C"
