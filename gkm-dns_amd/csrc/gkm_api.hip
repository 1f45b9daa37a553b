// gkm_api.hip — the C ABI of libgkm.so (declared in include/gkm.h).
// Host code only: argument checks, device allocation, copies and launches.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "gkm_internal.h"

using gkm::Ctx;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return fail(GKM_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

int check_params(const gkm_grid* g, const gkm_params* p) {
    if (!g || !p) return fail(GKM_E_ARG, "null grid or params");
    if (g->ghost != 3) return fail(GKM_E_ARG, "ghost must be 3 (got %d)", g->ghost);
    if (g->nx < 3 || g->ny < 3 || g->nz < 3)
        return fail(GKM_E_ARG, "box %dx%dx%d too small: each axis needs at least 3 interior cells", g->nx, g->ny,
                    g->nz);
    if (!(g->dx > 0) || !std::isfinite(g->dx)) return fail(GKM_E_ARG, "dx must be positive");
    if (!(p->gamma > 1.0)) return fail(GKM_E_ARG, "gamma must exceed 1");
    if (!(p->R > 0) || !(p->Pr > 0) || !(p->mu0 >= 0) || !(p->T0 > 0))
        return fail(GKM_E_ARG, "R, Pr, T0 must be positive and mu0 non-negative");
    if (p->K < 0) return fail(GKM_E_ARG, "K must be non-negative");
    if (p->visc_law != GKM_VISC_SUTHERLAND && p->visc_law != GKM_VISC_POWER) return fail(GKM_E_ARG, "bad visc_law");
    if (p->recon != GKM_RECON_SYMOO && p->recon != GKM_RECON_JS) return fail(GKM_E_ARG, "bad recon");
    if (p->energy_recon != GKM_ENERGY_T && p->energy_recon != GKM_ENERGY_E) return fail(GKM_E_ARG, "bad energy_recon");
    if (p->precision != GKM_PREC_FP64 && p->precision != GKM_PREC_MIXED) return fail(GKM_E_ARG, "bad precision");
    if (p->flux_scheme != GKM_SCHEME_WENO_BGK && p->flux_scheme != GKM_SCHEME_GKS)
        return fail(GKM_E_ARG, "bad flux_scheme");
    return GKM_OK;
}

Ctx* ctx(gkm_t* h) { return (Ctx*)h; }

int check_box(Ctx* h, int box) {
    if (!h) return fail(GKM_E_ARG, "null handle");
    if (box < 0 || box >= h->nbox) return fail(GKM_E_ARG, "box %d out of range [0, %d)", box, h->nbox);
    return GKM_OK;
}

void drop_graph(Ctx* h) {
    for (int f = 0; f < 2; f++) {
        if (h->graph_exec[f]) (void)hipGraphExecDestroy(h->graph_exec[f]);
        if (h->graph[f]) (void)hipGraphDestroy(h->graph[f]);
        h->graph_exec[f] = nullptr;
        h->graph[f] = nullptr;
        h->graph_ready[f] = false;
    }
}

// Upload per-box step parameters for the given dt.  They go through pinned
// staging; the kernels read them from the device array, so a captured graph
// stays valid when dt changes.
int upload_pars(Ctx* h, const double* dt) {
    bool same = h->pars_valid;
    for (int b = 0; b < h->nbox; b++) {
        // step_par zero-fills the struct, padding included, and the bytes are
        // copied as they are, so the comparison sees only the values
        gkm::StepPar p = gkm::step_par(h->grid, h->bpar[b], dt ? dt[b] : 0.0);
        if (!same || std::memcmp(&p, &h->h_pars[b], sizeof p) != 0) same = false;
        std::memcpy(&h->h_pars[b], &p, sizeof p);
    }
    if (same) return GKM_OK;
    HIP_TRY(hipMemcpyAsync(h->d_pars, h->h_pars, sizeof(gkm::StepPar) * h->nbox, hipMemcpyHostToDevice, h->stream));
    // the staging buffer is rewritten by the next call: wait for the copy
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->pars_valid = true;
    return GKM_OK;
}

void free_ctx(Ctx* h) {
    if (h->halo_send) (void)hipFree(h->halo_send);
    if (h->halo_recv) (void)hipFree(h->halo_recv);
    if (h->spl_grad) (void)hipFree(h->spl_grad);
    if (h->spl_inv) (void)hipFree(h->spl_inv);
    if (h->spl_wt) (void)hipFree(h->spl_wt);
    if (h->spl_part) (void)hipFree(h->spl_part);
    if (h->h_pars) (void)hipHostFree(h->h_pars);
    if (h->ps_host) (void)hipHostFree(h->ps_host);
    if (h->ps_row) (void)hipHostFree(h->ps_row);
    if (h->h_bad) (void)hipHostFree(h->h_bad);
    if (h->ev_bad) (void)hipEventDestroy(h->ev_bad);
    if (h->arena) (void)hipFree(h->arena);
    if (h->ev_state) (void)hipEventDestroy(h->ev_state);
    for (hipEvent_t e : h->ev_ps)
        if (e) (void)hipEventDestroy(e);
    if (h->stream2) (void)hipStreamDestroy(h->stream2);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

// GKM_E_NONPHYSICAL for the first box whose flag h_bad holds
int bad_report(const Ctx* h) {
    for (int b = 0; b < h->nbox; b++)
        if (h->h_bad[b])
            return fail(GKM_E_NONPHYSICAL,
                        "box %d: a step produced a non-physical state (rho <= 0, p <= 0 or a non-finite value); "
                        "replace it with gkm_set_state",
                        b);
    return GKM_OK;
}

// wait for the handle's stream, then read the non-physical flags
int check_physical(Ctx* h) {
    HIP_TRY(hipMemcpyAsync(h->h_bad, h->d_bad, sizeof(unsigned) * h->nbox, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return bad_report(h);
}

// the flags of the copy that ended the last stepping call, if it has landed
// (no wait); a copy still in flight is read by a later call
int poll_physical(Ctx* h) {
    if (!h->bad_copied) return GKM_OK;
    const hipError_t q = hipEventQuery(h->ev_bad);
    if (q == hipErrorNotReady) {
        (void)hipGetLastError();  // not an error: clear it from the thread's last-error state
        return GKM_OK;
    }
    if (q != hipSuccess) return fail(GKM_E_HIP, "hipEventQuery failed: %s", hipGetErrorString(q));
    return bad_report(h);
}

// queue the copy of the flags that poll_physical reads
int queue_flag_copy(Ctx* h) {
    HIP_TRY(hipMemcpyAsync(h->h_bad, h->d_bad, sizeof(unsigned) * h->nbox, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipEventRecord(h->ev_bad, h->stream));
    h->bad_copied = true;
    return GKM_OK;
}

}  // namespace

// error reporting for the other host translation units (gkm_io.cpp)
int gkm::report_error(int code, const char* msg) { return fail(code, "%s", msg); }

extern "C" {

const char* gkm_last_error(void) { return g_err.c_str(); }

const char* gkm_version(void) { return "gkm-mi355x 0.2 gfx950"; }

int gkm_create_batch(const gkm_grid* grid, const gkm_params* params, int nbox, gkm_t** out) {
    if (!out) return fail(GKM_E_ARG, "null output handle");
    *out = nullptr;
    if (!params || nbox < 1) return fail(GKM_E_ARG, "need nbox >= 1 parameter sets");
    for (int b = 0; b < nbox; b++) {
        int rc = check_params(grid, &params[b]);
        if (rc) return rc;
        if (params[b].recon != params[0].recon || params[b].energy_recon != params[0].energy_recon ||
            params[b].zbug_compat != params[0].zbug_compat || params[b].precision != params[0].precision ||
            params[b].flux_scheme != params[0].flux_scheme)
            return fail(GKM_E_ARG, "box %d: recon, energy_recon, zbug_compat, precision and flux_scheme must match box 0",
                        b);
    }
    // 32-bit byte offsets address every plane of the batch
    const double padded = (double)(grid->nx + 6) * (grid->ny + 6) * (grid->nz + 6);
    if (padded * nbox > 268435456.0)
        return fail(GKM_E_ARG, "batch too large: %.0f padded cells (limit 2^28)", padded * nbox);
    Ctx* h = new (std::nothrow) Ctx();
    if (!h) return fail(GKM_E_NOMEM, "host allocation failed");
    h->grid = *grid;
    h->par = params[0];
    h->bpar.assign(params, params + nbox);
    h->nbox = nbox;
    gkm::Geo& g = h->geo;
    g.nx = grid->nx;
    g.ny = grid->ny;
    g.nz = grid->nz;
    g.P0 = grid->ny + 6;
    g.P1 = grid->nx + 6;
    g.P2 = grid->nz + 6;
    g.s1 = (unsigned)g.P2;
    g.s0 = (unsigned)g.P1 * (unsigned)g.P2;
    g.ncell = (unsigned)g.P0 * g.s0;
    h->ncell = (long)g.ncell;
    hipError_t e = hipGetDevice(&h->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        h->stream = nullptr;
        free_ctx(h);
        return fail(GKM_E_HIP, "no usable HIP device: %s", hipGetErrorString(e));
    }
    const size_t n = (size_t)h->ncell * nbox;
    const size_t align = 256;
    auto up = [&](size_t b) { return (b + align - 1) / align * align; };
    const bool fp64 = params[0].precision == GKM_PREC_FP64;
    // flux blocks, measured at 256^3 over 64 / 128 / 256 threads: the fused Z
    // kernel is fastest at 128 for the mixed form (0.880 vs 0.900 ms at 256)
    // and at 64 for FP64 (1.945 vs 2.017 ms); the X and Y flux kernels at 256
    // (mixed, 0.649 vs 0.663 ms at 64) and 64 (FP64, 1.285 vs 1.302 ms)
    h->zflux_block = fp64 ? 64 : 128;
    h->flux_block = fp64 ? 64 : 256;
    const size_t plane = up(n * (fp64 ? sizeof(double) : sizeof(float)));
    // state (x2: current and next), interface values, fluxes, T and tau (x2)
    const int nplanes = fp64 ? (10 + 10 + 15 + 3) : (20 + 10 + 15 + 3);
    h->red_blocks = 2048;
    const size_t red = up(sizeof(double) * (size_t)h->red_blocks * gkm::kStatA) + up(sizeof(double) * 64);
    const size_t pars = up(sizeof(gkm::StepPar) * nbox);
    const size_t flags = up(sizeof(unsigned) * nbox);
    h->arena_bytes = plane * nplanes + red + pars + flags;
    e = hipMalloc(&h->arena, h->arena_bytes);
    if (e == hipSuccess) e = hipHostMalloc((void**)&h->h_pars, sizeof(gkm::StepPar) * nbox, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&h->h_bad, sizeof(unsigned) * nbox, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_bad, hipEventDisableTiming);
    if (e != hipSuccess) {
        const size_t bytes = h->arena_bytes;
        free_ctx(h);
        return fail(GKM_E_NOMEM, "allocation of %zu device bytes failed: %s", bytes, hipGetErrorString(e));
    }
    char* base = (char*)h->arena;
    auto take = [&](size_t b) {
        char* p = base;
        base += b;
        return p;
    };
    if (fp64) {
        for (int q = 0; q < 5; q++) h->Wd[q] = (double*)take(plane);
        for (int q = 0; q < 5; q++) h->Wd_nx[q] = (double*)take(plane);
        for (int q = 0; q < 10; q++) h->WId[q] = (double*)take(plane);
        for (int q = 0; q < 15; q++) h->Fd[q] = (double*)take(plane);
        h->Td = (double*)take(plane);
        h->taud = (double*)take(plane);
        h->taud_nx = (double*)take(plane);
    } else {
        for (int q = 0; q < 5; q++) h->Whi[q] = (float*)take(plane);
        for (int q = 0; q < 5; q++) h->Wlo[q] = (float*)take(plane);
        for (int q = 0; q < 5; q++) h->Whi_nx[q] = (float*)take(plane);
        for (int q = 0; q < 5; q++) h->Wlo_nx[q] = (float*)take(plane);
        for (int q = 0; q < 10; q++) h->WIf[q] = (float*)take(plane);
        for (int q = 0; q < 15; q++) h->Ff[q] = (float*)take(plane);
        h->Tf = (float*)take(plane);
        h->tauf = (float*)take(plane);
        h->tauf_nx = (float*)take(plane);
    }
    h->red_part = (double*)take(up(sizeof(double) * (size_t)h->red_blocks * gkm::kStatA));
    h->red_out = (double*)take(up(sizeof(double) * 64));
    h->d_pars = (gkm::StepPar*)take(pars);
    h->d_bad = (unsigned*)take(flags);
    std::memset(h->h_bad, 0, sizeof(unsigned) * nbox);
    // interface and flux planes start at zero so that unused border entries are defined
    e = hipMemsetAsync(h->arena, 0, h->arena_bytes, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        free_ctx(h);
        return fail(GKM_E_HIP, "initial clear failed: %s", hipGetErrorString(e));
    }
    int rc = upload_pars(h, nullptr);
    if (rc) {
        free_ctx(h);
        return rc;
    }
    *out = (gkm_t*)h;
    return GKM_OK;
}

int gkm_create(const gkm_grid* grid, const gkm_params* params, gkm_t** out) {
    return gkm_create_batch(grid, params, 1, out);
}

int gkm_get_grid(gkm_t* hh, gkm_grid* out) {
    Ctx* h = ctx(hh);
    if (!h || !out) return fail(GKM_E_ARG, "null argument");
    *out = h->grid;
    return GKM_OK;
}

int gkm_get_params_box(gkm_t* hh, int box, gkm_params* out) {
    Ctx* h = ctx(hh);
    int rc = check_box(h, box);
    if (rc) return rc;
    if (!out) return fail(GKM_E_ARG, "null output");
    *out = h->bpar[box];
    return GKM_OK;
}

int gkm_set_params_box(gkm_t* hh, int box, const gkm_params* p) {
    Ctx* h = ctx(hh);
    int rc = check_box(h, box);
    if (rc) return rc;
    rc = check_params(&h->grid, p);
    if (rc) return rc;
    if (p->recon != h->par.recon || p->energy_recon != h->par.energy_recon || p->zbug_compat != h->par.zbug_compat ||
        p->precision != h->par.precision || p->flux_scheme != h->par.flux_scheme)
        return fail(GKM_E_ARG, "recon, energy_recon, zbug_compat, precision and flux_scheme must match the handle's");
    h->bpar[box] = *p;
    if (box == 0) h->par = *p;
    h->pars_valid = false;
    gkm::invalidate_derived(h);  // Tw and tau depend on the gas constants
    return GKM_OK;
}

int gkm_batch_size(gkm_t* hh, int* nbox) {
    Ctx* h = ctx(hh);
    if (!h || !nbox) return fail(GKM_E_ARG, "null argument");
    *nbox = h->nbox;
    return GKM_OK;
}

int gkm_destroy(gkm_t* hh) {
    Ctx* h = ctx(hh);
    if (!h) return GKM_OK;
    (void)hipStreamSynchronize(h->stream);
    drop_graph(h);
    if (h->events_made)
        for (auto& ev : h->ev_step) (void)hipEventDestroy(ev);
    free_ctx(h);
    return GKM_OK;
}

int gkm_set_state_box(gkm_t* hh, int box, const double* const W[5]) {
    Ctx* h = ctx(hh);
    int rc = check_box(h, box);
    if (rc) return rc;
    if (!W) return fail(GKM_E_ARG, "null state");
    for (int q = 0; q < 5; q++)
        if (!W[q]) return fail(GKM_E_ARG, "null state plane %d", q);
    const size_t bytes = sizeof(double) * (size_t)h->ncell;
    const size_t off = (size_t)box * h->ncell;
    gkm::invalidate_derived(h);
    HIP_TRY(hipMemsetAsync(h->d_bad + box, 0, sizeof(unsigned), h->stream));
    if (h->par.precision == GKM_PREC_FP64) {
        for (int q = 0; q < 5; q++)
            HIP_TRY(hipMemcpyAsync(h->Wd[q] + off, W[q], bytes, hipMemcpyHostToDevice, h->stream));
    } else {
        // stage through the (float) flux planes: 15 float planes of the batch hold one double plane
        double* stage = (double*)h->Ff[0];
        for (int q = 0; q < 5; q++) {
            HIP_TRY(hipMemcpyAsync(stage, W[q], bytes, hipMemcpyHostToDevice, h->stream));
            gkm::launch_split(h, box, q, stage);
        }
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->h_bad[box] = 0;  // every flag copy queued before has landed
    return GKM_OK;
}

int gkm_get_state_box(gkm_t* hh, int box, double* const W[5]) {
    Ctx* h = ctx(hh);
    int rc = check_box(h, box);
    if (rc) return rc;
    if (!W) return fail(GKM_E_ARG, "null state");
    for (int q = 0; q < 5; q++)
        if (!W[q]) return fail(GKM_E_ARG, "null state plane %d", q);
    const size_t bytes = sizeof(double) * (size_t)h->ncell;
    const size_t off = (size_t)box * h->ncell;
    if (h->par.precision == GKM_PREC_FP64) {
        for (int q = 0; q < 5; q++)
            HIP_TRY(hipMemcpyAsync(W[q], h->Wd[q] + off, bytes, hipMemcpyDeviceToHost, h->stream));
    } else {
        double* stage = (double*)h->Ff[0];
        for (int q = 0; q < 5; q++) {
            gkm::launch_merge(h, box, q, stage);
            HIP_TRY(hipMemcpyAsync(W[q], stage, bytes, hipMemcpyDeviceToHost, h->stream));
        }
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    return GKM_OK;
}

int gkm_set_state(gkm_t* h, const double* const W[5]) { return gkm_set_state_box(h, 0, W); }
int gkm_get_state(gkm_t* h, double* const W[5]) { return gkm_get_state_box(h, 0, W); }

int gkm_fill_ghosts(gkm_t* hh) {
    Ctx* h = ctx(hh);
    if (!h) return fail(GKM_E_ARG, "null handle");
    gkm::launch_ghost(h, false);
    HIP_TRY(hipGetLastError());
    return GKM_OK;
}

int gkm_step_batch(gkm_t* hh, const double* dt, int nsteps) {
    Ctx* h = ctx(hh);
    if (!h || !dt) return fail(GKM_E_ARG, "null handle or dt");
    for (int b = 0; b < h->nbox; b++)
        if (!(dt[b] > 0) || !std::isfinite(dt[b])) return fail(GKM_E_ARG, "dt[%d] must be positive and finite", b);
    if (nsteps < 0) return fail(GKM_E_ARG, "nsteps must be >= 0");
    int rc = poll_physical(h);
    if (rc) return rc;
    rc = upload_pars(h, dt);
    if (rc) return rc;
    if (h->timing && !h->events_made) {
        for (auto& ev : h->ev_step) HIP_TRY(hipEventCreate(&ev));
        h->events_made = true;
    }
    h->timed_steps = 0;
    for (int s = 0; s < nsteps; s++) {
        const bool timed = h->timing && s < gkm::kTimedSteps;
        if (timed) {
            gkm::launch_step(h, h->ev_step + s * gkm::kEvPerStep);
            h->timed_steps = s + 1;
        } else if (h->graph_enabled) {
            // a captured step holds the buffer pointers it was captured with:
            // one graph per starting buffer.  The ghost fill a stale state
            // needs runs before, outside the graph, so that the graph holds
            // the same work on every replay.
            gkm::prepare_step(h);
            const int f = h->flip;
            const bool capture = !h->graph_ready[f];
            if (capture) {
                HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
                gkm::launch_step(h, nullptr);  // swaps the pointers for the work it records
                hipGraph_t gr = nullptr;
                hipError_t e = hipStreamEndCapture(h->stream, &gr);
                if (e == hipSuccess) e = hipGraphInstantiate(&h->graph_exec[f], gr, nullptr, nullptr, 0);
                if (e == hipSuccess) e = hipGraphLaunch(h->graph_exec[f], h->stream);
                if (e != hipSuccess) {
                    // the recorded step never ran: the current state is still the one it started from
                    if (h->graph_exec[f]) (void)hipGraphExecDestroy(h->graph_exec[f]);
                    if (gr) (void)hipGraphDestroy(gr);
                    h->graph_exec[f] = nullptr;
                    gkm::swap_state(h);
                    gkm::invalidate_derived(h);
                    return fail(GKM_E_HIP, "captured step failed: %s", hipGetErrorString(e));
                }
                h->graph[f] = gr;
                h->graph_ready[f] = true;
            } else {
                HIP_TRY(hipGraphLaunch(h->graph_exec[f], h->stream));
                gkm::swap_state(h);
            }
        } else {
            gkm::launch_step(h, nullptr);
        }
    }
    gkm::ensure_ghosts(h);
    HIP_TRY(hipGetLastError());
    if (nsteps > 0) return queue_flag_copy(h);
    return GKM_OK;
}

int gkm_set_dt_batch(gkm_t* hh, const double* dt) {
    Ctx* h = ctx(hh);
    if (!h || !dt) return fail(GKM_E_ARG, "null handle or dt");
    for (int b = 0; b < h->nbox; b++)
        if (!(dt[b] > 0) || !std::isfinite(dt[b])) return fail(GKM_E_ARG, "dt[%d] must be positive and finite", b);
    return upload_pars(h, dt);
}

int gkm_step(gkm_t* hh, double dt, int nsteps) {
    Ctx* h = ctx(hh);
    if (!h) return fail(GKM_E_ARG, "null handle");
    if (h->nbox != 1) return fail(GKM_E_ARG, "handle holds %d boxes: use gkm_step_batch", h->nbox);
    return gkm_step_batch(hh, &dt, nsteps);
}

int gkm_set_graph(gkm_t* hh, int enabled) {
    Ctx* h = ctx(hh);
    if (!h) return fail(GKM_E_ARG, "null handle");
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->graph_enabled = enabled != 0;
    if (!h->graph_enabled) drop_graph(h);
    return GKM_OK;
}

int gkm_set_tuning(gkm_t* hh, int flux_block, int recon_chunk) {
    Ctx* h = ctx(hh);
    if (!h) return fail(GKM_E_ARG, "null handle");
    if (flux_block != 64 && flux_block != 128 && flux_block != 256)
        return fail(GKM_E_ARG, "flux_block must be 64, 128 or 256 (got %d)", flux_block);
    if (recon_chunk != 4 && recon_chunk != 8 && recon_chunk != 16)
        return fail(GKM_E_ARG, "recon_chunk must be 4, 8 or 16 (got %d)", recon_chunk);
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->flux_block = flux_block;
    h->recon_chunk = recon_chunk;
    h->tune_gen++;  // groups holding the handle capture their step again
    drop_graph(h);
    return GKM_OK;
}

int gkm_set_tuning_z(gkm_t* hh, int zflux_block, int z_strip) {
    Ctx* h = ctx(hh);
    if (!h) return fail(GKM_E_ARG, "null handle");
    if (zflux_block != 64 && zflux_block != 128 && zflux_block != 256)
        return fail(GKM_E_ARG, "zflux_block must be 64, 128 or 256 (got %d)", zflux_block);
    if (z_strip != 4 && z_strip != 8 && z_strip != 16 && z_strip != 32 && z_strip != 64)
        return fail(GKM_E_ARG, "z_strip must be 4, 8, 16, 32 or 64 (got %d)", z_strip);
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->zflux_block = zflux_block;
    h->z_strip = z_strip;
    h->tune_gen++;  // groups holding the handle capture their step again
    drop_graph(h);
    return GKM_OK;
}

int gkm_sync(gkm_t* hh) {
    Ctx* h = ctx(hh);
    if (!h) return fail(GKM_E_ARG, "null handle");
    return check_physical(h);
}

// ---- groups ----------------------------------------------------------------
// One graph per combination of the handles' current state buffers (the
// WENO-BGK step alternates them; all handles of a group step together, so at
// most two combinations occur).
struct gkm_group {
    std::vector<Ctx*> hs;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, done = nullptr;
    std::vector<hipEvent_t> pre, join;  // per handle: before the group, end of its branch
    struct Cap {
        std::vector<int> flips;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Cap> caps;
    std::vector<unsigned> gens;  // the handles' tuning generations the captures were made with
    // gkm_group_step_persistent: class tables in device memory, one per
    // combination of the handles' current buffers (keyed like the captures)
    struct PTab {
        std::vector<int> flips;
        void* table = nullptr;
    };
    std::vector<PTab> ptabs;
};

static void drop_group_graphs(gkm_group* g) {
    for (auto& c : g->caps) {
        if (c.exec) (void)hipGraphExecDestroy(c.exec);
        if (c.graph) (void)hipGraphDestroy(c.graph);
    }
    g->caps.clear();
    for (auto& t : g->ptabs)
        if (t.table) (void)hipFree(t.table);
    g->ptabs.clear();
}

static void free_group(gkm_group* g) {
    drop_group_graphs(g);
    for (auto e : g->pre)
        if (e) (void)hipEventDestroy(e);
    for (auto e : g->join)
        if (e) (void)hipEventDestroy(e);
    if (g->fork) (void)hipEventDestroy(g->fork);
    if (g->done) (void)hipEventDestroy(g->done);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

int gkm_group_create(gkm_t* const* handles, int n, gkm_group_t** out) {
    if (!out) return fail(GKM_E_ARG, "null output");
    *out = nullptr;
    if (!handles || n < 1) return fail(GKM_E_ARG, "need n >= 1 handles");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    gkm_group* g = new (std::nothrow) gkm_group();
    if (!g) return fail(GKM_E_NOMEM, "host allocation failed");
    g->device = dev;
    for (int i = 0; i < n; i++) {
        Ctx* h = ctx(handles[i]);
        bool dup = false;
        for (Ctx* o : g->hs) dup |= o == h;
        if (!h || dup || h->device != dev) {
            delete g;
            return fail(GKM_E_ARG, "handle %d is null, repeated or on another device", i);
        }
        g->hs.push_back(h);
    }
    hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->done, hipEventDisableTiming);
    g->pre.assign(n, nullptr);
    g->join.assign(n, nullptr);
    for (int i = 0; i < n && e == hipSuccess; i++) {
        e = hipEventCreateWithFlags(&g->pre[i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&g->join[i], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        free_group(g);
        return fail(GKM_E_HIP, "group set-up failed: %s", hipGetErrorString(e));
    }
    *out = g;
    return GKM_OK;
}

int gkm_group_step(gkm_group_t* g, int nsteps) {
    if (!g) return fail(GKM_E_ARG, "null group");
    if (nsteps < 0) return fail(GKM_E_ARG, "nsteps must be >= 0");
    const int n = (int)g->hs.size();
    // launch shapes changed since the captures (gkm_set_tuning): capture again
    std::vector<unsigned> gens(n);
    for (int i = 0; i < n; i++) gens[i] = g->hs[i]->tune_gen;
    if (gens != g->gens) {
        HIP_TRY(hipStreamSynchronize(g->stream));
        drop_group_graphs(g);
        g->gens = gens;
    }
    for (int i = 0; i < n; i++) {
        Ctx* h = g->hs[i];
        int rc = poll_physical(h);
        if (rc) return rc;
        // every box needs a dt; constants changed since the last upload are
        // uploaded with the dt values last given
        std::vector<double> dt(h->nbox);
        for (int b = 0; b < h->nbox; b++) {
            dt[b] = h->h_pars[b].dt;
            if (!(dt[b] > 0))
                return fail(GKM_E_ARG, "handle %d box %d has no dt: call gkm_set_dt_batch or gkm_step_batch first", i,
                            b);
        }
        if (!h->pars_valid) {
            int rc = upload_pars(h, dt.data());
            if (rc) return rc;
        }
        gkm::prepare_step(h);  // outside the graph, as for a single handle
        // the group's work follows what is queued on the handle
        HIP_TRY(hipEventRecord(g->pre[i], h->stream));
        HIP_TRY(hipStreamWaitEvent(g->stream, g->pre[i], 0));
    }
    for (int s = 0; s < nsteps; s++) {
        // a capture is keyed by each handle's buffer parity and halo mode
        std::vector<int> flips(n);
        for (int i = 0; i < n; i++) flips[i] = g->hs[i]->flip + 2 * g->hs[i]->halo_mode;
        gkm_group::Cap* cap = nullptr;
        for (auto& c : g->caps)
            if (c.flips == flips) cap = &c;
        if (!cap) {
            // capture one step of every handle, each on its own stream as a branch
            gkm_group::Cap c;
            c.flips = flips;
            HIP_TRY(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
            (void)hipEventRecord(g->fork, g->stream);
            for (int i = 0; i < n; i++) {
                Ctx* h = g->hs[i];
                (void)hipStreamWaitEvent(h->stream, g->fork, 0);
                gkm::launch_step(h, nullptr);  // swaps the pointers for the work it records
                (void)hipEventRecord(g->join[i], h->stream);
                (void)hipStreamWaitEvent(g->stream, g->join[i], 0);
            }
            hipError_t e = hipStreamEndCapture(g->stream, &c.graph);
            if (e == hipSuccess) e = hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0);
            if (e == hipSuccess) e = hipGraphLaunch(c.exec, g->stream);
            if (e != hipSuccess) {
                // the recorded steps never ran: every handle keeps the state it started from
                if (c.exec) (void)hipGraphExecDestroy(c.exec);
                if (c.graph) (void)hipGraphDestroy(c.graph);
                for (Ctx* h : g->hs) {
                    gkm::swap_state(h);
                    gkm::invalidate_derived(h);
                }
                return fail(GKM_E_HIP, "group step capture failed: %s", hipGetErrorString(e));
            }
            g->caps.push_back(c);
        } else {
            HIP_TRY(hipGraphLaunch(cap->exec, g->stream));
            for (Ctx* h : g->hs) gkm::swap_state(h);
        }
    }
    // the handles' later work follows the group's; ghosts filled for reads
    HIP_TRY(hipEventRecord(g->done, g->stream));
    for (Ctx* h : g->hs) {
        HIP_TRY(hipStreamWaitEvent(h->stream, g->done, 0));
        gkm::ensure_ghosts(h);
        if (nsteps > 0) {
            int rc = queue_flag_copy(h);
            if (rc) return rc;
        }
    }
    HIP_TRY(hipGetLastError());
    return GKM_OK;
}

// One cooperative launch of the persistent kernel (k_group_persist) advances
// every box of every handle nsteps steps.  The handles must share the
// precision, reconstruction and energy mode, run the WENO-BGK scheme and keep
// the periodic fill; the per-stage path (gkm_group_step) has no such limits.
int gkm_group_step_persistent(gkm_group_t* g, int nsteps) {
    if (!g) return fail(GKM_E_ARG, "null group");
    if (nsteps < 0) return fail(GKM_E_ARG, "nsteps must be >= 0");
    const int n = (int)g->hs.size();
    if (n > gkm::kPersistMaxClasses)
        return fail(GKM_E_ARG, "the persistent step takes at most %d handles (got %d)", gkm::kPersistMaxClasses, n);
    const gkm_params& p0 = g->hs[0]->par;
    for (int i = 0; i < n; i++) {
        const gkm_params& p = g->hs[i]->par;
        if (p.flux_scheme != GKM_SCHEME_WENO_BGK || g->hs[i]->halo_mode != GKM_HALO_GATHER)
            return fail(GKM_E_ARG, "handle %d: the persistent step runs the WENO-BGK scheme with the periodic fill "
                                   "(GKM_HALO_GATHER); use gkm_group_step", i);
        if (p.precision != p0.precision || p.recon != p0.recon || p.energy_recon != p0.energy_recon)
            return fail(GKM_E_ARG, "handle %d: the persistent step needs one precision, reconstruction and energy "
                                   "mode across the group; use gkm_group_step", i);
    }
    // launch shapes changed since the tables were made (gkm_set_tuning)
    std::vector<unsigned> gens(n);
    for (int i = 0; i < n; i++) gens[i] = g->hs[i]->tune_gen;
    if (gens != g->gens) {
        HIP_TRY(hipStreamSynchronize(g->stream));
        drop_group_graphs(g);
        g->gens = gens;
    }
    for (int i = 0; i < n; i++) {
        Ctx* h = g->hs[i];
        int rc = poll_physical(h);
        if (rc) return rc;
        std::vector<double> dt(h->nbox);
        for (int b = 0; b < h->nbox; b++) {
            dt[b] = h->h_pars[b].dt;
            if (!(dt[b] > 0))
                return fail(GKM_E_ARG, "handle %d box %d has no dt: call gkm_set_dt_batch or gkm_step_batch first", i,
                            b);
        }
        if (!h->pars_valid) {
            rc = upload_pars(h, dt.data());
            if (rc) return rc;
        }
        gkm::prepare_step(h);
        HIP_TRY(hipEventRecord(g->pre[i], h->stream));
        HIP_TRY(hipStreamWaitEvent(g->stream, g->pre[i], 0));
    }
    if (nsteps > 0) {
        std::vector<int> flips(n);
        for (int i = 0; i < n; i++) flips[i] = g->hs[i]->flip;
        gkm_group::PTab* tab = nullptr;
        for (auto& t : g->ptabs)
            if (t.flips == flips) tab = &t;
        if (!tab) {
            const size_t one = gkm::persist_class_bytes(p0.precision);
            std::vector<unsigned char> host(one * (size_t)n);
            for (int i = 0; i < n; i++) gkm::persist_fill_class(g->hs[i], host.data() + one * (size_t)i);
            gkm_group::PTab t;
            t.flips = flips;
            HIP_TRY(hipMalloc(&t.table, host.size()));
            const hipError_t e = hipMemcpy(t.table, host.data(), host.size(), hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                (void)hipFree(t.table);
                return fail(GKM_E_HIP, "class table copy failed: %s", hipGetErrorString(e));
            }
            g->ptabs.push_back(t);
            tab = &g->ptabs.back();
        }
        int grid = 0;
        const hipError_t e = gkm::launch_persist(p0, tab->table, n, nsteps, g->stream, &grid);
        if (e != hipSuccess) return fail(GKM_E_HIP, "persistent launch failed: %s", hipGetErrorString(e));
        for (Ctx* h : g->hs)
            for (int s = 0; s < nsteps; s++) gkm::swap_state(h);
    }
    // the handles' later work follows the group's; ghosts filled for reads
    HIP_TRY(hipEventRecord(g->done, g->stream));
    for (Ctx* h : g->hs) {
        HIP_TRY(hipStreamWaitEvent(h->stream, g->done, 0));
        gkm::ensure_ghosts(h);
        if (nsteps > 0) {
            int rc = queue_flag_copy(h);
            if (rc) return rc;
        }
    }
    HIP_TRY(hipGetLastError());
    return GKM_OK;
}

int gkm_group_sync(gkm_group_t* g) {
    if (!g) return fail(GKM_E_ARG, "null group");
    HIP_TRY(hipStreamSynchronize(g->stream));
    for (Ctx* h : g->hs) {
        int rc = check_physical(h);
        if (rc) return rc;
    }
    return GKM_OK;
}

int gkm_group_destroy(gkm_group_t* g) {
    if (!g) return GKM_OK;
    (void)hipStreamSynchronize(g->stream);
    free_group(g);
    return GKM_OK;
}

int gkm_set_timing(gkm_t* hh, int mode) {
    Ctx* h = ctx(hh);
    if (!h) return fail(GKM_E_ARG, "null handle");
    h->timing = mode != GKM_TIMING_OFF;
    h->ev_mask = !h->timing ? 0u : mode == GKM_TIMING_FLUX ? gkm::kEvFluxXY : (1u << gkm::kEvPerStep) - 1u;
    return GKM_OK;
}

int gkm_last_timing(gkm_t* hh, double* ms, int* nsteps) {
    Ctx* h = ctx(hh);
    if (!h || !ms) return fail(GKM_E_ARG, "null argument");
    for (int i = 0; i < 5; i++) ms[i] = -1.0;
    if (nsteps) *nsteps = h->timed_steps;
    if (!h->timing || h->timed_steps == 0) return GKM_OK;
    HIP_TRY(hipStreamSynchronize(h->stream));
    double flux = 0, recon = 0, ghost = 0, upd = 0, prim = 0;
    if (h->ev_mask == gkm::kEvFluxXY) {  // GKM_TIMING_FLUX: the X and Y flux kernels only
        for (int s = 0; s < h->timed_steps; s++) {
            hipEvent_t* ev = h->ev_step + s * gkm::kEvPerStep;
            float t;
            for (int d = 0; d < 2; d++) {
                HIP_TRY(hipEventElapsedTime(&t, ev[2 + 3 * d], ev[3 + 3 * d]));
                flux += t;
            }
        }
        ms[0] = flux;
        return GKM_OK;
    }
    for (int s = 0; s < h->timed_steps; s++) {
        hipEvent_t* ev = h->ev_step + s * gkm::kEvPerStep;
        float t;
        HIP_TRY(hipEventElapsedTime(&t, ev[0], ev[12]));
        ghost += t;
        HIP_TRY(hipEventElapsedTime(&t, ev[12], ev[1]));
        prim += t;
        for (int d = 0; d < 3; d++) {
            HIP_TRY(hipEventElapsedTime(&t, ev[1 + 3 * d], ev[2 + 3 * d]));
            recon += t;
            HIP_TRY(hipEventElapsedTime(&t, ev[2 + 3 * d], ev[3 + 3 * d]));
            flux += t;
        }
        HIP_TRY(hipEventElapsedTime(&t, ev[9], ev[11]));
        upd += t;
    }
    ms[0] = flux;
    ms[1] = recon;
    ms[2] = ghost;
    ms[3] = upd;
    ms[4] = prim;
    return GKM_OK;
}

int gkm_last_sweep_timing(gkm_t* hh, double* ms) {
    Ctx* h = ctx(hh);
    if (!h || !ms) return fail(GKM_E_ARG, "null argument");
    for (int d = 0; d < 3; d++) ms[d] = -1.0;
    if (!h->timing || h->timed_steps == 0) return GKM_OK;
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (int d = 0; d < 3; d++) {
        if (!(h->ev_mask >> (3 + 3 * d) & 1u)) continue;  // not recorded in this mode
        double sum = 0;
        for (int s = 0; s < h->timed_steps; s++) {
            hipEvent_t* ev = h->ev_step + s * gkm::kEvPerStep;
            float t;
            HIP_TRY(hipEventElapsedTime(&t, ev[2 + 3 * d], ev[3 + 3 * d]));
            sum += t;
        }
        ms[d] = sum;
    }
    return GKM_OK;
}

int gkm_set_halo_mode(gkm_t* hh, int mode) {
    Ctx* h = ctx(hh);
    if (!h) return fail(GKM_E_ARG, "null handle");
    if (mode != GKM_HALO_GATHER && mode != GKM_HALO_PACK) return fail(GKM_E_ARG, "bad halo mode %d", mode);
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (mode == GKM_HALO_PACK && !h->halo_send) {
        // two 3-plane y faces of every plane (5 FP64 or 10 FP32) of every box
        const size_t elem = h->par.precision == GKM_PREC_FP64 ? sizeof(double) : 2 * sizeof(float);
        const size_t bytes = (size_t)2 * 3 * h->geo.s0 * 5 * elem * h->nbox;
        HIP_TRY(hipMalloc(&h->halo_send, bytes));
        HIP_TRY(hipMalloc(&h->halo_recv, bytes));
    }
    h->halo_mode = mode;
    gkm::invalidate_derived(h);
    drop_graph(h);
    return GKM_OK;
}

int gkm_compute_dt_box(gkm_t* hh, int box, double cfl, double* dt) {
    Ctx* h = ctx(hh);
    int rc = check_box(h, box);
    if (rc) return rc;
    if (!dt) return fail(GKM_E_ARG, "null argument");
    if (!(cfl > 0)) return fail(GKM_E_ARG, "cfl must be positive");
    double a[gkm::kStatA];
    HIP_TRY(gkm::launch_stats_a(h, box, a));
    rc = check_physical(h);
    if (rc) return rc;
    const double mx = a[27];
    if (!(mx > 0) || !std::isfinite(mx)) return fail(GKM_E_NONPHYSICAL, "signal speed %g is not positive/finite", mx);
    *dt = cfl * h->grid.dx / mx;
    return GKM_OK;
}

int gkm_compute_dt(gkm_t* h, double cfl, double* dt) { return gkm_compute_dt_box(h, 0, cfl, dt); }

int gkm_stats_box(gkm_t* hh, int box, gkm_stats_t* out) {
    Ctx* h = ctx(hh);
    int rc = check_box(h, box);
    if (rc) return rc;
    if (!out) return fail(GKM_E_ARG, "null argument");
    gkm::ensure_ghosts(h);
    double a[gkm::kStatA];
    HIP_TRY(gkm::launch_stats_a(h, box, a));
    const double N = (double)h->grid.nx * h->grid.ny * h->grid.nz;
    gkm_stats_t s;
    std::memset(&s, 0, sizeof s);
    s.rho_mean = a[0] / N;
    s.p_mean = a[1] / N;
    s.T_mean = a[2] / N;
    s.v_mean = a[3] / N;
    s.c_mean = a[4] / N;
    s.q2 = a[5] / N;
    s.tke = a[6] / N;
    s.ek = 0.5 * s.tke;
    s.mt = std::sqrt(s.q2) / s.c_mean;
    s.mass = a[7];
    s.mom_x = a[8];
    s.mom_y = a[9];
    s.mom_z = a[10];
    s.energy = a[11];
    s.dudx2 = a[12] / N; s.dudx3 = a[13] / N; s.dudx4 = a[14] / N;
    s.dvdy2 = a[15] / N; s.dvdy3 = a[16] / N; s.dvdy4 = a[17] / N;
    s.dwdz2 = a[18] / N; s.dwdz3 = a[19] / N; s.dwdz4 = a[20] / N;
    s.enstrophy = a[21] / N;
    s.dil2 = a[22] / N;
    s.diss_sol = a[23] / N;
    s.diss_dil = (4.0 / 3.0) * a[24] / N;
    s.diss_ref = a[25] / N;
    s.sijsij = a[26] / N;
    s.umax_plus_c = a[27];
    s.skew_u = s.dudx3 / std::pow(s.dudx2, 1.5);
    s.skew_v = s.dvdy3 / std::pow(s.dvdy2, 1.5);
    s.skew_w = s.dwdz3 / std::pow(s.dwdz2, 1.5);
    s.kurt_u = s.dudx4 / (s.dudx2 * s.dudx2);
    s.kurt_v = s.dvdy4 / (s.dvdy2 * s.dvdy2);
    s.kurt_w = s.dwdz4 / (s.dwdz2 * s.dwdz2);
    const double m[4] = {s.rho_mean, s.p_mean, s.T_mean, s.v_mean};
    double b[gkm::kStatB];
    HIP_TRY(gkm::launch_stats_b(h, box, m, b));
    s.rho_rms = std::sqrt(b[0] / N);
    s.p_rms = std::sqrt(b[1] / N);
    s.T_rms = std::sqrt(b[2] / N);
    s.v_rms = std::sqrt(b[3] / N);
    *out = s;
    rc = check_physical(h);
    if (rc) return rc;
    if (!(s.rho_mean > 0) || !(s.p_mean > 0) || !std::isfinite(s.energy))
        return fail(GKM_E_NONPHYSICAL, "non-physical state: <rho>=%g <p>=%g", s.rho_mean, s.p_mean);
    return GKM_OK;
}

int gkm_stats(gkm_t* h, gkm_stats_t* out) { return gkm_stats_box(h, 0, out); }

// the nine gradient planes of one box, the Thomas pivots of the three line
// lengths and the backward-sum weights of the y and x lines, allocated on
// first use (the pinned row last: a call that failed part way allocates the
// rest on the next one)
static int ensure_spline_buffers(Ctx* h) {
    if (h->ps_row) return GKM_OK;
    const gkm::Geo& g = h->geo;
    // Thomas pivots of the natural-spline slope system for each line length
    std::vector<double> inv;
    for (int m : {g.P0, g.P1, g.P2}) {
        double prev = 0.0;
        for (int i = 0; i < m; i++) {
            const double b = (i == 0 || i == m - 1) ? 2.0 : 4.0;
            prev = 1.0 / (i == 0 ? b : b - prev);
            inv.push_back(prev);
        }
    }
    // D_i = sum_j W[i][j] y_{i+j}: W[i][j] = (-1)^j inv_i ... inv_{i+j-1} for
    // i + j <= m - 1 (D_{m-1} = y_{m-1}), 0 beyond; rows of gkm::kSplRing
    constexpr int R = gkm::kSplRing;
    std::vector<double> wt;
    size_t first = 0;
    for (int m : {g.P0, g.P1}) {
        for (int i = 0; i < m; i++) {
            double w = 1.0;
            for (int j = 0; j < R; j++) {
                if (i + j <= m - 1) {
                    wt.push_back(w);
                    w *= -inv[first + i + j];
                } else {
                    wt.push_back(0.0);
                }
            }
        }
        first += (size_t)m;
    }
    if (!h->spl_grad) {
        hipError_t e = hipMalloc((void**)&h->spl_grad, sizeof(double) * 9 * (size_t)h->ncell);
        if (e != hipSuccess) {
            h->spl_grad = nullptr;
            return fail(GKM_E_NOMEM, "gradient planes (%zu bytes): %s", sizeof(double) * 9 * (size_t)h->ncell,
                        hipGetErrorString(e));
        }
    }
    if (!h->spl_inv) HIP_TRY(hipMalloc((void**)&h->spl_inv, sizeof(double) * inv.size()));
    HIP_TRY(hipMemcpy(h->spl_inv, inv.data(), sizeof(double) * inv.size(), hipMemcpyHostToDevice));
    if (!h->spl_wt) HIP_TRY(hipMalloc((void**)&h->spl_wt, sizeof(double) * wt.size()));
    HIP_TRY(hipMemcpy(h->spl_wt, wt.data(), sizeof(double) * wt.size(), hipMemcpyHostToDevice));
    if (!h->spl_part) HIP_TRY(hipMalloc((void**)&h->spl_part, sizeof(double) * gkm::print_stats_part_doubles(g)));
    HIP_TRY(hipHostMalloc((void**)&h->ps_row, sizeof(double) * (gkm::kPStat + gkm::kStatB), hipHostMallocDefault));
    return GKM_OK;
}

// the PrintStats quantities from the device sums a (kPStat) and the squared
// deviations b (kStatB) about the p, T and 1/rho means
static int fill_print_stats(const Ctx* h, const double* a, const double* b, gkm_print_stats_t* out) {
    const double N = (double)h->grid.nx * h->grid.ny * h->grid.nz;
    gkm_print_stats_t s;
    std::memset(&s, 0, sizeof s);
    s.tke = a[0] / N;
    s.diss = a[1] / N;
    s.sijsij = a[2] / N;
    const double u2 = a[3] / N, u3 = a[4] / N, u4 = a[5] / N;
    const double v2 = a[6] / N, v3 = a[7] / N, v4 = a[8] / N;
    const double w2 = a[9] / N, w3 = a[10] / N, w4 = a[11] / N;
    s.skew_u = u3 / std::pow(u2, 1.5);
    s.skew_v = v3 / std::pow(v2, 1.5);
    s.skew_w = w3 / std::pow(w2, 1.5);
    s.kurt_u = u4 / (u2 * u2);
    s.kurt_v = v4 / (v2 * v2);
    s.kurt_w = w4 / (w2 * w2);
    s.dudx2 = u2;
    s.dvdy2 = v2;
    s.dwdz2 = w2;
    s.p_mean = a[12] / N;
    s.T_mean = a[13] / N;
    s.v_mean = a[14] / N;
    s.p_rms = std::sqrt(b[1] / N);
    s.T_rms = std::sqrt(b[2] / N);
    s.v_rms = std::sqrt(b[3] / N);
    *out = s;
    if (!(s.p_mean > 0) || !(s.v_mean > 0) || !std::isfinite(s.tke))
        return fail(GKM_E_NONPHYSICAL, "non-physical state: <p>=%g <1/rho>=%g", s.p_mean, s.v_mean);
    return GKM_OK;
}

int gkm_print_stats_box(gkm_t* hh, int box, gkm_print_stats_t* out) {
    Ctx* h = ctx(hh);
    int rc = check_box(h, box);
    if (rc) return rc;
    if (!out) return fail(GKM_E_ARG, "null argument");
    rc = ensure_spline_buffers(h);
    if (rc) return rc;
    gkm::ensure_ghosts(h);
    // the sums and the non-physical flags into pinned memory, one wait for both
    double* a = h->ps_row;
    HIP_TRY(gkm::launch_print_stats_async(h, box, h->stream, a));
    HIP_TRY(hipMemcpyAsync(h->h_bad, h->d_bad, sizeof(unsigned) * h->nbox, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    rc = bad_report(h);
    const int rs = fill_print_stats(h, a, a + gkm::kPStat, out);
    return rc ? rc : rs;
}

// the loop of gkm_step_print_stats: statistics of state s on stream2 beside step s
static int step_stats_loop(gkm_t* hh, Ctx* h, double dt, int nsteps) {
    constexpr int row = gkm::kPStat + gkm::kStatB;
    gkm::prepare_step(h);  // ghosts (and primitives) of the first state, on the step stream
    for (int s = 0; s < nsteps; s++) {
        // statistics of state s on the second stream, once the step that wrote it is done
        HIP_TRY(hipEventRecord(h->ev_state, h->stream));
        HIP_TRY(hipStreamWaitEvent(h->stream2, h->ev_state, 0));
        HIP_TRY(gkm::launch_print_stats_async(h, 0, h->stream2, h->ps_host + (size_t)row * s));
        HIP_TRY(hipEventRecord(h->ev_ps[s & 1], h->stream2));
        // step s reads state s (as the statistics do) and writes the other buffer,
        // which the statistics of state s-1 read: it waits for those only
        if (s > 0) HIP_TRY(hipStreamWaitEvent(h->stream, h->ev_ps[(s - 1) & 1], 0));
        const int rc = gkm_step_batch(hh, &dt, 1);
        if (rc) return rc;
    }
    // the handle's later work (which may write the buffer the last statistics read) follows them
    if (nsteps > 0) HIP_TRY(hipStreamWaitEvent(h->stream, h->ev_ps[(nsteps - 1) & 1], 0));
    return GKM_OK;
}

int gkm_step_print_stats(gkm_t* hh, double dt, int nsteps, gkm_print_stats_t* out) {
    Ctx* h = ctx(hh);
    if (!h) return fail(GKM_E_ARG, "null handle");
    if (h->nbox != 1) return fail(GKM_E_ARG, "handle holds %d boxes: gkm_step_print_stats takes one box", h->nbox);
    if (nsteps < 0) return fail(GKM_E_ARG, "nsteps must be >= 0");
    if (nsteps > 0 && !out) return fail(GKM_E_ARG, "null argument");
    if (!(dt > 0) || !std::isfinite(dt)) return fail(GKM_E_ARG, "dt must be positive and finite");
    int rc = ensure_spline_buffers(h);
    if (rc) return rc;
    if (h->halo_mode == GKM_HALO_PACK) {
        // every step starts with the exchange, which rewrites the ghosts the
        // statistics read: no overlap, the loop as written
        for (int s = 0; s < nsteps; s++) {
            rc = gkm_print_stats_box(hh, 0, out + s);
            if (rc) return rc;
            rc = gkm_step_batch(hh, &dt, 1);
            if (rc) return rc;
        }
        return GKM_OK;
    }
    // created on first use (each one that a failed earlier attempt left missing)
    if (!h->stream2) HIP_TRY(hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking));
    if (!h->ev_state) HIP_TRY(hipEventCreateWithFlags(&h->ev_state, hipEventDisableTiming));
    for (hipEvent_t& e : h->ev_ps)
        if (!e) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    constexpr int row = gkm::kPStat + gkm::kStatB;
    if (h->ps_rows < nsteps) {
        if (h->ps_host) HIP_TRY(hipHostFree(h->ps_host));
        h->ps_host = nullptr;
        h->ps_rows = 0;
        HIP_TRY(hipHostMalloc((void**)&h->ps_host, sizeof(double) * row * (size_t)nsteps, hipHostMallocDefault));
        h->ps_rows = nsteps;
    }
    rc = upload_pars(h, &dt);
    if (rc) return rc;
    rc = step_stats_loop(hh, h, dt, nsteps);
    // nothing may still write the pinned rows once the call has returned
    const hipError_t e2 = hipStreamSynchronize(h->stream2);
    if (rc) return rc;
    if (e2 != hipSuccess) return fail(GKM_E_HIP, "statistics stream: %s", hipGetErrorString(e2));
    for (int s = 0; s < nsteps; s++) {
        const double* a = h->ps_host + (size_t)row * s;
        rc = fill_print_stats(h, a, a + gkm::kPStat, out + s);
        if (rc) return rc;
    }
    return poll_physical(h);
}

int gkm_print_stats(gkm_t* h, gkm_print_stats_t* out) { return gkm_print_stats_box(h, 0, out); }

int gkm_format_print_stats(const gkm_print_stats_t* s, double t_over_t0, double gamma, double p0, double T0,
                           double Mt, double den0, char* buf, int len) {
    if (!s || !buf || len <= 0) return fail(GKM_E_ARG, "null argument or empty buffer");
    const double prms = s->p_rms / (gamma * p0 * Mt * Mt);
    const double trms = s->T_rms / ((gamma - 1) * T0 * Mt * Mt);
    const double vrms = s->v_rms / (Mt * Mt / den0);
    // column separators as update.cpp:779-780 writes them (two spaces, four
    // before Vrms)
    const int w = snprintf(buf, (size_t)len, "%.8e  %.8e  %.8e  %.8e  %.8e  %.8e  %.8e  %.8e  %.8e  %.8e  %.8e    %.8e",
                           t_over_t0, s->tke, s->diss, s->skew_u, s->skew_v, s->skew_w, s->kurt_u, s->kurt_v,
                           s->kurt_w, prms, trms, vrms);
    if (w < 0 || w >= len) return fail(GKM_E_ARG, "buffer of %d bytes too small for the %d-byte line", len, w);
    return GKM_OK;
}

static int points_common(const gkm_params* p, int n, const double* in, size_t in_per, double* out, size_t out_per,
                         double** din, double** dout, hipStream_t* s) {
    if (!p || !in || !out || n < 0) return fail(GKM_E_ARG, "null argument or negative count");
    HIP_TRY(hipStreamCreateWithFlags(s, hipStreamNonBlocking));
    HIP_TRY(hipMalloc(din, sizeof(double) * in_per * (size_t)(n > 0 ? n : 1)));
    HIP_TRY(hipMalloc(dout, sizeof(double) * out_per * (size_t)(n > 0 ? n : 1)));
    HIP_TRY(hipMemcpyAsync(*din, in, sizeof(double) * in_per * n, hipMemcpyHostToDevice, *s));
    return GKM_OK;
}

int gkm_flux_points(const gkm_params* p, int n, const double* in, double dx, double* F) {
    double *din = nullptr, *dout = nullptr;
    hipStream_t s = nullptr;
    int rc = points_common(p, n, in, 52, F, 5, &din, &dout, &s);
    if (rc == GKM_OK && n > 0) {
        gkm::launch_flux_points(s, p->precision, n, din, dx, p->gamma, 1.0 / p->Pr, p->K, dout);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(F, dout, sizeof(double) * 5 * n, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = fail(GKM_E_HIP, "flux points: %s", hipGetErrorString(e));
    }
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    if (s) (void)hipStreamDestroy(s);
    return rc;
}

int gkm_weno_points(const gkm_params* p, int n, const double* in, double* out) {
    double *din = nullptr, *dout = nullptr;
    hipStream_t s = nullptr;
    int rc = points_common(p, n, in, 30, out, 10, &din, &dout, &s);
    if (rc == GKM_OK && n > 0) {
        gkm::launch_weno_points(s, p->precision, p->recon, n, din, p->gamma, p->R, p->energy_recon, dout);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(double) * 10 * n, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = fail(GKM_E_HIP, "weno points: %s", hipGetErrorString(e));
    }
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    if (s) (void)hipStreamDestroy(s);
    return rc;
}

int gkm_gks_points(const gkm_params* p, int n, const double* in, double* F) {
    double *din = nullptr, *dout = nullptr;
    hipStream_t s = nullptr;
    int rc = points_common(p, n, in, 22, F, 5, &din, &dout, &s);
    if (rc == GKM_OK && n > 0) {
        gkm::launch_gks_points(s, p->precision, n, din, p->gamma, 1.0 / p->Pr, p->K, dout);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(F, dout, sizeof(double) * 5 * n, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = fail(GKM_E_HIP, "gks points: %s", hipGetErrorString(e));
    }
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    if (s) (void)hipStreamDestroy(s);
    return rc;
}

}  // extern "C"
