// gkm_internal.h — handle layout and launcher declarations shared by the
// translation units of libgkm.so.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/gkm.h"
#include "gkm_kernels.h"

namespace gkm {

constexpr int kTimedSteps = 64;  // steps per gkm_step call that carry HIP events
// events of one step: 0 start, 12 after the ghost fill, 1 after the primitives,
// 1+3d / 2+3d / 3+3d around sweep d's WENO and flux launches, 11 after the update
constexpr int kEvPerStep = 13;
// events of a step: 0 start, 12 after the ghost fill, 1 after the primitives;
// sweep d (X, Y, Z) starts at 1 + 3 d: 2 + 3 d after its reconstruction, 3 + 3 d
// after its flux (for Z the fused flux + update); 11 end.  GKM_TIMING_FLUX
// records only 2, 3, 5, 6 (the X and Y flux kernels).
constexpr unsigned kEvFluxXY = (1u << 2) | (1u << 3) | (1u << 5) | (1u << 6);

struct Ctx {
    gkm_grid grid;
    gkm_params par;                 // box 0's parameters (shared options)
    std::vector<gkm_params> bpar;   // per-box parameters
    Geo geo;
    long ncell;  // padded cells of one box
    int nbox = 1;
    int device;
    hipStream_t stream;
    int flux_block = 256;  // X and Y flux kernels (gkm_set_tuning)
    int zflux_block = 128;  // the fused Z kernel: 128 (mixed) or 64 (FP64) at creation (gkm_set_tuning_z)
    int z_strip = 16;       // the fused Z kernel's z lines go in strips of this many columns of i
    int recon_chunk = 8;
    unsigned tune_gen = 0;  // counts gkm_set_tuning calls (group captures are keyed by it)
    // per-box step parameters: device array read by the kernels, pinned staging
    StepPar* d_pars = nullptr;
    StepPar* h_pars = nullptr;
    bool pars_valid = false;
    // captured steps, one per state buffer the step starts from (flip)
    bool graph_enabled = false;
    bool graph_ready[2] = {false, false};
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    // The WENO-BGK step reads the state from one buffer and writes the updated
    // interior to the other (the Z sweep's flux kernel applies the update);
    // Wd / Whi / Wlo always name the current state, the *_nx arrays the other
    // buffer, and flip counts the swaps mod 2.
    int flip = 0;
    double* Wd_nx[5] = {};
    float* Whi_nx[5] = {};
    float* Wlo_nx[5] = {};
    // FP64 form
    double* Wd[5] = {};
    double* WId[10] = {};
    double* Fd[15] = {};
    double* Td = nullptr;    // WENO temperature input per cell
    double* taud = nullptr;  // left-cell collision time mu/p per cell
    double* taud_nx = nullptr;  // the other buffer's (the Z sweep reads taud while writing it)
    // mixed form
    float* Whi[5] = {};
    float* Wlo[5] = {};
    float* WIf[10] = {};
    float* Ff[15] = {};
    float* Tf = nullptr;
    float* tauf = nullptr;
    float* tauf_nx = nullptr;
    // the current buffer's ghost cells, and for the WENO-BGK scheme its Tw and
    // tau planes, are those of its interior (written by the last step's Z sweep
    // or by a ghost fill); anything that changes the state or the gas
    // constants clears it (invalidate_derived)
    bool derived_valid = false;
    // GKM_HALO_PACK: the current buffer's Tw and tau hold its interior cells'
    // values (the last WENO-BGK step's Z sweep wrote them), so a step only
    // recomputes the ghost shell's after the exchange; cleared with
    // derived_valid
    bool prims_valid = false;
    // non-physical flags, one per box, raised by the fused Z kernel on a new
    // cell with rho or T not positive and finite; cleared by gkm_set_state_box.
    // Every gkm_step_batch call ends with a copy of the flags to the pinned
    // h_bad (event ev_bad), which the next call reads without waiting.
    unsigned* d_bad = nullptr;
    unsigned* h_bad = nullptr;
    hipEvent_t ev_bad = nullptr;
    bool bad_copied = false;
    // one device allocation holds every plane
    void* arena = nullptr;
    size_t arena_bytes = 0;
    // reductions
    double* red_part = nullptr;  // partial sums per block
    double* red_out = nullptr;   // final values
    int red_blocks = 0;
    // PrintStats: nine FP64 gradient planes of one box and the Thomas pivots
    // of the three line lengths (P0, P1, P2 values), allocated on first use
    double* spl_grad = nullptr;
    double* spl_inv = nullptr;
    double* spl_wt = nullptr;  // backward-sum weights of the y (P0 rows) and x (P1 rows) lines, 32 per row
    double* spl_part = nullptr;  // partial-sum rows of the fused PrintStats passes (print_stats_part_doubles)
    // gkm_step_print_stats: a second stream (default priority) for the
    // statistics, its events and the pinned host rows of the sums
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_state = nullptr, ev_ps[2] = {nullptr, nullptr};
    double* ps_host = nullptr;
    int ps_rows = 0;
    double* ps_row = nullptr;  // pinned: the row of one gkm_print_stats call
    // halo mode and the y-face exchange buffers of GKM_HALO_PACK
    int halo_mode = GKM_HALO_GATHER;
    void* halo_send = nullptr;
    void* halo_recv = nullptr;
    // timing
    bool timing = false;
    unsigned ev_mask = 0;  // which of the kEvPerStep events of a timed step are recorded
    int timed_steps = 0;
    hipEvent_t ev_step[kTimedSteps * kEvPerStep];
    bool events_made = false;
};

// sets the message gkm_last_error() returns on this thread; returns code
int report_error(int code, const char* msg);

StepPar step_par(const gkm_grid& grid, const gkm_params& p, double dt);
void launch_ghost(Ctx* h, bool prim);  // prim = false: ghosts only
void prepare_step(Ctx* h);              // ghosts and primitives valid before a step (periodic fill)
void ensure_ghosts(Ctx* h);             // ghosts valid for a read of the state
// one step; for the WENO-BGK scheme it leaves the new state in the other
// buffer and swaps the current-state pointers (swap_state)
void launch_step(Ctx* h, hipEvent_t* ev);  // ev: kEvPerStep events or nullptr
void swap_state(Ctx* h);
// the state, the gas constants or the halo mode changed: ghosts and per-cell primitives are stale
inline void invalidate_derived(Ctx* h) {
    h->derived_valid = false;
    h->prims_valid = false;
}
// persistent group step: the size of one class entry, an entry made from a
// handle's current buffers, and the cooperative launch over ncls entries
// (device table) for nsteps steps; grid_out receives the block count
constexpr int kPersistMaxClasses = 16;
size_t persist_class_bytes(int precision);
void persist_fill_class(Ctx* h, void* dst);
hipError_t launch_persist(const gkm_params& par, const void* table, int ncls, int nsteps, hipStream_t s,
                          int* grid_out);
void launch_split(Ctx* h, int box, int q, const double* dsrc);
void launch_merge(Ctx* h, int box, int q, double* ddst);
void launch_flux_points(hipStream_t s, int prec, int n, const double* rec, double dx, double gam, double invPr, int K,
                        double* out);
void launch_gks_points(hipStream_t s, int prec, int n, const double* rec, double gam, double invPr, int K,
                       double* out);
void launch_weno_points(hipStream_t s, int prec, int recon, int n, const double* rec, double gam, double R,
                        int energy_recon, double* out);

// statistics: number of accumulated quantities of each pass
constexpr int kStatA = 28;  // sums + one max (slot 27)
constexpr int kStatB = 4;
hipError_t launch_stats_a(Ctx* h, int box, double* host_out /*kStatA*/);
hipError_t launch_stats_b(Ctx* h, int box, const double means[4], double* host_out /*kStatB*/);
constexpr int kPStat = 15;  // PrintStats sums
constexpr int kSplRing = 32;  // terms of the backward sums of the x and y line slopes
// spline slopes into h->spl_grad (allocated by the caller), then the sums
hipError_t launch_print_stats_async(Ctx* h, int box, hipStream_t s, double* host_out /*kPStat + kStatB*/);
// doubles of h->spl_part the fused PrintStats passes need for this grid
size_t print_stats_part_doubles(const Geo& g);

}  // namespace gkm
