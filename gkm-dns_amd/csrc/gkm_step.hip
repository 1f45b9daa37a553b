// gkm_step.hip — launchers of the kernels of one GKM-DNS time step on gfx950
// (gkm_step_dev.h): periodic ghost fill fused with the per-cell primitives,
// WENO reconstruction per sweep, BGK flux per sweep, conservative update,
// state split/merge for the mixed form; the step, graph and persistent-kernel
// plumbing.
#include "gkm_step_dev.h"

namespace gkm {

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
    launch_recon_win_k<Real, St, RECON, ENERGY, CHUNK>(dim3(blocks_for(b.count(), 256), h->nbox), h->stream, h->geo,
                                                        b, sn, sa, n_normal, st, Tw, wi, h->d_pars);
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
