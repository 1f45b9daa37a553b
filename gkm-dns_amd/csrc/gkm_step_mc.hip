// gkm_step_mc.hip — the FP64 windowed X/Y WENO reconstruction
// (k_recon_win<double>, gkm_step_dev.h) compiled on its own with the
// instruction scheduler set for long memory clauses
// (-mllvm -amdgpu-sched-strategy=max-memory-clause, the Makefile's per-file
// flags): the three FP64 reconstructions of a 256^3 step measured 1.552 ->
// 1.540 ms against the default schedule.  Same kernel source, same bits.
#include "gkm_step_dev.h"

namespace gkm {

template <class Real, class St, int RECON, int ENERGY, int CHUNK>
void launch_recon_win_k(dim3 grid, hipStream_t s, Geo g, Box3 b, unsigned sn, unsigned sa, int n_normal, St st,
                        const Real* Tw, Planes10<Real> wi, const StepPar* pars) {
    hipLaunchKernelGGL((k_recon_win<Real, St, RECON, ENERGY, CHUNK>), grid, dim3(256), 0, s, g, b, sn, sa, n_normal,
                       st, Tw, wi, pars);
}

template void launch_recon_win_k<double, StateD, 0, 0, 4>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateD, const double*, Planes10<double>, const StepPar*);
template void launch_recon_win_k<double, StateD, 0, 0, 8>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateD, const double*, Planes10<double>, const StepPar*);
template void launch_recon_win_k<double, StateD, 0, 0, 16>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateD, const double*, Planes10<double>, const StepPar*);
template void launch_recon_win_k<double, StateD, 0, 1, 4>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateD, const double*, Planes10<double>, const StepPar*);
template void launch_recon_win_k<double, StateD, 0, 1, 8>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateD, const double*, Planes10<double>, const StepPar*);
template void launch_recon_win_k<double, StateD, 0, 1, 16>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateD, const double*, Planes10<double>, const StepPar*);
template void launch_recon_win_k<double, StateD, 1, 0, 4>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateD, const double*, Planes10<double>, const StepPar*);
template void launch_recon_win_k<double, StateD, 1, 0, 8>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateD, const double*, Planes10<double>, const StepPar*);
template void launch_recon_win_k<double, StateD, 1, 0, 16>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateD, const double*, Planes10<double>, const StepPar*);
template void launch_recon_win_k<double, StateD, 1, 1, 4>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateD, const double*, Planes10<double>, const StepPar*);
template void launch_recon_win_k<double, StateD, 1, 1, 8>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateD, const double*, Planes10<double>, const StepPar*);
template void launch_recon_win_k<double, StateD, 1, 1, 16>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateD, const double*, Planes10<double>, const StepPar*);

}  // namespace gkm
