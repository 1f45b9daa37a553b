// gkm_step_ilp.hip — the FP32 windowed X/Y WENO reconstruction
// (k_recon_win<float>, gkm_step_dev.h) compiled on its own with the
// instruction scheduler set for instruction-level parallelism
// (-mllvm -amdgpu-sched-strategy=max-ilp, the Makefile's per-file flags).
// That schedule keeps more of a window's loads in flight (134 VGPRs at 3 waves
// per SIMD instead of 66 at 7) and measured 290 -> 271 us per sweep at 256^3;
// the FP64 windows and every other kernel were slower under it, so they stay
// in the other translation units.  Same kernel source, same bits.
#include "gkm_step_dev.h"

namespace gkm {

template <class Real, class St, int RECON, int ENERGY, int CHUNK>
void launch_recon_win_k(dim3 grid, hipStream_t s, Geo g, Box3 b, unsigned sn, unsigned sa, int n_normal, St st,
                        const Real* Tw, Planes10<Real> wi, const StepPar* pars) {
    hipLaunchKernelGGL((k_recon_win<Real, St, RECON, ENERGY, CHUNK>), grid, dim3(256), 0, s, g, b, sn, sa, n_normal,
                       st, Tw, wi, pars);
}

template void launch_recon_win_k<float, StateM, 0, 0, 4>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateM, const float*, Planes10<float>, const StepPar*);
template void launch_recon_win_k<float, StateM, 0, 0, 8>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateM, const float*, Planes10<float>, const StepPar*);
template void launch_recon_win_k<float, StateM, 0, 0, 16>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateM, const float*, Planes10<float>, const StepPar*);
template void launch_recon_win_k<float, StateM, 0, 1, 4>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateM, const float*, Planes10<float>, const StepPar*);
template void launch_recon_win_k<float, StateM, 0, 1, 8>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateM, const float*, Planes10<float>, const StepPar*);
template void launch_recon_win_k<float, StateM, 0, 1, 16>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateM, const float*, Planes10<float>, const StepPar*);
template void launch_recon_win_k<float, StateM, 1, 0, 4>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateM, const float*, Planes10<float>, const StepPar*);
template void launch_recon_win_k<float, StateM, 1, 0, 8>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateM, const float*, Planes10<float>, const StepPar*);
template void launch_recon_win_k<float, StateM, 1, 0, 16>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateM, const float*, Planes10<float>, const StepPar*);
template void launch_recon_win_k<float, StateM, 1, 1, 4>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateM, const float*, Planes10<float>, const StepPar*);
template void launch_recon_win_k<float, StateM, 1, 1, 8>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateM, const float*, Planes10<float>, const StepPar*);
template void launch_recon_win_k<float, StateM, 1, 1, 16>(dim3, hipStream_t, Geo, Box3, unsigned, unsigned, int,
                                                          StateM, const float*, Planes10<float>, const StepPar*);

}  // namespace gkm
