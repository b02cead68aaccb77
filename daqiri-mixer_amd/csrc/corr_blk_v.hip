// corr_blk_v.hip -- BLOCKED correlator kernels with the integrate fused into the epilogue
// (corr_epi.h), pipeline variant 3.
#include "corr_epi.h"

namespace mixer {
hipError_t launch_xcorr_v(hipStream_t s, const XcorrPlan& p, const XSrc& src, const VOutArgs& vo,
                          uint32_t* wrap, int C, int n_tb) {
    if (p.kind != XcorrPlan::BLOCKED || p.n_slots != 1 || 32 % vo.gs) return hipErrorInvalidValue;
    const int nb = p.n_elem / 32;
    if (nb < 12 || nb > 32 || nb % 4) return hipErrorInvalidValue;
    return launch_blk_v_t<3>(s, p, src, vo, wrap, C, n_tb);
}
}  // namespace mixer
