// corr_fold_m16.hip -- FOLD correlator kernels in the 16x16x64 MFMA tile shape.
#include "corr_impl.h"

namespace mixer {
hipError_t launch_fold_m16(hipStream_t s, const XcorrPlan& p, const XSrc& src, int32_t* acc,
                           uint32_t* wrap, int C, int n_tb, int accumulate) {
    return launch_fold_pf<6>(s, p, src, acc, wrap, C, n_tb, accumulate);
}
}  // namespace mixer
