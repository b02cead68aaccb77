// corr_fold_pf5.hip -- FOLD correlator kernels with pipeline variant 5.
#include "corr_impl.h"

namespace mixer {
hipError_t launch_fold_5(hipStream_t s, const XcorrPlan& p, const XSrc& src, int32_t* acc,
                          uint32_t* wrap, int C, int n_tb, int accumulate) {
    return launch_fold_pf<5>(s, p, src, acc, wrap, C, n_tb, accumulate);
}
}  // namespace mixer
