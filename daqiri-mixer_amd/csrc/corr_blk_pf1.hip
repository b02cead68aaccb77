// corr_blk_pf1.hip -- BLOCKED correlator kernels with 1 prefetch stage.
#include "corr_impl.h"

namespace mixer {
hipError_t launch_blk_1(hipStream_t s, const XcorrPlan& p, const int8_t* tiles, int32_t* acc,
                          int C, int n_tb, int accumulate) {
    return launch_blk_pf<1>(s, p, tiles, acc, C, n_tb, accumulate);
}
}  // namespace mixer
