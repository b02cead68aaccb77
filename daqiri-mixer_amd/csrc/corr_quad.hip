// corr_quad.hip -- QUAD correlator kernels (256-element arrays straight from the heaps).
#include "corr_quad.h"

namespace mixer {
hipError_t launch_quad_1(hipStream_t s, const XcorrPlan& p, const XSrc& src, int32_t* acc,
                         uint32_t* wrap, int C, int n_tb, int accumulate) {
    return launch_quad_t<1>(s, p, src, acc, wrap, C, n_tb, accumulate);
}
hipError_t launch_quad_3(hipStream_t s, const XcorrPlan& p, const XSrc& src, int32_t* acc,
                         uint32_t* wrap, int C, int n_tb, int accumulate) {
    return launch_quad_t<3>(s, p, src, acc, wrap, C, n_tb, accumulate);
}
}  // namespace mixer
