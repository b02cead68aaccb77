// kernels_corr.hip -- dispatch of the correlator launch plan to the kernel variants
// (corr_impl.h; instantiated per pipeline variant in corr_{fold,blk}_pf{1,2,3,4,5}.hip and, for the
// 16x16x64 tile shape, corr_fold_m16.hip).
#include "kernels.h"

namespace mixer {

#define MIXER_DECL(kind, pf)                                                                  \
    hipError_t launch_##kind##_##pf(hipStream_t s, const XcorrPlan& p, const XSrc& src,      \
                                    int32_t* acc, uint32_t* wrap, int C, int n_tb,           \
                                    int accumulate);
MIXER_DECL(fold, 1)
MIXER_DECL(fold, 2)
MIXER_DECL(fold, 3)
MIXER_DECL(fold, 4)
MIXER_DECL(fold, 5)
MIXER_DECL(blk, 1)
MIXER_DECL(blk, 2)
MIXER_DECL(blk, 3)
MIXER_DECL(blk, 4)
MIXER_DECL(blk, 5)
#undef MIXER_DECL

hipError_t launch_fold_m16(hipStream_t s, const XcorrPlan& p, const XSrc& src, int32_t* acc,
                           uint32_t* wrap, int C, int n_tb, int accumulate);
hipError_t launch_quad_1(hipStream_t s, const XcorrPlan& p, const XSrc& src, int32_t* acc,
                         uint32_t* wrap, int C, int n_tb, int accumulate);
hipError_t launch_quad_3(hipStream_t s, const XcorrPlan& p, const XSrc& src, int32_t* acc,
                         uint32_t* wrap, int C, int n_tb, int accumulate);

hipError_t launch_xcorr(hipStream_t s, const XcorrPlan& p, const XSrc& src, int32_t* acc,
                        uint32_t* wrap, int C, int n_tb, int accumulate) {
    if (p.kind == XcorrPlan::QUAD) {
        if (p.prefetch == 3) return launch_quad_3(s, p, src, acc, wrap, C, n_tb, accumulate);
        if (p.prefetch == 1) return launch_quad_1(s, p, src, acc, wrap, C, n_tb, accumulate);
        return hipErrorInvalidValue;
    }
    const bool fold = p.kind == XcorrPlan::FOLD;
    if (fold && p.mfma == 16) return launch_fold_m16(s, p, src, acc, wrap, C, n_tb, accumulate);
    switch (p.prefetch) {
    case 1: return fold ? launch_fold_1(s, p, src, acc, wrap, C, n_tb, accumulate)
                        : launch_blk_1(s, p, src, acc, wrap, C, n_tb, accumulate);
    case 2: return fold ? launch_fold_2(s, p, src, acc, wrap, C, n_tb, accumulate)
                        : launch_blk_2(s, p, src, acc, wrap, C, n_tb, accumulate);
    case 3: return fold ? launch_fold_3(s, p, src, acc, wrap, C, n_tb, accumulate)
                        : launch_blk_3(s, p, src, acc, wrap, C, n_tb, accumulate);
    case 4: return fold ? launch_fold_4(s, p, src, acc, wrap, C, n_tb, accumulate)
                        : launch_blk_4(s, p, src, acc, wrap, C, n_tb, accumulate);
    case 5: return fold ? launch_fold_5(s, p, src, acc, wrap, C, n_tb, accumulate)
                        : launch_blk_5(s, p, src, acc, wrap, C, n_tb, accumulate);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace mixer
