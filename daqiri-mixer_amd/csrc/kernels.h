// kernels.h -- internal launch interface of the gfx950 kernels (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mixer {

// How the correlator covers the lower triangle of one channel.
//   FOLD     arrays of <= 8 blocks of 32 elements: one workgroup per (channel, slot), one wave
//            per folded row pair (r, nb-1-r); balanced (nb+1 tiles per wave).
//   BLOCKED  larger arrays: one wave per 2x4 tile block, four per workgroup, an explicit task
//            list (I0, J0) on the device.
// n_slots splits the time axis of one call into partial int32 sums (summed exactly by the
// integrate and image kernels) so that small channel counts still fill all 256 CUs.
struct XcorrPlan {
    enum Kind { FOLD = 0, BLOCKED = 1 } kind;
    int n_elem;
    int n_slots;
    int n_tasks;          // BLOCKED only
    const void* d_tasks;  // BLOCKED only: int2 (I0, J0) per task
};

hipError_t launch_corner_turn(hipStream_t s, const uint8_t* heaps, int8_t* tiles, uint32_t* wrap,
                              int N, int C, int n_tb, int c_base, int c_count);
hipError_t launch_xcorr(hipStream_t s, const XcorrPlan& plan, const int8_t* tiles, int32_t* acc,
                        int C, int n_tb, int accumulate);
hipError_t launch_integrate(hipStream_t s, const int32_t* acc, float* v_store, float* v_add, int N,
                            int C, int n_slots);
size_t image_lds_bytes(int ng);
hipError_t set_image_lds_limit(int ng);
hipError_t launch_image(hipStream_t s, const int32_t* acc, const float2* tw, float* img,
                        float* vmax, float* vmin, int N, int C, int n_slots, int grid_x, int ng);
hipError_t launch_xcorr_plain(hipStream_t s, const int8_t* tiles, int32_t* vre, int32_t* vim, int N,
                              int n_tb, int c);

}  // namespace mixer
