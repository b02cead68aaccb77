// kernels.h -- internal launch interface of the gfx950 kernels (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mixer {

// Tiles per integrate work unit (a run along one tile diagonal).
constexpr int kSegTiles = 2;

// How the correlator covers the lower triangle of one channel.
//   QUAD     arrays of a multiple of 4 blocks (N = 128, 256, 384, ..., 1024) staged straight
//            from the heaps (prefetch 1 or 3): one workgroup per (4 adjacent channels, item of
//            the triangle: a 4-block diagonal triangle or a 2x4-tile rectangle) so that every
//            staged load is a whole 128-byte line of a heap plane.
//   FOLD     arrays of <= 8 blocks of 32 elements: one workgroup per (channel, slot), one wave
//            per folded row pair (r, nb-1-r); balanced (nb+1 tiles per wave).
//   BLOCKED  larger arrays: one workgroup per 4x8-tile super-block, one wave per 2x2 tile
//            block (8 waves), an explicit super-block list (I0, J0) on the device.
// n_slots splits the time axis of one call into partial int32 sums (summed exactly by the
// integrate and image kernels) so that small channel counts still fill all 256 CUs.
struct XcorrPlan {
    enum Kind { FOLD = 0, BLOCKED = 1, QUAD = 2 } kind;
    int n_elem;
    int n_slots;
    int prefetch;         // time-loop pipeline: 1 = registers one block ahead, 2/4 = LDS DMA stages
    int ct_ch;            // corner-turn channels per workgroup: 8, 16 or 32
    int fused;            // contexts: 1 = correlate straight from the heaps, 0 = via tiles
    int mfma;             // MFMA tile shape: 32 (v_mfma_i32_32x32x32_i8) or 16 (16x16x64, FOLD)
    int n_tasks;          // BLOCKED only
    const void* d_tasks;  // BLOCKED only: int2 (I0, J0) per task
    int n_quad;           // QUAD items: int2 (d0, -1) triangle or (r0, c0) rectangle
    const void* d_quad;
    int fuse_integ;       // BLOCKED, one slot, prefetch 3: the integrate may run in the epilogue
    int n_segs;           // integrate work items
    const void* d_segs;   // int4 (d, j0, j1, 0): tiles (J + d, J), j0 <= J < j1
};

// Where the correlator reads its 32-sample element rows. A row (channel c, time block tb,
// plane p, element e) is 32 contiguous bytes both in the corner-turned tiles
// [C][n_tb][2][N][32] and in the raw heaps [n_tb][N][re C*32 | im C*32], so the correlator can
// stage straight from the heaps (the corner-turn becomes part of its loads) or from tiles.
struct XSrc {
    const int8_t* base;  // element 0, plane 0, time block 0 of channel 0
    size_t chan;         // bytes between channels
    size_t tb;           // bytes between time blocks
    int plane;           // bytes between the re and im planes
    int elem;            // bytes between elements
    __device__ __forceinline__ size_t off(int p, int e, int h) const {
        return (size_t)p * plane + (size_t)e * elem + h * 16;
    }
    // the same within one time block (< 2^32 bytes: plane <= 8 KiB, elem <= 16 KiB, e < 1024), so
    // a load is a uniform 64-bit base plus one 32-bit register
    __device__ __forceinline__ uint32_t off32(int p, int e, int h) const {
        return (uint32_t)p * (uint32_t)plane + (uint32_t)e * (uint32_t)elem + (uint32_t)h * 16u;
    }
};
__host__ __device__ inline XSrc xsrc_tiles(const int8_t* tiles, int N, int n_tb) {
    const size_t slab = (size_t)N * 64;
    return XSrc{tiles, (size_t)n_tb * slab, slab, (int)(slab / 2), 32};
}
__host__ __device__ inline XSrc xsrc_heaps(const uint8_t* heaps, int N, int C) {
    const size_t payload = (size_t)C * 64;
    return XSrc{(const int8_t*)heaps, 32, (size_t)N * payload, C * 32, (int)payload};
}

// Outputs of the integrate when it is fused into the BLOCKED correlator's epilogue (one launch and
// one time slot cover the integration): what k_integrate writes from the accumulator.
struct VOutArgs {
    float* v_store;         // V [C][2][N][N] f32, overwritten (may be null)
    float* v_add;           // V [C][2][N][N] f32, added to (the cumulative cube; may be null)
    unsigned long long* G;  // image grid [C][ng * ng][2] int64 (may be null: no gridding)
    int N, gs, ng;
};

hipError_t launch_corner_turn(hipStream_t s, const uint8_t* heaps, int8_t* tiles, uint32_t* wrap,
                              int N, int C, int n_tb, int c_base, int c_count, int cch);
// wrap (optional): counts workgroups that staged an imaginary -128
hipError_t launch_xcorr(hipStream_t s, const XcorrPlan& plan, const XSrc& src, int32_t* acc,
                        uint32_t* wrap, int C, int n_tb, int accumulate);
// BLOCKED plans of one time slot: correlate the whole integration and write V (and grid it) from
// the epilogue, without the int32 accumulator (prefetch 3)
hipError_t launch_xcorr_v(hipStream_t s, const XcorrPlan& plan, const XSrc& src, const VOutArgs& vo,
                          uint32_t* wrap, int C, int n_tb);
// integrate (+ grid into G when non-null) the plan's accumulator slots, one workgroup per
// tile-diagonal segment of the plan and channel
hipError_t launch_integrate(hipStream_t s, const XcorrPlan& p, const int32_t* acc, float* v_store,
                            float* v_add, unsigned long long* G, int N, int C, int grid_x, int ng);
size_t image_lds_bytes(int ng);
hipError_t set_image_lds_limit(int ng);
// image of the int64 grid G (cleared); Gf (optional) accumulates G for a frame image
hipError_t launch_image(hipStream_t s, unsigned long long* G, const float2* tw, float* img,
                        float* vmax, float* vmin, int C, int ng, unsigned long long* Gf = nullptr);
hipError_t launch_xcorr_plain(hipStream_t s, const int8_t* tiles, int32_t* vre, int32_t* vim, int N,
                              int n_tb, int c);

hipError_t launch_fill_random(hipStream_t s, uint8_t* p, size_t bytes, uint64_t seed);

// One array ("job") of a batch, as the batched kernels see it (device memory).
struct BatchJobDev {
    const uint8_t* heaps;        // [n_tb][n_elem][payload]
    int8_t* tiles;               // [n_chan][n_tb][2][n_elem][32]
    int32_t* acc;                // [n_slots][n_chan][tiles][2][32][32]
    float* v;                    // [n_chan][2][n_elem][n_elem]
    unsigned long long* G;       // [n_chan][ng*ng][2]
    float* img;                  // [n_chan][ng][ng], then vmax[n_chan], vmin[n_chan]
    const float2* tw;            // FFT twiddles for ng
    const int2* sblocks;         // BLOCKED super-blocks (I0, J0)
    const int4* segs;            // integrate tile-diagonal segments
    int n_segs;
    uint32_t* wrap;              // corner-turn -128 counter
    int n_elem, n_chan, n_tb, n_slots, grid_x, ng, nb, n_sb;
    int ct_wg0, int_wg0, img_wg0;  // first workgroup of this job in each batched launch
    int ch_per_item;               // FOLD: channels one correlator work item covers (>= 1)
    int fused;                     // 1: correlate straight from the heaps, 0: from the tiles
};

// Correlator work item: (job, first channel, super-block index or 0, time slot). The items are
// split into one queue per XCD: queue x is items [start[x], start[x] + size[x]).
struct BatchQueues {
    int start[8];
    int size[8];
};
hipError_t launch_batch_reset(hipStream_t s, int* counters);
hipError_t launch_batch_corner_turn(hipStream_t s, const BatchJobDev* jobs, int n_jobs, int n_wgs);
// fold items (arrays of <= 8 element blocks) / BLOCKED items (more); counters: 8 ints each
// mfma: 32 = v_mfma_i32_32x32x32_i8 tiles, 16 = v_mfma_i32_16x16x64_i8 (fold items)
hipError_t launch_batch_xcorr(hipStream_t s, const BatchJobDev* jobs, const int4* items,
                              const BatchQueues& q, int* counters, int n_wgs, int mfma);
hipError_t launch_batch_xcorr_blk(hipStream_t s, const BatchJobDev* jobs, const int4* items,
                                  const BatchQueues& q, int* counters, int n_wgs);
hipError_t launch_batch_integrate(hipStream_t s, const BatchJobDev* jobs, int n_jobs, int n_wgs,
                                  int max_ng);
hipError_t launch_batch_image(hipStream_t s, const BatchJobDev* jobs, int n_jobs, int n_wgs,
                              int max_ng);
hipError_t set_batch_lds_limits(int max_ng);
// channels packed into one correlator workgroup for an array of nb blocks (FOLD with <= 5 blocks)
int batch_fold_groups(int nb);

}  // namespace mixer
