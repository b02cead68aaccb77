// kernels.hip -- the X-engine hot path as hand-written CDNA4 (gfx950) kernels.
//
//   k_corner_turn   heaps [tb][e][re C*32 | im C*32]  ->  tiles [c][tb][re|im][e][32]
//                   (replaces k_packAB, src/rx_host_corr.cu:56-72)
//   (the correlator k_xcorr_fold / k_xcorr_blk is in kernels_corr.hip)
//   k_integrate     int32 slots -> f32 full-square cube (replaces k_flush_cube,
//                   src/rx_host_corr.cu:86-95)
//   k_image         exact integer gridding + ng x ng FP32 Stockham FFT in LDS + fftshift/Re/
//                   min/max (replaces k_diff/k_grid_cube/cuFFT/host publish,
//                   src/rx_host_corr.cu:74-85,205-206,219-235)
//   k_xcorr_plain   one 32x32 full-square tile per wave, reference math, per-channel launch
//                   (the baseline form)
//
#include "device_common.h"
#include "kernels.h"
#include "kernels_body.h"

namespace mixer {

// ------------------------------------------------------------------------------------------
// Single-array launches of the bodies in kernels_body.h
// ------------------------------------------------------------------------------------------
template <int CCH>
__global__ __launch_bounds__(256) void k_corner_turn(const uint8_t* __restrict__ heaps,
                                                     int8_t* __restrict__ tiles,
                                                     uint32_t* __restrict__ wrap, int N, int C,
                                                     int n_tb, int c_base, int c_end) {
    corner_turn_body<CCH>(heaps, tiles, wrap, N, C, n_tb, blockIdx.x, blockIdx.y * 32,
                          c_base + blockIdx.z * CCH, c_end);
}

__global__ __launch_bounds__(256) void k_integrate(const int32_t* __restrict__ acc,
                                                   float* __restrict__ v_store,
                                                   float* __restrict__ v_add,
                                                   unsigned long long* __restrict__ G, int N,
                                                   int C, int n_slots, int grid_x, int ng,
                                                   const int4* __restrict__ segs) {
    integrate_seg(acc, v_store, v_add, G, N, C, n_slots, grid_x, ng, segs[blockIdx.x], blockIdx.y);
}

__global__ __launch_bounds__(256) void k_image(unsigned long long* __restrict__ G,
                                               unsigned long long* __restrict__ Gf,
                                               const float2* __restrict__ tw,
                                               float* __restrict__ img, float* __restrict__ vmax,
                                               float* __restrict__ vmin, int ng) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    image_body(G, Gf, tw, img, vmax, vmin, ng, blockIdx.x, smem);
}

// ------------------------------------------------------------------------------------------
// Plain form: one wave per 32x32 output tile of the FULL square of one channel, four MFMAs
// per 32-sample step exactly as cgemm.cuh:59-62, written to V[N][N] (overwrite).
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_xcorr_plain(const int8_t* __restrict__ tiles,
                                                     int32_t* __restrict__ vre,
                                                     int32_t* __restrict__ vim, int N, int n_tb,
                                                     int c) {
    const int nb = N / 32;
    const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile >= nb * nb) return;
    const int I = tile / nb, J = tile % nb;
    const int lane = threadIdx.x & 63, lrow = lane & 31, kg = lane >> 5;
    v16i cr = (v16i){0}, ci = (v16i){0};
    const size_t plane = (size_t)N * 32;
    const int8_t* base = tiles + (size_t)c * n_tb * 2 * plane + lrow * 32 + kg * 16;
    for (int kb = 0; kb < n_tb; ++kb, base += 2 * plane) {
        const v4i ar = *(const v4i*)(base + I * 1024);
        const v4i ai = *(const v4i*)(base + plane + I * 1024);
        const v4i br = *(const v4i*)(base + J * 1024);
        const v4i bi = neg8x16(*(const v4i*)(base + plane + J * 1024));
        cr = mfma(ar, br, cr);
        cr = mfma(neg8x16(ai), bi, cr);
        ci = mfma(ar, bi, ci);
        ci = mfma(ai, br, ci);
    }
    const size_t o = (size_t)c * N * N;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = I * 32 + 8 * (r >> 2) + 4 * kg + (r & 3), col = J * 32 + lrow;
        vre[o + (size_t)row * N + col] = cr[r];
        vim[o + (size_t)row * N + col] = ci[r];
    }
}

// ------------------------------------------------------------------------------------------
// Seeded random int8 heaps in [-127, 127] (autotuning and large benchmark inputs): a
// splitmix64 hash of (seed, 8-byte word index), one byte per 8 bits reduced mod 255.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fill_random(uint8_t* __restrict__ p, size_t n_words,
                                                     uint64_t seed) {
    for (size_t w = (size_t)blockIdx.x * 256 + threadIdx.x; w < n_words; w += (size_t)gridDim.x * 256) {
        uint64_t z = seed + (w + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        uint64_t out = 0;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int v = (int)((z >> (8 * b)) & 0xFF) % 255 - 127;
            out |= (uint64_t)(uint8_t)(int8_t)v << (8 * b);
        }
        ((uint64_t*)p)[w] = out;
    }
}

// ------------------------------------------------------------------------------------------
// Host-side launchers
// ------------------------------------------------------------------------------------------
hipError_t launch_corner_turn(hipStream_t s, const uint8_t* heaps, int8_t* tiles, uint32_t* wrap,
                              int N, int C, int n_tb, int c_base, int c_count, int cch) {
    switch (cch) {
    case 8:
        hipLaunchKernelGGL(k_corner_turn<8>, dim3(n_tb, N / 32, (c_count + 7) / 8), dim3(256), 0, s,
                           heaps, tiles, wrap, N, C, n_tb, c_base, c_base + c_count);
        break;
    case 16:
        hipLaunchKernelGGL(k_corner_turn<16>, dim3(n_tb, N / 32, (c_count + 15) / 16), dim3(256), 0,
                           s, heaps, tiles, wrap, N, C, n_tb, c_base, c_base + c_count);
        break;
    case 32:
        hipLaunchKernelGGL(k_corner_turn<32>, dim3(n_tb, N / 32, (c_count + 31) / 32), dim3(256), 0,
                           s, heaps, tiles, wrap, N, C, n_tb, c_base, c_base + c_count);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_integrate(hipStream_t s, const XcorrPlan& p, const int32_t* acc, float* v_store,
                            float* v_add, unsigned long long* G, int N, int C, int grid_x, int ng) {
    hipLaunchKernelGGL(k_integrate, dim3(p.n_segs, C), dim3(256), 0, s, acc, v_store, v_add, G, N,
                       C, p.n_slots, grid_x, ng, (const int4*)p.d_segs);
    return hipGetLastError();
}

size_t image_lds_bytes(int ng) { return (size_t)ng * ng * 16; }

hipError_t launch_image(hipStream_t s, unsigned long long* G, const float2* tw, float* img,
                        float* vmax, float* vmin, int C, int ng, unsigned long long* Gf) {
    hipLaunchKernelGGL(k_image, dim3(C), dim3(256), image_lds_bytes(ng), s, G, Gf, tw, img, vmax,
                       vmin, ng);
    return hipGetLastError();
}

hipError_t set_image_lds_limit(int ng) {
    return hipFuncSetAttribute((const void*)k_image, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)image_lds_bytes(ng));
}

hipError_t launch_xcorr_plain(hipStream_t s, const int8_t* tiles, int32_t* vre, int32_t* vim, int N,
                              int n_tb, int c) {
    const int nb = N / 32;
    hipLaunchKernelGGL(k_xcorr_plain, dim3((nb * nb + 3) / 4), dim3(256), 0, s, tiles, vre, vim, N,
                       n_tb, c);
    return hipGetLastError();
}

hipError_t launch_fill_random(hipStream_t s, uint8_t* p, size_t bytes, uint64_t seed) {
    if (bytes % 8) return hipErrorInvalidValue;
    const size_t words = bytes / 8;
    size_t blocks = (words + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_random, dim3((unsigned)blocks), dim3(256), 0, s, p, words, seed);
    return hipGetLastError();
}

}  // namespace mixer
