// noalloc.hip -- mixer_image_noalloc: mixer_image without any device allocation, for callers that
// cannot call hipMalloc (code under HIP graph capture, tight serving loops, harnesses that forbid
// it). mixer_image allocates its int64 grid workspace per call (hipMallocAsync); here the
// workspace is `scratch`, a file-scope device array of fixed size that every GPU the library
// runs on holds until the process ends. Its device address comes from hipGetSymbolAddress, once
// per device (the array's name in host code is not a device address). Every call on a GPU shares
// that GPU's array, so calls that overlap on one GPU race; each call clears what it uses first.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.h"
#include "kernels.h"
#include "mixer.h"

namespace mixer {
namespace {

// Workspace layout: the int64 image grid [C][ng*ng][2] (at most kGridBytes: 256 channels at
// ng = 64, or 1024 channels at ng = 32), then the integrate segment list of the array's block
// count and the FFT twiddles of its ng, both written by the prep kernel of every call.
constexpr size_t kGridBytes = (size_t)16 << 20;
constexpr int kMaxSegs = 32 * 33 / 2;  // >= segments of nb = 32 element blocks (272)
constexpr size_t kSegBytes = (size_t)kMaxSegs * sizeof(int4);
constexpr size_t kTwBytes = 32 * sizeof(float2);  // ng / 2 <= 32 twiddles
constexpr size_t kScratchBytes = kGridBytes + kSegBytes + kTwBytes;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Number of integrate segments of nb element blocks (the list segs_for(nb) builds on the host).
int seg_count(int nb) {
    int n = 0;
    for (int d = 0; d < nb; ++d) n += (nb - d + kSegTiles - 1) / kSegTiles;
    return n;
}

// Clears the grid (grid-stride over all workgroups) and writes the call's tables (workgroup 0):
// the segments in the order segs_for(nb) gives them, and tw[q] = exp(-2 pi i q / ng) computed in
// double precision and rounded once to f32, as the host table of get_twiddles is.
__global__ __launch_bounds__(256) void k_noalloc_prep(unsigned long long* __restrict__ grid,
                                                      size_t n_words, int4* __restrict__ segs, int nb,
                                                      float2* __restrict__ tw, int ng) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n_words; i += (size_t)gridDim.x * 256)
        grid[i] = 0ull;
    if (blockIdx.x != 0) return;
    if (threadIdx.x < nb) {  // thread d writes the segments of tile diagonal d
        const int d = threadIdx.x;
        int k = 0;
        for (int e = 0; e < d; ++e) k += (nb - e + kSegTiles - 1) / kSegTiles;
        for (int j0 = 0; j0 < nb - d; j0 += kSegTiles) {
            const int j1 = j0 + kSegTiles < nb - d ? j0 + kSegTiles : nb - d;
            segs[k++] = make_int4(d, j0, j1, 0);
        }
    }
    for (int q = threadIdx.x; q < ng / 2; q += 256) {
        const double a = 2.0 * M_PI * (double)q / (double)ng;
        tw[q] = make_float2((float)cos(a), (float)-sin(a));
    }
}

std::mutex g_mtx;
constexpr int kMaxDevices = 64;
unsigned char* g_ptr[kMaxDevices] = {};  // device address of `scratch` on each device
bool g_lds[kMaxDevices][7] = {};         // image LDS limit set for ng = 2^k

// The device address of this device's `scratch` (looked up once per device, with it current).
int scratch_ptr(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kMaxDevices) return MIXER_ERR_NODEV;
    std::lock_guard<std::mutex> lk(g_mtx);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_status(e);
        g_ptr[dev] = (unsigned char*)p;
    }
    *out = g_ptr[dev];
    return MIXER_OK;
}

int image_lds_ready(int dev, int ng) {
    int k = 0;
    while ((1 << k) < ng) ++k;
    std::lock_guard<std::mutex> lk(g_mtx);
    if (!g_lds[dev][k]) {
        MIXER_HIP(set_image_lds_limit(ng));
        g_lds[dev][k] = true;
    }
    return MIXER_OK;
}

}  // namespace
}  // namespace mixer

using namespace mixer;

extern "C" int mixer_image_noalloc(void* stream, const mixer_array_cfg* c, const int32_t* d_acc,
                                   float* d_img, float* d_vmax, float* d_vmin) {
    int st = mixer_validate_cfg(c);
    if (st) return st;
    if (!d_acc || !d_img || !d_vmax || !d_vmin) return MIXER_ERR_ARG;
    const size_t gbytes = (size_t)c->n_chan * c->ng * c->ng * 16;
    if (gbytes > kGridBytes) {
        fprintf(stderr, "[mixer] mixer_image_noalloc: grid of %zu bytes (%d channels x %d x %d) exceeds "
                        "the %zu-byte static workspace\n", gbytes, c->n_chan, c->ng, c->ng, kGridBytes);
        return MIXER_ERR_ARG;
    }
    int dev = 0;
    MIXER_HIP(hipGetDevice(&dev));
    unsigned char* base = nullptr;
    MIXER_TRY(scratch_ptr(dev, &base));
    MIXER_TRY(image_lds_ready(dev, c->ng));
    unsigned long long* grid = (unsigned long long*)base;
    int4* segs = (int4*)(base + kGridBytes);
    float2* tw = (float2*)(base + kGridBytes + kSegBytes);
    const int nb = c->n_elem / 32;
    hipStream_t s = (hipStream_t)stream;
    const size_t words = gbytes / 8;
    const int wgs = (int)std::min<size_t>(1024, (words + 255) / 256 + 1);
    hipLaunchKernelGGL(k_noalloc_prep, dim3(wgs), dim3(256), 0, s, grid, words, segs, nb, tw, c->ng);
    MIXER_HIP(hipGetLastError());
    XcorrPlan p{};
    p.n_elem = c->n_elem;
    p.n_slots = mixer_acc_slots(c);  // the accumulator layout mixer_xcorr(_heaps) wrote
    p.n_segs = seg_count(nb);
    p.d_segs = segs;
    MIXER_HIP(launch_integrate(s, p, d_acc, nullptr, nullptr, grid, c->n_elem, c->n_chan, c->grid_x, c->ng));
    return hip_status(launch_image(s, grid, tw, d_img, d_vmax, d_vmin, c->n_chan, c->ng));
}
