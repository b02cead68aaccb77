// tune.cpp -- the launch autotuner (mixer_autotune): times the per-integration pipeline for
// each candidate launch shape on seeded random device data and keeps the fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <vector>

#include "internal.h"
#include "kernels.h"
#include "mixer.h"

using namespace mixer;

namespace {

struct Bufs {
    uint8_t* heaps = nullptr;
    int8_t* tiles = nullptr;
    int32_t* acc = nullptr;
    float* v = nullptr;
    float* img = nullptr;
    unsigned long long* G = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Bufs() {
        if (s) (void)hipStreamSynchronize(s);
        (void)hipFree(heaps); (void)hipFree(tiles); (void)hipFree(acc); (void)hipFree(v); (void)hipFree(img); (void)hipFree(G);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
    }
};

// One pass of (corner-turn +) correlate + integrate + image (or the corner-turn alone), as a
// context runs it under the tuning in effect.
int run_once(const mixer_array_cfg& c, Bufs& b, int n_tb, bool ct_only) {
    XcorrPlan p;
    MIXER_TRY(get_plan(c, &p));
    if (!p.fused || ct_only)
        MIXER_HIP(launch_corner_turn(b.s, b.heaps, b.tiles, nullptr, c.n_elem, c.n_chan, n_tb, 0,
                                     c.n_chan, p.ct_ch));
    if (ct_only) return MIXER_OK;
    const XSrc xs = p.fused ? xsrc_heaps(b.heaps, c.n_elem, c.n_chan) : xsrc_tiles(b.tiles, c.n_elem, n_tb);
    const float2* tw = nullptr;
    MIXER_TRY(get_twiddles(c.ng, &tw));
    if (p.fuse_integ) {
        const VOutArgs vo{b.v, nullptr, b.G, c.n_elem, c.grid_x, c.ng};
        MIXER_HIP(launch_xcorr_v(b.s, p, xs, vo, nullptr, c.n_chan, n_tb));
    } else {
        MIXER_HIP(launch_xcorr(b.s, p, xs, b.acc, nullptr, c.n_chan, n_tb, 0));
        MIXER_HIP(launch_integrate(b.s, p, b.acc, b.v, nullptr, b.G, c.n_elem, c.n_chan, c.grid_x, c.ng));
    }
    const size_t nimg = (size_t)c.n_chan * c.ng * c.ng;
    MIXER_HIP(launch_image(b.s, b.G, tw, b.img, b.img + nimg, b.img + nimg + c.n_chan, c.n_chan, c.ng));
    return MIXER_OK;
}

int time_it(const mixer_array_cfg& c, Bufs& b, int n_tb, bool ct_only, float* ms) {
    MIXER_TRY(run_once(c, b, n_tb, ct_only));  // warm-up (and first-use setup)
    const int reps = 3;
    MIXER_HIP(hipEventRecord(b.e0, b.s));
    for (int r = 0; r < reps; ++r) MIXER_TRY(run_once(c, b, n_tb, ct_only));
    MIXER_HIP(hipEventRecord(b.e1, b.s));
    MIXER_HIP(hipEventSynchronize(b.e1));
    MIXER_HIP(hipEventElapsedTime(ms, b.e0, b.e1));
    *ms /= reps;
    return MIXER_OK;
}

}  // namespace

extern "C" int mixer_autotune(const mixer_array_cfg* cfg, int device, mixer_tuning* best) {
    MIXER_TRY(mixer_validate_cfg(cfg));
    if (!best) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(device));
    mixer_array_cfg c = *cfg;
    // time on at most 4096 samples: the ranking of launch shapes is set by the per-step work
    c.k_update = std::min(c.k_update, 4096);
    c.k_integ = c.k_update;
    c.flags = 0;
    const int n_tb = c.k_update / c.t_pkt;
    const size_t heap_bytes = (size_t)n_tb * c.n_elem * mixer_payload_bytes(&c);
    const size_t nb = c.n_elem / 32;
    Bufs b;
    MIXER_HIP(hipStreamCreateWithFlags(&b.s, hipStreamNonBlocking));
    MIXER_HIP(hipEventCreate(&b.e0));
    MIXER_HIP(hipEventCreate(&b.e1));
    MIXER_HIP(hipMalloc(&b.heaps, heap_bytes));
    MIXER_HIP(hipMalloc(&b.tiles, mixer_tiles_bytes(&c, n_tb)));
    MIXER_HIP(hipMalloc(&b.acc, (size_t)8 * c.n_chan * (nb * (nb + 1) / 2) * 2048 * sizeof(int32_t)));
    MIXER_HIP(hipMalloc(&b.v, (size_t)c.n_chan * 2 * c.n_elem * c.n_elem * sizeof(float)));
    MIXER_HIP(hipMalloc(&b.img, ((size_t)c.n_chan * c.ng * c.ng + 2 * c.n_chan) * sizeof(float)));
    MIXER_HIP(hipMalloc(&b.G, (size_t)c.n_chan * c.ng * c.ng * 16));
    MIXER_HIP(hipMemset(b.G, 0, (size_t)c.n_chan * c.ng * c.ng * 16));
    MIXER_HIP(launch_fill_random(b.s, b.heaps, heap_bytes, 0x4D49585455ull));

    const mixer_tuning saved = tuning_for(c.n_elem);
    const bool had = has_tuning(c.n_elem);
    mixer_tuning t{0, 0, 32, 0, 32, 0}, win{0, 0, 32, 0, 32, 0};
    float best_ms = 1e30f;
    int st = MIXER_OK;
    // the 32x32x32 shape with each pipeline variant, then (arrays of <= 8 element blocks) the
    // 16x16x64 shape, whose two-block stages have a pipeline of their own
    const int n_pf = nb <= 8 ? 6 : 5;
    for (int staging : {1, 2}) {
        for (int slots : {1, 2, 4, 8}) {
            if (slots > n_tb) continue;
            for (int k = 0; k < n_pf; ++k) {
                t.slots = slots;
                t.prefetch = k < 5 ? k + 1 : 0;
                t.mfma = k < 5 ? 32 : 16;
                t.staging = staging;
                if ((st = mixer_set_tuning(c.n_elem, &t))) break;
                float ms = 0;
                if ((st = time_it(c, b, n_tb, false, &ms))) break;
                if (ms < best_ms) { best_ms = ms; win = t; }
            }
            if (st) break;
        }
        if (st) break;
    }
    if (!st && win.staging == 2) {  // the corner-turn width matters only with a separate pass
        best_ms = 1e30f;
        for (int ct : {8, 16, 32}) {
            t = win;
            t.ct_chan = ct;
            if ((st = mixer_set_tuning(c.n_elem, &t))) break;
            float ms = 0;
            if ((st = time_it(c, b, n_tb, true, &ms))) break;
            if (ms < best_ms) { best_ms = ms; win.ct_chan = ct; }
        }
    }
    if (!st) {  // a blocked winner with one slot: integrate fused into the epilogue or separate
        XcorrPlan wp;
        if (!(st = mixer_set_tuning(c.n_elem, &win)) && !(st = get_plan(c, &wp)) && wp.fuse_integ) {
            float fused_ms = 0, sep_ms = 0;
            t = win;
            t.integrate = 1;
            if (!(st = time_it(c, b, n_tb, false, &fused_ms)) && !(st = mixer_set_tuning(c.n_elem, &t)) &&
                !(st = time_it(c, b, n_tb, false, &sep_ms)))
                win.integrate = sep_ms < fused_ms ? 1 : 2;
        }
    }
    if (st) {  // leave the table as it was
        const mixer_tuning automatic{0, 0, 0, 0, 0, 0};
        (void)mixer_set_tuning(c.n_elem, had ? &saved : &automatic);
        return st;
    }
    MIXER_TRY(mixer_set_tuning(c.n_elem, &win));
    *best = win;
    return MIXER_OK;
}
