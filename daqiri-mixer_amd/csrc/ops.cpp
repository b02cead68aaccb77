// ops.cpp -- configuration, launch plans, the device-pointer operator entry points and the HIP
// plumbing of the C ABI (include/mixer.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "internal.h"
#include "kernels.h"
#include "mixer.h"

namespace mixer {

int hip_status(hipError_t e) {
    if (e == hipSuccess) return MIXER_OK;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return MIXER_ERR_NOMEM;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return MIXER_ERR_NODEV;
    fprintf(stderr, "[mixer] HIP error %d: %s\n", (int)e, hipGetErrorString(e));
    return MIXER_ERR_HIP;
}

namespace {
std::mutex g_tune_mtx;
std::map<int, mixer_tuning> g_tuning;  // n_elem -> tuning (0 fields = automatic)
}  // namespace

mixer_tuning tuning_for(int n_elem) {
    std::lock_guard<std::mutex> lk(g_tune_mtx);
    auto it = g_tuning.find(n_elem);
    if (it == g_tuning.end()) return mixer_tuning{0, 0, 0, 0, 0, 0};
    return it->second;
}

bool has_tuning(int n_elem) {
    std::lock_guard<std::mutex> lk(g_tune_mtx);
    return g_tuning.count(n_elem) != 0;
}

// Time slots: split each channel's time axis so the correlator grid reaches ~2 workgroups per
// CU-equivalent of work (256 CUs). Depends on the geometry only, so every call on the same cfg
// (correlate, integrate, image) agrees on the accumulator layout.
int plan_slots(const mixer_array_cfg& c) {
    const int nb = c.n_elem / 32;
    const int n_tb = c.k_update / c.t_pkt;
    const mixer_tuning t = tuning_for(c.n_elem);
    if (t.slots > 0) return t.slots < n_tb ? t.slots : n_tb;
    long units;  // workgroups per slot
    const int kind = plan_kind(c);
    if (kind == XcorrPlan::QUAD) {
        units = (long)quad_items_for(nb).size() * ((c.n_chan + 3) / 4);
    } else if (nb <= 8) {
        units = c.n_chan;
        // small folds run several workgroups per CU
        if (nb <= 5) units = (units + 3) / 4;
    } else {
        units = (long)sblocks_for(nb).size() * c.n_chan;
    }
    int slots = (int)((256 + units - 1) / units);
    if (slots < 1) slots = 1;
    if (slots > 8) slots = 8;
    if (slots > n_tb) slots = n_tb;
    return slots;
}

bool plan_fused(const mixer_array_cfg& c) {
    const mixer_tuning t = tuning_for(c.n_elem);
    return t.staging != 2;
}

int plan_kind(const mixer_array_cfg& c) {
    const int nb = c.n_elem / 32;
    const mixer_tuning t = tuning_for(c.n_elem);
    const int pf = t.prefetch > 0 ? t.prefetch : 3;
    // straight from the heaps with register staging, arrays of 4 or 8 blocks use 4-channel QUAD
    // items (256 x 128 x 8192: 0.216 vs 0.25 ms for the per-channel fold). Larger arrays stay
    // on the per-channel BLOCKED layout: their QUAD rectangles move twice the bytes through L2
    // per tile and measured slower (1024 x 256 x 4096: 2.91 vs 2.80 ms). The direct-to-LDS
    // variants keep the per-channel layouts.
    if (t.mfma == 16 && nb <= 8) return XcorrPlan::FOLD;  // the 16x16x64 shape runs on FOLD
    if (plan_fused(c) && (nb == 4 || nb == 8) && (pf == 1 || pf == 3)) return XcorrPlan::QUAD;
    return nb <= 8 ? XcorrPlan::FOLD : XcorrPlan::BLOCKED;
}

// QUAD items (nb % 4 == 0): the 4-block diagonal triangles (d0 = 4k, -1) and below them the
// rectangles (r0, c0): rows r0, r0 + 1 (r0 = 4k, 4k + 2) x columns c0..c0+3 (c0 = 4j, j < k).
std::vector<std::pair<int, int>> quad_items_for(int nb) {
    std::vector<std::pair<int, int>> v;
    for (int k = 0; k < nb / 4; ++k) v.push_back({4 * k, -1});
    for (int k = 1; k < nb / 4; ++k)
        for (int r0 = 4 * k; r0 < 4 * k + 4; r0 += 2)
            for (int j = 0; j < k; ++j) v.push_back({r0, 4 * j});
    return v;
}

// Integrate work units: runs of at most kSegTiles tiles (J + d, J) along tile diagonal d (d
// ascending, then j0 ascending). A workgroup loads its tiles one after another, so short runs
// keep the longest chain short (256 x 128: 43 -> 35 us going from 8 to 2 tiles; 1024 x 256:
// 660 -> 632 us).
std::vector<int4> segs_for(int nb) {
    std::vector<int4> v;
    for (int d = 0; d < nb; ++d)
        for (int j0 = 0; j0 < nb - d; j0 += kSegTiles)
            v.push_back(make_int4(d, j0, std::min(j0 + kSegTiles, nb - d), 0));
    return v;
}

// Work items (I0, J0) of the BLOCKED correlator (nb % 4 == 0) covering the lower triangle
// exactly once: the diagonal triangles of 8 element blocks (I0 = 8k, J0 = -1; 36 tiles on the
// 8-wave FOLD layout), the full 4x8-tile super-blocks below them (I0 % 4 == 0, J0 % 8 == 0,
// J0 + 8 <= 8 * (I0 / 8)), and for nb % 8 == 4 the last 4-block triangle as a super-block
// (I0 = J0 = nb - 4, whose waves past the diagonal idle).
std::vector<std::pair<int, int>> sblocks_for(int nb) {
    std::vector<std::pair<int, int>> v;
    const int q = nb / 8;
    for (int d = 0; d < q; ++d) v.push_back({8 * d, -1});
    for (int I0 = 8; I0 < nb; I0 += 4)
        for (int J0 = 0; J0 + 8 <= (I0 / 8) * 8; J0 += 8) v.push_back({I0, J0});
    if (nb % 8) v.push_back({8 * q, 8 * q});
    return v;
}

namespace {
std::mutex g_cache_mtx;
std::map<std::pair<int, int>, void*> g_sblock_cache;  // (device, nb) -> int2 list
std::map<std::pair<int, int>, void*> g_seg_cache;     // (device, nb) -> int4 list
std::map<std::pair<int, int>, void*> g_quad_cache;    // (device, nb) -> int2 list
std::map<std::pair<int, int>, void*> g_tw_cache;      // (device, ng) -> float2 twiddles
std::map<std::pair<int, int>, bool> g_lds_set;        // (device, ng) -> attribute set
}  // namespace

int current_device() {
    int d = 0;
    (void)hipGetDevice(&d);
    return d;
}

int get_plan(const mixer_array_cfg& c, XcorrPlan* p) {
    const int nb = c.n_elem / 32;
    p->n_elem = c.n_elem;
    p->n_slots = plan_slots(c);
    const mixer_tuning t = tuning_for(c.n_elem);
    // automatic choices, from timings of whole integrations on MI355X: every size correlates
    // fastest straight from the heaps with loads issued two blocks ahead (256 x 128 x 8192:
    // 0.31 ms vs 0.45 ms with a corner-turn pass; 1024 x 256 x 4096: 3.97 vs 4.12 ms)
    p->fused = plan_fused(c) ? 1 : 0;
    p->prefetch = t.prefetch > 0 ? t.prefetch : 3;
    p->ct_ch = t.ct_chan > 0 ? t.ct_chan : 8;
    p->mfma = (t.mfma == 16 && nb <= 8) ? 16 : 32;
    p->n_tasks = 0;
    p->d_tasks = nullptr;
    {
        const auto sg = segs_for(nb);
        p->n_segs = (int)sg.size();
        std::lock_guard<std::mutex> lk(g_cache_mtx);
        auto key = std::make_pair(current_device(), nb);
        auto it = g_seg_cache.find(key);
        if (it == g_seg_cache.end()) {
            void* d = nullptr;
            int st = hip_status(hipMalloc(&d, sg.size() * sizeof(int4)));
            if (st) return st;
            st = hip_status(hipMemcpy(d, sg.data(), sg.size() * sizeof(int4), hipMemcpyHostToDevice));
            if (st) return st;
            it = g_seg_cache.emplace(key, d).first;
        }
        p->d_segs = it->second;
    }
    p->kind = (XcorrPlan::Kind)plan_kind(c);
    // BLOCKED with one time slot and register staging two blocks ahead: the integrate can run in
    // the correlator's epilogue when one launch covers an integration (1024 x 256 x 4096: no
    // 1.1 GB accumulator store and read-back, no integrate launch)
    p->fuse_integ = p->kind == XcorrPlan::BLOCKED && p->n_slots == 1 && p->prefetch == 3 &&
                    p->fused && t.integrate != 1;
    p->n_quad = 0;
    p->d_quad = nullptr;
    if (nb % 4 == 0) {
        const auto qi = quad_items_for(nb);
        p->n_quad = (int)qi.size();
        std::lock_guard<std::mutex> lk(g_cache_mtx);
        auto key = std::make_pair(current_device(), nb);
        auto it = g_quad_cache.find(key);
        if (it == g_quad_cache.end()) {
            std::vector<int> h;
            for (auto& q : qi) { h.push_back(q.first); h.push_back(q.second); }
            void* d = nullptr;
            int st = hip_status(hipMalloc(&d, h.size() * sizeof(int)));
            if (st) return st;
            st = hip_status(hipMemcpy(d, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
            if (st) return st;
            it = g_quad_cache.emplace(key, d).first;
        }
        p->d_quad = it->second;
    }
    if (nb <= 8) return MIXER_OK;
    auto sb = sblocks_for(nb);
    p->n_tasks = (int)sb.size();
    std::lock_guard<std::mutex> lk(g_cache_mtx);
    auto key = std::make_pair(current_device(), nb);
    auto it = g_sblock_cache.find(key);
    if (it == g_sblock_cache.end()) {
        std::vector<int> h;
        for (auto& q : sb) { h.push_back(q.first); h.push_back(q.second); }
        void* d = nullptr;
        int st = hip_status(hipMalloc(&d, h.size() * sizeof(int)));
        if (st) return st;
        st = hip_status(hipMemcpy(d, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
        if (st) return st;
        it = g_sblock_cache.emplace(key, d).first;
    }
    p->d_tasks = it->second;
    return MIXER_OK;
}

// exp(-2 pi i q / ng), q < ng/2, from FP64 libm rounded once to f32.
int get_twiddles(int ng, const float2** out) {
    std::lock_guard<std::mutex> lk(g_cache_mtx);
    const int dev = current_device();
    auto key = std::make_pair(dev, ng);
    auto it = g_tw_cache.find(key);
    if (it == g_tw_cache.end()) {
        std::vector<float2> h(ng / 2);
        for (int q = 0; q < ng / 2; ++q) {
            const double a = 2.0 * M_PI * (double)q / (double)ng;
            h[q] = make_float2((float)std::cos(a), (float)-std::sin(a));
        }
        void* d = nullptr;
        int st = hip_status(hipMalloc(&d, h.size() * sizeof(float2)));
        if (st) return st;
        st = hip_status(hipMemcpy(d, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice));
        if (st) return st;
        it = g_tw_cache.emplace(key, d).first;
    }
    if (!g_lds_set[key]) {
        int st = hip_status(set_image_lds_limit(ng));
        if (st) return st;
        g_lds_set[key] = true;
    }
    *out = (const float2*)it->second;
    return MIXER_OK;
}

}  // namespace mixer

using namespace mixer;

extern "C" {

void mixer_default_cfg(mixer_array_cfg* c) {
    if (!c) return;
    memset(c, 0, sizeof(*c));
    c->n_elem = 256;      // MIXER_NANT, src/config.h:26
    c->grid_x = 16;       // MIXER_GRID, src/config.h:27
    c->n_chan = 128;      // MIXER_ACTIVE_NCH, src/config.h:29
    c->t_pkt = 32;        // MIXER_TPKT, src/config.h:30
    c->hdr_bytes = 64;    // MIXER_HDR_BYTES, src/config.h:32
    c->seq_byte = 48;     // MIXER_SEQ_BYTE, src/config.h:35
    c->k_integ = 8192;    // 8192 samples per 20 Hz integration
    c->k_update = 8192;   // one corner-turn + GEMM launch per integration (KPCH=1024 in the reference)
    c->ng = 32;           // NG, src/rx_host_corr.cu:51
    c->flags = 0;
    c->pkt_pitch = 16384; // MR_STRIDE, src/rx_host_corr.cu:47
}

int mixer_validate_cfg(const mixer_array_cfg* c) {
    if (!c) return MIXER_ERR_ARG;
    const int N = c->n_elem, nb = N / 32;
    if (N < 32 || N > 1024 || N % 32) return MIXER_ERR_ARG;
    if (nb > 8 && nb % 4) return MIXER_ERR_ARG;
    if (c->grid_x < 1 || c->grid_x > 32 || 32 % c->grid_x) return MIXER_ERR_ARG;  // rows of 32-element tiles
    if (c->n_chan < 1 || c->n_chan > 1024) return MIXER_ERR_ARG;
    if (c->t_pkt != 32) return MIXER_ERR_ARG;
    if (c->seq_byte < 0 || c->hdr_bytes < c->seq_byte + 4) return MIXER_ERR_ARG;
    if (c->k_update <= 0 || c->k_update % c->t_pkt) return MIXER_ERR_ARG;
    if (c->k_integ <= 0 || c->k_integ % c->k_update) return MIXER_ERR_ARG;
    // |Re| per sample <= 128*128 + 128*128 = 2^15, so an int32 sum is exact below 2^16 samples
    if (c->k_integ > 65504) return MIXER_ERR_OVERFLOW;
    const int gy = N / c->grid_x;
    if (c->ng < 8 || c->ng > 64 || (c->ng & (c->ng - 1))) return MIXER_ERR_ARG;
    if (c->ng < 2 * c->grid_x || c->ng < 2 * gy) return MIXER_ERR_ARG;
    if (c->pkt_pitch < (int64_t)mixer_wire_bytes(c)) return MIXER_ERR_ARG;
    return MIXER_OK;
}

size_t mixer_payload_bytes(const mixer_array_cfg* c) { return (size_t)2 * c->n_chan * c->t_pkt; }
size_t mixer_wire_bytes(const mixer_array_cfg* c) { return (size_t)c->hdr_bytes + mixer_payload_bytes(c); }

const char* mixer_version(void) { return "mixer-mi355x 0.1 (gfx950)"; }

const char* mixer_strerror(int s) {
    switch (s) {
    case MIXER_OK: return "ok";
    case MIXER_ERR_ARG: return "invalid argument or unsupported geometry";
    case MIXER_ERR_HIP: return "HIP runtime error";
    case MIXER_ERR_NOMEM: return "out of memory";
    case MIXER_ERR_STATE: return "invalid state";
    case MIXER_ERR_BUDGET: return "HBM budget exceeded";
    case MIXER_ERR_NODEV: return "no usable gfx950 device";
    case MIXER_ERR_OVERFLOW: return "samples per integration exceed the exact int32 range";
    default: return "unknown status";
    }
}

int mixer_device_count(int* n) {
    if (!n) return MIXER_ERR_ARG;
    *n = 0;
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess) return MIXER_OK;
    for (int d = 0; d < cnt; ++d) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, d) == hipSuccess && strncmp(p.gcnArchName, "gfx950", 6) == 0) ++*n;
    }
    return MIXER_OK;
}

size_t mixer_tiles_bytes(const mixer_array_cfg* c, int n_tb) {
    return (size_t)c->n_chan * (size_t)n_tb * 2 * (size_t)c->n_elem * 32;
}

int mixer_corner_turn(void* stream, const mixer_array_cfg* c, const uint8_t* d_heaps, int n_tb,
                      int8_t* d_tiles, uint32_t* d_wrap) {
    int st = mixer_validate_cfg(c);
    if (st) return st;
    if (!d_heaps || !d_tiles || n_tb <= 0) return MIXER_ERR_ARG;
    XcorrPlan p;
    if ((st = get_plan(*c, &p))) return st;
    return hip_status(launch_corner_turn((hipStream_t)stream, d_heaps, d_tiles, d_wrap, c->n_elem,
                                         c->n_chan, n_tb, 0, c->n_chan, p.ct_ch));
}

int mixer_acc_slots(const mixer_array_cfg* c) {
    if (mixer_validate_cfg(c)) return 0;
    return plan_slots(*c);
}

size_t mixer_acc_bytes(const mixer_array_cfg* c) {
    if (mixer_validate_cfg(c)) return 0;
    const size_t nb = c->n_elem / 32;
    return (size_t)plan_slots(*c) * c->n_chan * (nb * (nb + 1) / 2) * 2048 * sizeof(int32_t);
}

int mixer_xcorr(void* stream, const mixer_array_cfg* c, const int8_t* d_tiles, int n_tb,
                int32_t* d_acc, int accumulate) {
    int st = mixer_validate_cfg(c);
    if (st) return st;
    if (!d_tiles || !d_acc || n_tb <= 0) return MIXER_ERR_ARG;
    XcorrPlan p;
    if ((st = get_plan(*c, &p))) return st;
    if (n_tb < p.n_slots) return MIXER_ERR_ARG;
    if (p.kind == XcorrPlan::QUAD)  // tiles: the per-channel layouts
        p.kind = c->n_elem / 32 <= 8 ? XcorrPlan::FOLD : XcorrPlan::BLOCKED;
    return hip_status(launch_xcorr((hipStream_t)stream, p, xsrc_tiles(d_tiles, c->n_elem, n_tb),
                                   d_acc, nullptr, c->n_chan, n_tb, accumulate ? 1 : 0));
}

int mixer_xcorr_heaps(void* stream, const mixer_array_cfg* c, const uint8_t* d_heaps, int n_tb,
                      int32_t* d_acc, int accumulate, uint32_t* d_wrap) {
    int st = mixer_validate_cfg(c);
    if (st) return st;
    if (!d_heaps || !d_acc || n_tb <= 0) return MIXER_ERR_ARG;
    XcorrPlan p;
    if ((st = get_plan(*c, &p))) return st;
    if (n_tb < p.n_slots) return MIXER_ERR_ARG;
    return hip_status(launch_xcorr((hipStream_t)stream, p, xsrc_heaps(d_heaps, c->n_elem, c->n_chan),
                                   d_acc, d_wrap, c->n_chan, n_tb, accumulate ? 1 : 0));
}

int mixer_integrate(void* stream, const mixer_array_cfg* c, const int32_t* d_acc, float* d_v,
                    int accumulate) {
    int st = mixer_validate_cfg(c);
    if (st) return st;
    if (!d_acc || !d_v) return MIXER_ERR_ARG;
    XcorrPlan p;
    if ((st = get_plan(*c, &p))) return st;
    return hip_status(launch_integrate((hipStream_t)stream, p, d_acc, accumulate ? nullptr : d_v,
                                       accumulate ? d_v : nullptr, nullptr, c->n_elem, c->n_chan,
                                       c->grid_x, c->ng));
}

// Gridding runs as the grid-only form of the integrate kernel into a stream-ordered int64
// workspace, then the per-channel FFT kernel reads (and clears) it.
int mixer_image(void* stream, const mixer_array_cfg* c, const int32_t* d_acc, float* d_img,
                float* d_vmax, float* d_vmin) {
    int st = mixer_validate_cfg(c);
    if (st) return st;
    if (!d_acc || !d_img || !d_vmax || !d_vmin) return MIXER_ERR_ARG;
    const float2* tw = nullptr;
    if ((st = get_twiddles(c->ng, &tw))) return st;
    XcorrPlan p;
    if ((st = get_plan(*c, &p))) return st;
    hipStream_t s = (hipStream_t)stream;
    const size_t gbytes = (size_t)c->n_chan * c->ng * c->ng * 16;
    void* g = nullptr;
    if ((st = hip_status(hipMallocAsync(&g, gbytes, s)))) return st;
    st = hip_status(hipMemsetAsync(g, 0, gbytes, s));
    if (!st)
        st = hip_status(launch_integrate(s, p, d_acc, nullptr, nullptr, (unsigned long long*)g,
                                         c->n_elem, c->n_chan, c->grid_x, c->ng));
    if (!st)
        st = hip_status(launch_image(s, (unsigned long long*)g, tw, d_img, d_vmax, d_vmin, c->n_chan,
                                     c->ng));
    const int st2 = hip_status(hipFreeAsync(g, s));
    return st ? st : st2;
}

int mixer_set_tuning(int n_elem, const mixer_tuning* t) {
    if (!t || n_elem < 32 || n_elem > 1024 || n_elem % 32) return MIXER_ERR_ARG;
    if (t->slots < 0 || t->slots > 8) return MIXER_ERR_ARG;
    if (t->prefetch < 0 || t->prefetch > 5) return MIXER_ERR_ARG;
    if (t->ct_chan != 0 && t->ct_chan != 8 && t->ct_chan != 16 && t->ct_chan != 32) return MIXER_ERR_ARG;
    if (t->staging < 0 || t->staging > 2) return MIXER_ERR_ARG;
    if (t->mfma != 0 && t->mfma != 16 && t->mfma != 32) return MIXER_ERR_ARG;
    if (t->integrate < 0 || t->integrate > 2) return MIXER_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_tune_mtx);
    g_tuning[n_elem] = *t;
    return MIXER_OK;
}

int mixer_get_tuning(const mixer_array_cfg* c, mixer_tuning* t) {
    int st = mixer_validate_cfg(c);
    if (st) return st;
    if (!t) return MIXER_ERR_ARG;
    XcorrPlan p;
    if ((st = get_plan(*c, &p))) return st;
    t->slots = p.n_slots;
    t->prefetch = p.prefetch;
    t->ct_chan = p.ct_ch;
    t->staging = p.fused ? 1 : 2;
    t->mfma = p.mfma;
    t->integrate = p.fuse_integ ? 2 : 1;
    return MIXER_OK;
}

int mixer_xcorr_plain(void* stream, const mixer_array_cfg* c, const uint8_t* d_heaps, int n_tb,
                      int8_t* d_tiles, int32_t* d_vre, int32_t* d_vim) {
    int st = mixer_validate_cfg(c);
    if (st) return st;
    if (!d_heaps || !d_tiles || !d_vre || !d_vim || n_tb <= 0) return MIXER_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    for (int ch = 0; ch < c->n_chan; ++ch) {
        if ((st = hip_status(launch_corner_turn(s, d_heaps, d_tiles, nullptr, c->n_elem, c->n_chan,
                                                n_tb, ch, 1, 8))))
            return st;
        if ((st = hip_status(launch_xcorr_plain(s, d_tiles, d_vre, d_vim, c->n_elem, n_tb, ch))))
            return st;
    }
    return MIXER_OK;
}

// ---- plumbing --------------------------------------------------------------------------------
int mixer_set_device(int device) { return hip_status(hipSetDevice(device)); }
int mixer_dev_alloc(size_t bytes, void** ptr) {
    if (!ptr) return MIXER_ERR_ARG;
    return hip_status(hipMalloc(ptr, bytes));
}
int mixer_dev_free(void* ptr) { return hip_status(hipFree(ptr)); }
int mixer_host_alloc(size_t bytes, void** ptr) {
    if (!ptr) return MIXER_ERR_ARG;
    return hip_status(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
}
int mixer_host_free(void* ptr) { return hip_status(hipHostFree(ptr)); }
int mixer_memcpy(void* dst, const void* src, size_t bytes, int kind, void* stream) {
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                    : kind == 1 ? hipMemcpyDeviceToHost
                                : hipMemcpyDeviceToDevice;
    return hip_status(hipMemcpyAsync(dst, src, bytes, k, (hipStream_t)stream));
}
int mixer_memset(void* ptr, int value, size_t bytes, void* stream) {
    return hip_status(hipMemsetAsync(ptr, value, bytes, (hipStream_t)stream));
}
int mixer_stream_create(void** s) {
    if (!s) return MIXER_ERR_ARG;
    return hip_status(hipStreamCreateWithFlags((hipStream_t*)s, hipStreamNonBlocking));
}
int mixer_stream_destroy(void* s) { return hip_status(hipStreamDestroy((hipStream_t)s)); }
int mixer_stream_sync(void* s) { return hip_status(hipStreamSynchronize((hipStream_t)s)); }
int mixer_event_create(void** e) {
    if (!e) return MIXER_ERR_ARG;
    return hip_status(hipEventCreate((hipEvent_t*)e));
}
int mixer_event_destroy(void* e) { return hip_status(hipEventDestroy((hipEvent_t)e)); }
int mixer_event_record(void* e, void* s) { return hip_status(hipEventRecord((hipEvent_t)e, (hipStream_t)s)); }
int mixer_event_elapsed_ms(void* a, void* b, float* ms) {
    int st = hip_status(hipEventSynchronize((hipEvent_t)b));
    if (st) return st;
    return hip_status(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
}
int mixer_device_sync(void) { return hip_status(hipDeviceSynchronize()); }

}  // extern "C"
