// engine.cpp -- the X-engine context: the reference's gpu_thread (ingest + correlate,
// src/rx_host_corr.cu:143-175) and its monitor/dump loop (src/rx_host_corr.cu:213-239) as a
// library object on the HIP runtime.
//
// Streams: s_copy carries host->device header-strip copies (and device->device leftovers),
// s_comp runs corner-turn -> correlate (every V-update) and integrate -> image (every
// integration), s_pub carries dumps and image read-outs. Events order them:
//   ring_free[b]  s_comp finished the corner-turn reading ring buffer b -> s_copy may refill b
//   copy_done     s_copy filled the current V-update -> s_comp may corner-turn it
//   integ_ev[b]   V/image buffer b complete -> s_pub may read it. Recorded by the reader at
//                 the tail of s_comp (under the context mutex, so the latest integration is
//                 already enqueued there), not after every integration: an event record's
//                 system-scope release idles s_comp for ~13 us, and most integrations are
//                 never read.
//   pub_done[b]   s_pub finished reading buffer b -> s_comp may overwrite it
// The int32 accumulator is written only by s_comp, so it has a single writer by construction
// (the reference serialises two compute streams with vacc_ev, src/rx_host_corr.cu:150,166).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <vector>

#include "internal.h"
#include "kernels.h"
#include "mixer.h"

using namespace mixer;

namespace {

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

uint32_t read_seq_be(const uint8_t* p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}

struct TimedLaunch {
    int stage;
    hipEvent_t a, b;
};

}  // namespace

struct mixer_ctx {
    mixer_array_cfg cfg{};
    int device = 0;
    XcorrPlan plan{};
    int n_slots = 1;
    int n_tb_upd = 0;       // time blocks per V-update
    int upd_per_integ = 0;  // V-updates per integration
    bool fuse_integ = false;  // one correlator launch per integration writes V and the grid
    size_t payload = 0;     // bytes per packet payload
    size_t tb_bytes = 0;    // bytes per time block (n_elem payloads)
    size_t v_bytes = 0;     // f32 cube bytes
    size_t img_floats = 0;

    hipStream_t s_comp = nullptr, s_copy = nullptr, s_pub = nullptr;
    uint8_t* d_ring[2] = {nullptr, nullptr};
    hipEvent_t ring_free[2] = {nullptr, nullptr};
    bool ring_used[2] = {false, false};
    int ring_cur = 0, ring_fill = 0;
    bool copy_wait = false;  // s_copy must wait for ring_free[ring_cur] before its next copy
    int8_t* d_tiles = nullptr;
    int32_t* d_acc = nullptr;
    uint32_t* d_wrap = nullptr;
    unsigned long long* d_G = nullptr;  // int64 image grid [C][ng*ng][2], cleared by the image kernel
    float* d_v[2] = {nullptr, nullptr};
    float* d_vcum = nullptr;
    float* d_img[2] = {nullptr, nullptr};  // [C][ng][ng] then vmax[C], vmin[C]
    hipEvent_t integ_ev[2] = {nullptr, nullptr};
    hipEvent_t pub_done[2] = {nullptr, nullptr};
    // frame (MIXER_CFG_FRAME_DV): int64 grid summed since the last mixer_read_frame, its image
    unsigned long long* d_Gf = nullptr;
    float* d_imgf = nullptr;
    hipEvent_t frame_ev = nullptr, pubf_done = nullptr;
    uint64_t frame_mark = 0;  // integrations covered by frames read so far
    hipEvent_t pub_done_cum = nullptr;
    hipEvent_t copy_done = nullptr;
    const float2* d_tw = nullptr;

    int upd_in_integ = 0;
    std::atomic<uint64_t> n_integ{0};  // integrations enqueued
    uint64_t last_arrival_ns = 0;

    // tokens: events recorded after the last use of a caller's buffer. Token t uses slot
    // t % NTOK; a slot is re-recorded only after its previous token's event has completed, so a
    // token whose slot holds a later token is complete.
    static constexpr int NTOK = 256;
    hipEvent_t tok_ev[NTOK] = {};
    uint64_t tok_owner[NTOK] = {};
    uint64_t tok_next = 1;

    // bounded ingest (mixer_set_ingest_slots): a time block of mixer_submit_heaps holds a slot
    // until its header-strip copy has completed; batches of slots with their copy events
    int slots_max = 0;  // 0: unbounded
    int slots_busy = 0;
    struct SlotBatch { hipEvent_t ev; int n; };
    std::deque<SlotBatch> slot_inflight;
    std::vector<hipEvent_t> slot_pool;  // submitting thread only

    // latency: per enqueued integration, (event, arrival)
    struct Pending { uint64_t idx; hipEvent_t ev; uint64_t arrival; };
    std::deque<Pending> pending;
    std::vector<hipEvent_t> ev_pool;

    // timing (guarded by tmtx: the submitting thread appends, mixer_get_timing resolves)
    std::mutex tmtx;
    unsigned timing = 0;  // bit s: time stage s
    std::vector<TimedLaunch> timed;
    double t_ms[4] = {0, 0, 0, 0};
    uint64_t t_n[4] = {0, 0, 0, 0};

    std::mutex mtx;  // guards the integration/dump handshake and every field of st
    mixer_stats st{};
    uint64_t frames_published = 0;
};

namespace {

// The imaginary -128 counter the correlator and the corner-turn add to, or none when the context
// was created with MIXER_CFG_NO_WRAP_CHECK.
uint32_t* wrap_of(const mixer_ctx* x) {
    return (x->cfg.flags & MIXER_CFG_NO_WRAP_CHECK) ? nullptr : x->d_wrap;
}

int take_event(mixer_ctx* x, hipEvent_t* e) {
    if (!x->ev_pool.empty()) {
        *e = x->ev_pool.back();
        x->ev_pool.pop_back();
        return MIXER_OK;
    }
    // completion only (the host reads nothing through these): no system-scope release, whose
    // L2 writeback idles the compute stream for ~20 us before the next integration
    return hip_status(hipEventCreateWithFlags(e, hipEventDisableTiming | hipEventDisableSystemFence));
}

// Resolve completed timing pairs (all of them if wait). Caller holds tmtx.
int resolve_timing(mixer_ctx* x, bool wait) {
    size_t k = 0;
    for (; k < x->timed.size(); ++k) {
        TimedLaunch& t = x->timed[k];
        if (!wait && hipEventQuery(t.b) != hipSuccess) break;
        MIXER_HIP(hipEventSynchronize(t.b));
        float ms = 0;
        MIXER_HIP(hipEventElapsedTime(&ms, t.a, t.b));
        x->t_ms[t.stage] += ms;
        x->t_n[t.stage] += 1;
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    x->timed.erase(x->timed.begin(), x->timed.begin() + k);
    return MIXER_OK;
}

int time_begin(mixer_ctx* x, hipEvent_t* a, int stage) {
    std::lock_guard<std::mutex> lk(x->tmtx);
    if (!(x->timing & (1u << stage))) return MIXER_OK;
    MIXER_HIP(hipEventCreate(a));
    return hip_status(hipEventRecord(*a, x->s_comp));
}
int time_end(mixer_ctx* x, int stage, hipEvent_t a) {
    std::lock_guard<std::mutex> lk(x->tmtx);
    if (!(x->timing & (1u << stage)) || !a) return MIXER_OK;
    hipEvent_t b;
    MIXER_HIP(hipEventCreate(&b));
    MIXER_HIP(hipEventRecord(b, x->s_comp));
    x->timed.push_back({stage, a, b});
    if (x->timed.size() > 4096) return resolve_timing(x, false);
    return MIXER_OK;
}

// Poll finished integrations for latency statistics.
void poll_pending(mixer_ctx* x) {
    while (!x->pending.empty()) {
        auto& p = x->pending.front();
        if (hipEventQuery(p.ev) != hipSuccess) break;
        const uint64_t t = now_ns();
        x->st.last_dump_latency_ms = t > p.arrival ? (double)(t - p.arrival) * 1e-6 : 0.0;
        x->ev_pool.push_back(p.ev);
        x->pending.pop_front();
    }
}

// Integrate + image of the completed integration (s_comp). With fused_src (one correlator launch
// covers the integration, fuse_integ) the correlator itself writes V and the grid from its
// epilogue, after the same waits for the reader.
int finish_integration(mixer_ctx* x, const XSrc* fused_src) {
    std::lock_guard<std::mutex> lk(x->mtx);
    const uint64_t idx = x->n_integ.load();
    const int b = (int)(idx % 2);
    MIXER_HIP(hipStreamWaitEvent(x->s_comp, x->pub_done[b], 0));
    if (x->d_vcum) MIXER_HIP(hipStreamWaitEvent(x->s_comp, x->pub_done_cum, 0));
    hipEvent_t t0 = nullptr;
    const bool image = !(x->cfg.flags & MIXER_CFG_NO_IMAGE);
    if (fused_src) {
        MIXER_TRY(time_begin(x, &t0, 1));
        const VOutArgs vo{x->d_v[b], x->d_vcum, image ? x->d_G : nullptr, x->cfg.n_elem,
                          x->cfg.grid_x, x->cfg.ng};
        MIXER_HIP(launch_xcorr_v(x->s_comp, x->plan, *fused_src, vo, wrap_of(x), x->cfg.n_chan,
                                 x->n_tb_upd));
        MIXER_TRY(time_end(x, 1, t0));
        x->st.vupdates++;
        x->st.time_blocks += x->n_tb_upd;
    } else {
        MIXER_TRY(time_begin(x, &t0, 2));
        MIXER_HIP(launch_integrate(x->s_comp, x->plan, x->d_acc, x->d_v[b], x->d_vcum,
                                   image ? x->d_G : nullptr, x->cfg.n_elem, x->cfg.n_chan,
                                   x->cfg.grid_x, x->cfg.ng));
        MIXER_TRY(time_end(x, 2, t0));
    }
    if (image) {
        MIXER_TRY(time_begin(x, &t0, 3));
        float* img = x->d_img[b];
        MIXER_HIP(launch_image(x->s_comp, x->d_G, x->d_tw, img, img + x->img_floats,
                               img + x->img_floats + x->cfg.n_chan, x->cfg.n_chan, x->cfg.ng,
                               x->d_Gf));
        MIXER_TRY(time_end(x, 3, t0));
    }
    // latency probe, at most one in flight: each event record idles the compute stream for a
    // few microseconds, and at real-time rates the previous probe has long completed
    if (x->pending.empty()) {
        hipEvent_t e;
        MIXER_TRY(take_event(x, &e));
        MIXER_HIP(hipEventRecord(e, x->s_comp));
        x->pending.push_back({idx, e, x->last_arrival_ns});
    }
    x->n_integ.store(idx + 1);
    x->st.integrations = idx + 1;
    x->upd_in_integ = 0;
    return MIXER_OK;
}

// Corner-turn of one V-update whose heaps start at src (device); nothing when the correlator
// stages straight from the heaps.
int run_update(mixer_ctx* x, const uint8_t* src) {
    if (x->plan.fused) return MIXER_OK;
    hipEvent_t t0 = nullptr;
    MIXER_TRY(time_begin(x, &t0, 0));
    MIXER_HIP(launch_corner_turn(x->s_comp, src, x->d_tiles, wrap_of(x), x->cfg.n_elem,
                                 x->cfg.n_chan, x->n_tb_upd, 0, x->cfg.n_chan, x->plan.ct_ch));
    MIXER_TRY(time_end(x, 0, t0));
    return MIXER_OK;
}
// Correlate one V-update (from the heaps at src when fused, else from the tiles).
int run_gemm(mixer_ctx* x, const uint8_t* src) {
    if (x->fuse_integ) {
        const XSrc xs = xsrc_heaps(src, x->cfg.n_elem, x->cfg.n_chan);
        return finish_integration(x, &xs);
    }
    hipEvent_t t0 = nullptr;
    MIXER_TRY(time_begin(x, &t0, 1));
    const XSrc xs = x->plan.fused ? xsrc_heaps(src, x->cfg.n_elem, x->cfg.n_chan)
                                  : xsrc_tiles(x->d_tiles, x->cfg.n_elem, x->n_tb_upd);
    MIXER_HIP(launch_xcorr(x->s_comp, x->plan, xs, x->d_acc, x->plan.fused ? wrap_of(x) : nullptr,
                           x->cfg.n_chan, x->n_tb_upd, x->upd_in_integ > 0 ? 1 : 0));
    MIXER_TRY(time_end(x, 1, t0));
    x->upd_in_integ++;
    {
        std::lock_guard<std::mutex> lk(x->mtx);
        x->st.vupdates++;
        x->st.time_blocks += x->n_tb_upd;
    }
    if (x->upd_in_integ == x->upd_per_integ) MIXER_TRY(finish_integration(x, nullptr));
    return MIXER_OK;
}

// The current ring buffer is full: hand it to s_comp.
int dispatch_ring(mixer_ctx* x) {
    const int b = x->ring_cur;
    MIXER_HIP(hipEventRecord(x->copy_done, x->s_copy));
    MIXER_HIP(hipStreamWaitEvent(x->s_comp, x->copy_done, 0));
    const bool flop = !(x->cfg.flags & MIXER_CFG_NOFLOP);
    if (flop) MIXER_TRY(run_update(x, x->d_ring[b]));
    if (flop && x->plan.fused) MIXER_TRY(run_gemm(x, x->d_ring[b]));  // reads the ring itself
    MIXER_HIP(hipEventRecord(x->ring_free[b], x->s_comp));
    x->ring_used[b] = true;
    if (flop && !x->plan.fused) MIXER_TRY(run_gemm(x, x->d_ring[b]));
    if (!flop) {
        std::lock_guard<std::mutex> lk(x->mtx);
        x->st.time_blocks += x->n_tb_upd;
    }
    x->ring_cur ^= 1;
    x->ring_fill = 0;
    // s_copy waits for the compute that still reads the next ring buffer just before its next
    // copy into it, not here: an event recorded on s_copy meanwhile (a caller's token, a slot
    // batch) then covers the copies already enqueued and not that compute
    x->copy_wait = x->ring_used[x->ring_cur];
    return MIXER_OK;
}

// Called before every copy into the current ring buffer.
int ring_copy_wait(mixer_ctx* x) {
    if (!x->copy_wait) return MIXER_OK;
    x->copy_wait = false;
    return hip_status(hipStreamWaitEvent(x->s_copy, x->ring_free[x->ring_cur], 0));
}

int new_token(mixer_ctx* x, hipStream_t s, uint64_t* token) {
    const uint64_t t = x->tok_next;
    const int k = (int)(t % mixer_ctx::NTOK);
    hipEvent_t& e = x->tok_ev[k];
    if (!e) MIXER_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventDisableSystemFence));
    // the slot's previous token must be complete before its event is re-recorded (back-pressure
    // when more than NTOK tokens are outstanding)
    else MIXER_HIP(hipEventSynchronize(e));
    MIXER_HIP(hipEventRecord(e, s));
    x->tok_owner[k] = t;
    x->tok_next = t + 1;
    if (token) *token = t;
    return MIXER_OK;
}

// Return the slots whose header-strip copies have completed (the reference's gpu_thread polls
// its in-flight slots the same way before taking the next one, src/rx_host_corr.cu:154).
void retire_slots(mixer_ctx* x) {
    while (!x->slot_inflight.empty() && hipEventQuery(x->slot_inflight.front().ev) == hipSuccess) {
        x->slots_busy -= x->slot_inflight.front().n;
        x->slot_pool.push_back(x->slot_inflight.front().ev);
        x->slot_inflight.pop_front();
    }
}

void destroy_ctx(mixer_ctx* x) {
    if (!x) return;
    (void)hipSetDevice(x->device);
    if (x->s_comp) (void)hipStreamSynchronize(x->s_comp);
    if (x->s_copy) (void)hipStreamSynchronize(x->s_copy);
    if (x->s_pub) (void)hipStreamSynchronize(x->s_pub);
    {
        std::lock_guard<std::mutex> lk(x->tmtx);
        resolve_timing(x, true);
    }
    for (auto& p : x->pending) (void)hipEventDestroy(p.ev);
    for (auto e : x->ev_pool) (void)hipEventDestroy(e);
    for (auto e : x->tok_ev) if (e) (void)hipEventDestroy(e);
    for (auto& sb : x->slot_inflight) (void)hipEventDestroy(sb.ev);
    for (auto e : x->slot_pool) (void)hipEventDestroy(e);
    for (int b = 0; b < 2; ++b) {
        if (x->d_ring[b]) (void)hipFree(x->d_ring[b]);
        if (x->d_v[b]) (void)hipFree(x->d_v[b]);
        if (x->d_img[b]) (void)hipFree(x->d_img[b]);
        if (x->ring_free[b]) (void)hipEventDestroy(x->ring_free[b]);
        if (x->integ_ev[b]) (void)hipEventDestroy(x->integ_ev[b]);
        if (x->pub_done[b]) (void)hipEventDestroy(x->pub_done[b]);
    }
    if (x->pub_done_cum) (void)hipEventDestroy(x->pub_done_cum);
    if (x->frame_ev) (void)hipEventDestroy(x->frame_ev);
    if (x->pubf_done) (void)hipEventDestroy(x->pubf_done);
    if (x->d_Gf) (void)hipFree(x->d_Gf);
    if (x->d_imgf) (void)hipFree(x->d_imgf);
    if (x->copy_done) (void)hipEventDestroy(x->copy_done);
    if (x->d_vcum) (void)hipFree(x->d_vcum);
    if (x->d_tiles) (void)hipFree(x->d_tiles);
    if (x->d_acc) (void)hipFree(x->d_acc);
    if (x->d_wrap) (void)hipFree(x->d_wrap);
    if (x->d_G) (void)hipFree(x->d_G);
    if (x->s_comp) (void)hipStreamDestroy(x->s_comp);
    if (x->s_copy) (void)hipStreamDestroy(x->s_copy);
    if (x->s_pub) (void)hipStreamDestroy(x->s_pub);
    delete x;
}

int create_ctx(const mixer_array_cfg* cfg, int device, mixer_ctx* x) {
    x->cfg = *cfg;
    x->device = device;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return MIXER_ERR_NODEV;
    MIXER_HIP(hipSetDevice(device));
    if ((cfg->flags & MIXER_CFG_AUTOTUNE) && !has_tuning(cfg->n_elem)) {
        mixer_tuning best;
        MIXER_TRY(mixer_autotune(cfg, device, &best));
    }
    MIXER_TRY(get_plan(*cfg, &x->plan));
    x->n_slots = x->plan.n_slots;
    x->n_tb_upd = cfg->k_update / cfg->t_pkt;
    x->upd_per_integ = cfg->k_integ / cfg->k_update;
    x->fuse_integ = x->plan.fuse_integ && x->upd_per_integ == 1;
    x->payload = mixer_payload_bytes(cfg);
    x->tb_bytes = x->payload * cfg->n_elem;
    x->v_bytes = (size_t)cfg->n_chan * 2 * cfg->n_elem * cfg->n_elem * sizeof(float);
    x->img_floats = (size_t)cfg->n_chan * cfg->ng * cfg->ng;
    MIXER_HIP(hipStreamCreateWithFlags(&x->s_comp, hipStreamNonBlocking));
    MIXER_HIP(hipStreamCreateWithFlags(&x->s_copy, hipStreamNonBlocking));
    MIXER_HIP(hipStreamCreateWithFlags(&x->s_pub, hipStreamNonBlocking));
    for (int b = 0; b < 2; ++b) {
        MIXER_HIP(hipMalloc(&x->d_ring[b], x->tb_bytes * x->n_tb_upd));
        MIXER_HIP(hipMalloc(&x->d_v[b], x->v_bytes));
        MIXER_HIP(hipMalloc(&x->d_img[b], (x->img_floats + 2 * cfg->n_chan) * sizeof(float)));
        MIXER_HIP(hipEventCreateWithFlags(&x->ring_free[b], hipEventDisableTiming));
        MIXER_HIP(hipEventCreateWithFlags(&x->integ_ev[b], hipEventDisableTiming));
        MIXER_HIP(hipEventCreateWithFlags(&x->pub_done[b], hipEventDisableTiming));
        MIXER_HIP(hipEventRecord(x->pub_done[b], x->s_pub));
    }
    MIXER_HIP(hipEventCreateWithFlags(&x->pub_done_cum, hipEventDisableTiming));
    MIXER_HIP(hipEventRecord(x->pub_done_cum, x->s_pub));
    MIXER_HIP(hipEventCreateWithFlags(&x->copy_done, hipEventDisableTiming));
    if (cfg->flags & MIXER_CFG_CUMULATIVE) {
        MIXER_HIP(hipMalloc(&x->d_vcum, x->v_bytes));
        MIXER_HIP(hipMemset(x->d_vcum, 0, x->v_bytes));  // zeroed once (src/rx_host_corr.cu:188)
    }
    if (!x->plan.fused) MIXER_HIP(hipMalloc(&x->d_tiles, mixer_tiles_bytes(cfg, x->n_tb_upd)));
    if (!x->fuse_integ) MIXER_HIP(hipMalloc(&x->d_acc, mixer_acc_bytes(cfg)));
    MIXER_HIP(hipMalloc(&x->d_wrap, sizeof(uint32_t)));
    MIXER_HIP(hipMemset(x->d_wrap, 0, sizeof(uint32_t)));
    const size_t gbytes = (size_t)cfg->n_chan * cfg->ng * cfg->ng * 16;
    MIXER_HIP(hipMalloc(&x->d_G, gbytes));
    MIXER_HIP(hipMemset(x->d_G, 0, gbytes));
    if ((cfg->flags & MIXER_CFG_FRAME_DV) && !(cfg->flags & MIXER_CFG_NO_IMAGE)) {
        MIXER_HIP(hipMalloc(&x->d_Gf, gbytes));
        MIXER_HIP(hipMemset(x->d_Gf, 0, gbytes));
        MIXER_HIP(hipMalloc(&x->d_imgf, (x->img_floats + 2 * cfg->n_chan) * sizeof(float)));
        MIXER_HIP(hipEventCreateWithFlags(&x->frame_ev, hipEventDisableTiming));
        MIXER_HIP(hipEventCreateWithFlags(&x->pubf_done, hipEventDisableTiming));
        MIXER_HIP(hipEventRecord(x->pubf_done, x->s_pub));
    }
    MIXER_TRY(get_twiddles(cfg->ng, &x->d_tw));
    MIXER_HIP(hipDeviceSynchronize());
    return MIXER_OK;
}

}  // namespace

extern "C" {

int mixer_create(const mixer_array_cfg* cfg, int device, mixer_ctx** out) {
    if (!out) return MIXER_ERR_ARG;
    *out = nullptr;
    MIXER_TRY(mixer_validate_cfg(cfg));
    mixer_ctx* x = new mixer_ctx();
    int st = create_ctx(cfg, device, x);
    if (st) {
        destroy_ctx(x);
        return st;
    }
    *out = x;
    return MIXER_OK;
}

int mixer_destroy(mixer_ctx* x) {
    destroy_ctx(x);
    return MIXER_OK;
}

int mixer_get_cfg(mixer_ctx* x, mixer_array_cfg* c) {
    if (!x || !c) return MIXER_ERR_ARG;
    *c = x->cfg;
    return MIXER_OK;
}

void* mixer_compute_stream(mixer_ctx* x) { return x ? (void*)x->s_comp : nullptr; }

// Input seam: src/rx_host_corr.cu:134-141 (alignment, groups of n_elem packets) and :159 (the
// header-strip copy), with per-packet seq checks the reference does not make. Consecutive
// valid time blocks are consecutive rows at the source pitch and consecutive ring slots, so a
// run of them (up to kMaxCopyRows packets, within one ring buffer) goes as one 2-D copy.
constexpr int kMaxCopyRows = 16384;

int mixer_submit_heaps(mixer_ctx* x, const uint8_t* pkts, size_t pitch, int npkts,
                       uint64_t arrival_ns, uint64_t* token) {
    if (!x || !pkts || npkts < 0 || pitch < mixer_wire_bytes(&x->cfg)) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(x->device));
    const int N = x->cfg.n_elem, sb = x->cfg.seq_byte, hdr = x->cfg.hdr_bytes;
    retire_slots(x);
    const uint8_t* run0 = nullptr;  // first packet of the pending run of valid blocks
    int run_n = 0;                  // its time blocks, the last run_n filled ring slots
    uint64_t d_heaps = 0, d_missed = 0, d_gaps = 0, d_misalign = 0;  // stats of this call
    int slots_taken = 0;
    auto flush = [&]() -> int {
        if (!run_n) return MIXER_OK;
        MIXER_HIP(hipMemcpy2DAsync(x->d_ring[x->ring_cur] + (size_t)(x->ring_fill - run_n) * x->tb_bytes,
                                   x->payload, run0 + hdr, pitch, x->payload, (size_t)N * run_n,
                                   hipMemcpyHostToDevice, x->s_copy));
        run_n = 0;
        return MIXER_OK;
    };
    auto body = [&]() -> int {
        int i = 0;
        if (npkts > 0) {
            const uint32_t seq0 = read_seq_be(pkts + sb);
            const int off = (int)((N - seq0 % N) % N);
            if (off) d_misalign++;
            i = off;
        }
        while (i + N <= npkts) {
            const uint8_t* g = pkts + (size_t)i * pitch;
            const uint32_t s0 = read_seq_be(g + sb);
            if (s0 % N) {  // lost alignment after a gap: resynchronise on the next block start
                MIXER_TRY(flush());
                ++i;
                continue;
            }
            int bad = -1;
            for (int k = 1; k < N; ++k)
                if (read_seq_be(g + (size_t)k * pitch + sb) != s0 + (uint32_t)k) { bad = k; break; }
            if (bad >= 0) {
                d_gaps++;
                MIXER_TRY(flush());
                i += bad;
                continue;
            }
            // no free slot: the block is dropped and counted (src/rx_host_corr.cu:139-140)
            if (x->slots_max > 0 && x->slots_busy >= x->slots_max) {
                d_missed++;
                MIXER_TRY(flush());
                i += N;
                continue;
            }
            if (x->slots_max > 0) { x->slots_busy++; slots_taken++; }
            if (run_n == 0) {
                MIXER_TRY(ring_copy_wait(x));
                run0 = g;
            }
            ++run_n;
            d_heaps += N;
            i += N;
            x->last_arrival_ns = arrival_ns;
            ++x->ring_fill;
            if (x->ring_fill == x->n_tb_upd || N * (run_n + 1) > kMaxCopyRows) MIXER_TRY(flush());
            if (x->ring_fill == x->n_tb_upd) MIXER_TRY(dispatch_ring(x));
        }
        return flush();
    };
    int st = body();
    if (!st && slots_taken > 0) {  // the slots of this call are free once its copies are done
        hipEvent_t e = nullptr;
        if (!x->slot_pool.empty()) {
            e = x->slot_pool.back();
            x->slot_pool.pop_back();
        } else {
            st = hip_status(hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventDisableSystemFence));
        }
        if (!st) st = hip_status(hipEventRecord(e, x->s_copy));
        if (!st) x->slot_inflight.push_back({e, slots_taken});
    }
    if (!st && token) st = new_token(x, x->s_copy, token);
    std::lock_guard<std::mutex> lk(x->mtx);
    x->st.heaps_in += d_heaps;
    x->st.imissed += d_missed;
    x->st.seq_gaps += d_gaps;
    x->st.misalign += d_misalign;
    poll_pending(x);
    return st;
}

int mixer_submit_device(mixer_ctx* x, const uint8_t* d, int n_tb, uint64_t arrival_ns,
                        uint64_t* token) {
    if (!x || !d || n_tb < 0) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(x->device));
    x->last_arrival_ns = arrival_ns;
    int done = 0;
    bool copied = false;  // a copy of the caller's buffer is queued on s_copy
    uint64_t d_heaps = 0;
    auto body = [&]() -> int {
        // complete a partly filled ring buffer first
        while (done < n_tb && x->ring_fill > 0) {
            copied = true;
            const int take = std::min(n_tb - done, x->n_tb_upd - x->ring_fill);
            MIXER_HIP(hipMemcpyAsync(x->d_ring[x->ring_cur] + (size_t)x->ring_fill * x->tb_bytes,
                                     d + (size_t)done * x->tb_bytes, (size_t)take * x->tb_bytes,
                                     hipMemcpyDeviceToDevice, x->s_copy));
            done += take;
            x->ring_fill += take;
            d_heaps += (uint64_t)take * x->cfg.n_elem;
            if (x->ring_fill == x->n_tb_upd) MIXER_TRY(dispatch_ring(x));
        }
        // whole V-updates straight from the caller's buffer
        while (n_tb - done >= x->n_tb_upd) {
            if (!(x->cfg.flags & MIXER_CFG_NOFLOP)) {
                MIXER_TRY(run_update(x, d + (size_t)done * x->tb_bytes));
                MIXER_TRY(run_gemm(x, d + (size_t)done * x->tb_bytes));
            }
            done += x->n_tb_upd;
            d_heaps += (uint64_t)x->n_tb_upd * x->cfg.n_elem;
        }
        // leftovers into the ring
        if (done < n_tb) {
            const int take = n_tb - done;
            copied = true;
            MIXER_TRY(ring_copy_wait(x));
            MIXER_HIP(hipMemcpyAsync(x->d_ring[x->ring_cur], d + (size_t)done * x->tb_bytes,
                                     (size_t)take * x->tb_bytes, hipMemcpyDeviceToDevice, x->s_copy));
            x->ring_fill = take;
            d_heaps += (uint64_t)take * x->cfg.n_elem;
        }
        // the caller's buffer is free once both streams are past it (a cross-stream wait only
        // when something was copied: it costs the compute stream an idle gap of tens of
        // microseconds)
        if (copied) {
            MIXER_HIP(hipEventRecord(x->copy_done, x->s_copy));
            MIXER_HIP(hipStreamWaitEvent(x->s_comp, x->copy_done, 0));
        }
        if (token) MIXER_TRY(new_token(x, x->s_comp, token));
        return MIXER_OK;
    };
    const int st = body();
    std::lock_guard<std::mutex> lk(x->mtx);
    x->st.heaps_in += d_heaps;
    poll_pending(x);
    return st;
}

int mixer_token_done(mixer_ctx* x, uint64_t t) {
    if (!x || t == 0 || t >= x->tok_next) return MIXER_ERR_ARG;
    const int k = (int)(t % mixer_ctx::NTOK);
    if (x->tok_owner[k] != t) return 1;  // re-recorded by a later token only once t completed
    hipError_t e = hipEventQuery(x->tok_ev[k]);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return hip_status(e);
}

int mixer_token_wait(mixer_ctx* x, uint64_t t) {
    if (!x || t == 0 || t >= x->tok_next) return MIXER_ERR_ARG;
    const int k = (int)(t % mixer_ctx::NTOK);
    if (x->tok_owner[k] != t) return MIXER_OK;
    return hip_status(hipEventSynchronize(x->tok_ev[k]));
}

// Bounded ingest: the reference's fixed pool of NSLOT = 96 slots (src/rx_host_corr.cu:46,
// 120-121,139-140,154).
int mixer_set_ingest_slots(mixer_ctx* x, int n_slots) {
    if (!x || n_slots < 0) return MIXER_ERR_ARG;
    x->slots_max = n_slots;
    return MIXER_OK;
}

int mixer_sync(mixer_ctx* x) {
    if (!x) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(x->device));
    MIXER_HIP(hipStreamSynchronize(x->s_copy));
    MIXER_HIP(hipStreamSynchronize(x->s_comp));
    MIXER_HIP(hipStreamSynchronize(x->s_pub));
    std::lock_guard<std::mutex> lk(x->mtx);
    poll_pending(x);
    return MIXER_OK;
}

// Output seam: the --dump record (src/rx_host_corr.cu:210-227), taken between integrations.
int mixer_dump(mixer_ctx* x, float* host, int cumulative, uint64_t* integ_index) {
    if (!x || !host) return MIXER_ERR_ARG;
    if (cumulative && !x->d_vcum) return MIXER_ERR_STATE;
    MIXER_HIP(hipSetDevice(x->device));
    uint64_t n;
    {
        std::lock_guard<std::mutex> lk(x->mtx);
        n = x->n_integ.load();
        if (integ_index) *integ_index = n;
        if (n == 0) return MIXER_OK;
        const int b = (int)((n - 1) % 2);
        MIXER_HIP(hipEventRecord(x->integ_ev[b], x->s_comp));
        MIXER_HIP(hipStreamWaitEvent(x->s_pub, x->integ_ev[b], 0));
        MIXER_HIP(hipMemcpyAsync(host, cumulative ? x->d_vcum : x->d_v[b], x->v_bytes,
                                 hipMemcpyDeviceToHost, x->s_pub));
        MIXER_HIP(hipEventRecord(cumulative ? x->pub_done_cum : x->pub_done[b], x->s_pub));
    }
    MIXER_HIP(hipStreamSynchronize(x->s_pub));
    return MIXER_OK;
}

int mixer_read_image(mixer_ctx* x, float* img, float* vmax, float* vmin, uint64_t* integ_index) {
    if (!x) return MIXER_ERR_ARG;
    if (x->cfg.flags & MIXER_CFG_NO_IMAGE) return MIXER_ERR_STATE;
    MIXER_HIP(hipSetDevice(x->device));
    const size_t C = x->cfg.n_chan;
    {
        std::lock_guard<std::mutex> lk(x->mtx);
        const uint64_t n = x->n_integ.load();
        if (integ_index) *integ_index = n;
        if (n == 0) return MIXER_OK;
        const int b = (int)((n - 1) % 2);
        MIXER_HIP(hipEventRecord(x->integ_ev[b], x->s_comp));
        MIXER_HIP(hipStreamWaitEvent(x->s_pub, x->integ_ev[b], 0));
        const float* d = x->d_img[b];
        if (img) MIXER_HIP(hipMemcpyAsync(img, d, x->img_floats * 4, hipMemcpyDeviceToHost, x->s_pub));
        if (vmax) MIXER_HIP(hipMemcpyAsync(vmax, d + x->img_floats, C * 4, hipMemcpyDeviceToHost, x->s_pub));
        if (vmin) MIXER_HIP(hipMemcpyAsync(vmin, d + x->img_floats + C, C * 4, hipMemcpyDeviceToHost, x->s_pub));
        MIXER_HIP(hipEventRecord(x->pub_done[b], x->s_pub));
    }
    MIXER_HIP(hipStreamSynchronize(x->s_pub));
    return MIXER_OK;
}

// The reference's monitor frame: the image of dV, all that was integrated since its previous
// frame (src/rx_host_corr.cu:218-222), from the exact int64 frame grid.
int mixer_read_frame(mixer_ctx* x, float* img, float* vmax, float* vmin, uint64_t* n_in_frame) {
    if (!x) return MIXER_ERR_ARG;
    if (!x->d_Gf) return MIXER_ERR_STATE;
    MIXER_HIP(hipSetDevice(x->device));
    const size_t C = x->cfg.n_chan;
    {
        std::lock_guard<std::mutex> lk(x->mtx);
        const uint64_t n = x->n_integ.load(), nf = n - x->frame_mark;
        if (n_in_frame) *n_in_frame = nf;
        if (nf == 0) return MIXER_OK;
        x->frame_mark = n;
        // on the compute stream, behind the integrations it covers: image the frame grid (and
        // clear it), once the previous frame's copy has read the frame image buffer
        MIXER_HIP(hipStreamWaitEvent(x->s_comp, x->pubf_done, 0));
        float* d = x->d_imgf;
        MIXER_HIP(launch_image(x->s_comp, x->d_Gf, x->d_tw, d, d + x->img_floats, d + x->img_floats + C,
                               x->cfg.n_chan, x->cfg.ng));
        MIXER_HIP(hipEventRecord(x->frame_ev, x->s_comp));
        MIXER_HIP(hipStreamWaitEvent(x->s_pub, x->frame_ev, 0));
        if (img) MIXER_HIP(hipMemcpyAsync(img, d, x->img_floats * 4, hipMemcpyDeviceToHost, x->s_pub));
        if (vmax) MIXER_HIP(hipMemcpyAsync(vmax, d + x->img_floats, C * 4, hipMemcpyDeviceToHost, x->s_pub));
        if (vmin) MIXER_HIP(hipMemcpyAsync(vmin, d + x->img_floats + C, C * 4, hipMemcpyDeviceToHost, x->s_pub));
        MIXER_HIP(hipEventRecord(x->pubf_done, x->s_pub));
    }
    MIXER_HIP(hipStreamSynchronize(x->s_pub));
    return MIXER_OK;
}

size_t mixer_corr32_bytes(const mixer_array_cfg* c) {
    return 24 + 4 * ((size_t)c->n_chan * c->ng * c->ng + 2 * (size_t)c->n_chan);
}

// Corr32 header (src/rx_host_corr.cu:98,104): magic, w, h, nch, then write_seq at offset 16.
int mixer_corr32_init(void* region, const mixer_array_cfg* c) {
    if (!region || !c) return MIXER_ERR_ARG;
    memset(region, 0, mixer_corr32_bytes(c));
    uint32_t* h = (uint32_t*)region;
    h[0] = 0x43523343u;
    h[1] = (uint32_t)c->ng;
    h[2] = (uint32_t)c->ng;
    h[3] = (uint32_t)c->n_chan;
    return MIXER_OK;
}

// The reference stores write_seq = frame+1 with release order after the frame
// (src/rx_host_corr.cu:236) and has no writer flag; here bit 63 is set for the duration of the
// write so a reader that checks it never uses a torn frame.
int mixer_publish_image(mixer_ctx* x, void* region, size_t bytes) {
    if (!x || !region || bytes < mixer_corr32_bytes(&x->cfg)) return MIXER_ERR_ARG;
    uint32_t* h = (uint32_t*)region;
    if (h[0] != 0x43523343u) MIXER_TRY(mixer_corr32_init(region, &x->cfg));
    const size_t C = x->cfg.n_chan;
    std::vector<float> img(x->img_floats + 2 * C);
    uint64_t n = 0;
    if (x->d_Gf)  // dV since the previous frame, as the reference's monitor publishes
        MIXER_TRY(mixer_read_frame(x, img.data(), img.data() + x->img_floats,
                                   img.data() + x->img_floats + C, &n));
    else
        MIXER_TRY(mixer_read_image(x, img.data(), img.data() + x->img_floats,
                                   img.data() + x->img_floats + C, &n));
    if (n == 0) return MIXER_OK;
    uint64_t* ws = (uint64_t*)((uint8_t*)region + 16);
    const uint64_t frame = ++x->frames_published;
    __atomic_store_n(ws, frame | (1ull << 63), __ATOMIC_RELEASE);
    __atomic_thread_fence(__ATOMIC_SEQ_CST);
    memcpy((uint8_t*)region + 24, img.data(), img.size() * sizeof(float));
    __atomic_store_n(ws, frame, __ATOMIC_RELEASE);
    return MIXER_OK;
}

int mixer_get_stats(mixer_ctx* x, mixer_stats* s) {
    if (!x || !s) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(x->device));
    uint32_t w = 0;
    MIXER_HIP(hipMemcpy(&w, x->d_wrap, sizeof(w), hipMemcpyDeviceToHost));
    std::lock_guard<std::mutex> lk(x->mtx);
    poll_pending(x);
    x->st.wrap128 = w;
    *s = x->st;
    return MIXER_OK;
}

int mixer_set_timing(mixer_ctx* x, int enable) {
    if (!x) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(x->device));
    std::lock_guard<std::mutex> lk(x->tmtx);
    MIXER_TRY(resolve_timing(x, true));
    for (int i = 0; i < 4; ++i) { x->t_ms[i] = 0; x->t_n[i] = 0; }
    x->timing = (unsigned)enable & 0xFu;
    return MIXER_OK;
}

int mixer_get_timing(mixer_ctx* x, int stage, double* total_ms, uint64_t* launches) {
    if (!x || stage < 0 || stage > 3) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(x->device));
    std::lock_guard<std::mutex> lk(x->tmtx);
    MIXER_TRY(resolve_timing(x, true));
    if (total_ms) *total_ms = x->t_ms[stage];
    if (launches) *launches = x->t_n[stage];
    return MIXER_OK;
}

}  // extern "C"
