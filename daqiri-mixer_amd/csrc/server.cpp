// server.cpp -- the serving loop (mixer_server_*): arrays admitted under an HBM budget, chunks
// of their heaps correlated as they arrive, integrations integrated + imaged and published to
// pinned host memory, arrival-to-dump latency timed by a completion thread; plus a Poisson load
// generator (mixer_serve_run).
//
// Ordering, per array a (stream a.s) and the shared device->host stream s_d2h:
//   a.s    [corner-turn] correlate ... correlate | wait pub[b] | integrate -> V[b] + grid,
//          image -> img[b] | record integ[b]
//   s_d2h  wait integ[b] | copy img[b] -> pinned host slot b | record pub[b], done event
// The completion thread synchronises the done events in submission order (one stream, so they
// complete in that order) and timestamps each dump.
//
// Chunk coalescing: a chunk that arrives while the array's previous correlator launch is still
// running is held back, and the chunks that follow it contiguously in memory join it, so that
// one longer launch correlates them all (every launch reads and writes the whole int32
// accumulator, which for short chunks of large arrays costs more than the chunk's own input:
// 1024 x 235 with 1440-sample chunks runs 14.5 ms per 11520-sample integration, 8.7 ms in one
// launch). A held run is launched when the next chunk finds the array idle, when it would not
// extend, when the integration completes, and by drain / token_wait. Runs never cross an
// integration. The int32 sums do not depend on how the time range is split.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <queue>
#include <random>
#include <thread>
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

struct SrvArray {
    mixer_array_cfg cfg{};
    XcorrPlan plan{};
    int n_tb_upd = 0, upd_per_integ = 0;
    size_t tb_bytes = 0, v_bytes = 0, img_floats = 0, footprint = 0;
    uint8_t* d_in = nullptr;
    int in_chunks = 0, in_next = 0;
    int8_t* d_tiles = nullptr;
    int32_t* d_acc = nullptr;
    uint32_t* d_wrap = nullptr;
    unsigned long long* d_G = nullptr;
    float* d_v[2] = {nullptr, nullptr};
    float* d_img[2] = {nullptr, nullptr};
    float* h_img[2] = {nullptr, nullptr};  // pinned
    hipEvent_t ev_integ[2] = {nullptr, nullptr};
    hipEvent_t ev_pub[2] = {nullptr, nullptr};
    hipStream_t s = nullptr;
    const float2* tw = nullptr;
    int upd_in_integ = 0;       // updates of the current integration submitted
    int upd_launched = 0;       // ... of which launched
    const uint8_t* run_ptr = nullptr;  // held run: run_tb time blocks from run_ptr
    int run_tb = 0;
    hipEvent_t ev_launch = nullptr;    // after the last correlator launch
    uint64_t tok_launched = 0;         // tokens up to this one have their events recorded
    uint64_t n_integ = 0;  // integrations enqueued
    // token t uses slot t % NTOK; a slot's event is re-recorded only after the previous token's
    // event has completed, so a token whose slot holds a later token is complete
    static constexpr int NTOK = 64;
    hipEvent_t tok[NTOK] = {};
    uint64_t tok_owner[NTOK] = {};
    uint64_t tok_next = 1;
    std::vector<void*> dev;
};

struct Done {
    SrvArray* a;
    uint64_t arrival;
    hipEvent_t ev;
};

size_t chunk_bytes_of(const mixer_array_cfg& c) {
    return (size_t)(c.k_update / c.t_pkt) * c.n_elem * mixer_payload_bytes(&c);
}
int ring_chunks_of(const mixer_array_cfg& c) { return 2 * (c.k_integ / c.k_update); }

uint64_t footprint_of(const mixer_array_cfg& c) {
    const size_t C = c.n_chan, N = c.n_elem, ng = c.ng;
    uint64_t f = (uint64_t)ring_chunks_of(c) * chunk_bytes_of(c);
    if (!plan_fused(c)) f += mixer_tiles_bytes(&c, c.k_update / c.t_pkt);
    f += mixer_acc_bytes(&c);
    f += 2 * C * 2 * N * N * sizeof(float);             // V, double-buffered
    f += 2 * (C * ng * ng + 2 * C) * sizeof(float);      // images, double-buffered
    f += C * ng * ng * 16;                               // int64 image grid
    f += 256;                                            // counter
    return f;
}

}  // namespace

struct mixer_server {
    int device = 0;
    uint64_t budget = 0, used = 0;
    std::vector<hipStream_t> streams;
    hipStream_t s_d2h = nullptr;
    int rr = 0;
    std::vector<SrvArray*> arrays;  // index = array id; nullptr once released
    std::mutex m;
    std::condition_variable cv, cv_done;
    std::deque<Done> q;
    bool stop = false;
    std::thread th;
    std::vector<hipEvent_t> ev_pool;
    std::vector<float> lat_ms;
    uint64_t admitted = 0, rejected = 0, chunks = 0, bytes_in = 0, integrations = 0;
    std::atomic<uint64_t> launches{0};
    bool coalesce = true;
    uint64_t outstanding = 0;
    uint64_t last_done_ns = 0;
};

namespace {

void completion_loop(mixer_server* S) {
    (void)hipSetDevice(S->device);
    for (;;) {
        Done d;
        {
            std::unique_lock<std::mutex> lk(S->m);
            S->cv.wait(lk, [&] { return S->stop || !S->q.empty(); });
            if (S->q.empty()) return;  // stop requested and nothing left
            d = S->q.front();
        }
        const hipError_t e = hipEventSynchronize(d.ev);
        const uint64_t t = now_ns();
        std::lock_guard<std::mutex> lk(S->m);
        S->q.pop_front();
        if (e == hipSuccess) {
            S->lat_ms.push_back(t > d.arrival ? (float)((t - d.arrival) * 1e-6) : 0.0f);
            S->integrations++;
        }
        S->last_done_ns = t;
        S->ev_pool.push_back(d.ev);
        S->outstanding--;
        S->cv_done.notify_all();
    }
}

int take_event(mixer_server* S, hipEvent_t* e) {
    {
        std::lock_guard<std::mutex> lk(S->m);
        if (!S->ev_pool.empty()) {
            *e = S->ev_pool.back();
            S->ev_pool.pop_back();
            return MIXER_OK;
        }
    }
    return hip_status(hipEventCreateWithFlags(e, hipEventDisableTiming));
}

SrvArray* get_array(mixer_server* S, int id) {
    std::lock_guard<std::mutex> lk(S->m);
    if (id < 0 || id >= (int)S->arrays.size()) return nullptr;
    return S->arrays[id];
}

void free_array(SrvArray* a) {
    if (!a) return;
    if (a->s) (void)hipStreamSynchronize(a->s);
    for (int b = 0; b < 2; ++b) {
        if (a->ev_integ[b]) (void)hipEventDestroy(a->ev_integ[b]);
        if (a->ev_pub[b]) (void)hipEventDestroy(a->ev_pub[b]);
        if (a->h_img[b]) (void)hipHostFree(a->h_img[b]);
    }
    for (auto e : a->tok)
        if (e) (void)hipEventDestroy(e);
    if (a->ev_launch) (void)hipEventDestroy(a->ev_launch);
    for (void* p : a->dev) (void)hipFree(p);
    delete a;
}

int dalloc(SrvArray* a, void** p, size_t bytes) {
    MIXER_HIP(hipMalloc(p, bytes));
    a->dev.push_back(*p);
    return MIXER_OK;
}

int build_array(mixer_server* S, const mixer_array_cfg& c, int id, SrvArray* a) {
    a->cfg = c;
    MIXER_TRY(get_plan(c, &a->plan));
    a->n_tb_upd = c.k_update / c.t_pkt;
    a->upd_per_integ = c.k_integ / c.k_update;
    a->tb_bytes = (size_t)c.n_elem * mixer_payload_bytes(&c);
    a->v_bytes = (size_t)c.n_chan * 2 * c.n_elem * c.n_elem * sizeof(float);
    a->img_floats = (size_t)c.n_chan * c.ng * c.ng;
    a->footprint = footprint_of(c);
    a->in_chunks = ring_chunks_of(c);
    const size_t cb = chunk_bytes_of(c);
    MIXER_TRY(dalloc(a, (void**)&a->d_in, cb * a->in_chunks));
    if (!a->plan.fused) MIXER_TRY(dalloc(a, (void**)&a->d_tiles, mixer_tiles_bytes(&c, a->n_tb_upd)));
    MIXER_TRY(dalloc(a, (void**)&a->d_acc, mixer_acc_bytes(&c)));
    MIXER_TRY(dalloc(a, (void**)&a->d_wrap, 256));
    const size_t gbytes = (size_t)c.n_chan * c.ng * c.ng * 16;
    MIXER_TRY(dalloc(a, (void**)&a->d_G, gbytes));
    for (int b = 0; b < 2; ++b) {
        MIXER_TRY(dalloc(a, (void**)&a->d_v[b], a->v_bytes));
        MIXER_TRY(dalloc(a, (void**)&a->d_img[b], (a->img_floats + 2 * c.n_chan) * sizeof(float)));
        MIXER_HIP(hipHostMalloc((void**)&a->h_img[b], (a->img_floats + 2 * c.n_chan) * sizeof(float),
                                hipHostMallocDefault));
        MIXER_HIP(hipEventCreateWithFlags(&a->ev_integ[b], hipEventDisableTiming));
        MIXER_HIP(hipEventCreateWithFlags(&a->ev_pub[b], hipEventDisableTiming));
        MIXER_HIP(hipEventRecord(a->ev_pub[b], S->s_d2h));
    }
    MIXER_HIP(hipEventCreateWithFlags(&a->ev_launch, hipEventDisableTiming));
    MIXER_TRY(get_twiddles(c.ng, &a->tw));
    MIXER_HIP(hipMemset(a->d_wrap, 0, 256));
    MIXER_HIP(hipMemset(a->d_G, 0, gbytes));
    // the input ring starts as seeded int8 noise in [-127, 127] (what mixer_serve_run replays)
    MIXER_HIP(launch_fill_random(nullptr, a->d_in, cb * a->in_chunks, 0x4D495835ull + (uint64_t)id));
    MIXER_HIP(hipDeviceSynchronize());
    return MIXER_OK;
}

// Integrate + image the array's completed integration and queue its dump (submitting thread).
int finish_integration(mixer_server* S, SrvArray* a, uint64_t arrival) {
    const mixer_array_cfg& c = a->cfg;
    const int b = (int)(a->n_integ % 2);
    MIXER_HIP(hipStreamWaitEvent(a->s, a->ev_pub[b], 0));
    MIXER_HIP(launch_integrate(a->s, a->plan, a->d_acc, a->d_v[b], nullptr, a->d_G, c.n_elem,
                               c.n_chan, c.grid_x, c.ng));
    float* img = a->d_img[b];
    MIXER_HIP(launch_image(a->s, a->d_G, a->tw, img, img + a->img_floats, img + a->img_floats + c.n_chan,
                           c.n_chan, c.ng));
    MIXER_HIP(hipEventRecord(a->ev_integ[b], a->s));
    MIXER_HIP(hipStreamWaitEvent(S->s_d2h, a->ev_integ[b], 0));
    MIXER_HIP(hipMemcpyAsync(a->h_img[b], img, (a->img_floats + 2 * c.n_chan) * sizeof(float),
                             hipMemcpyDeviceToHost, S->s_d2h));
    MIXER_HIP(hipEventRecord(a->ev_pub[b], S->s_d2h));
    hipEvent_t e;
    MIXER_TRY(take_event(S, &e));
    MIXER_HIP(hipEventRecord(e, S->s_d2h));
    {
        std::lock_guard<std::mutex> lk(S->m);
        S->q.push_back({a, arrival, e});
        S->outstanding++;
    }
    S->cv.notify_one();
    a->n_integ++;
    a->upd_in_integ = 0;
    a->upd_launched = 0;
    return MIXER_OK;
}

// Correlate the held run (if any) on the array's stream and record the events of the tokens
// submitted so far.
int launch_run(mixer_server* S, SrvArray* a) {
    const mixer_array_cfg& c = a->cfg;
    uint32_t* const wrap = (c.flags & MIXER_CFG_NO_WRAP_CHECK) ? nullptr : a->d_wrap;
    if (a->run_tb > 0) {
        if (a->plan.fused) {
            MIXER_HIP(launch_xcorr(a->s, a->plan, xsrc_heaps(a->run_ptr, c.n_elem, c.n_chan),
                                   a->d_acc, wrap, c.n_chan, a->run_tb,
                                   a->upd_launched > 0 ? 1 : 0));
            a->upd_launched += a->run_tb / a->n_tb_upd;
            S->launches++;
        } else {  // the tile buffer holds one update: one corner-turn + correlator per update
            for (int k = 0; k < a->run_tb; k += a->n_tb_upd) {
                const uint8_t* src = a->run_ptr + (size_t)k * a->tb_bytes;
                MIXER_HIP(launch_corner_turn(a->s, src, a->d_tiles, wrap, c.n_elem, c.n_chan,
                                             a->n_tb_upd, 0, c.n_chan, a->plan.ct_ch));
                MIXER_HIP(launch_xcorr(a->s, a->plan, xsrc_tiles(a->d_tiles, c.n_elem, a->n_tb_upd),
                                       a->d_acc, nullptr, c.n_chan, a->n_tb_upd,
                                       a->upd_launched > 0 ? 1 : 0));
                a->upd_launched++;
                S->launches++;
            }
        }
        MIXER_HIP(hipEventRecord(a->ev_launch, a->s));
        a->run_tb = 0;
        a->run_ptr = nullptr;
    }
    for (uint64_t t = a->tok_launched + 1; t < a->tok_next; ++t) {
        const int k = (int)(t % SrvArray::NTOK);
        hipEvent_t& e = a->tok[k];
        if (!e) MIXER_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventDisableSystemFence));
        else MIXER_HIP(hipEventSynchronize(e));  // the slot's previous token completes first
        MIXER_HIP(hipEventRecord(e, a->s));
        a->tok_owner[k] = t;
    }
    a->tok_launched = a->tok_next - 1;
    return MIXER_OK;
}

int submit(mixer_server* S, SrvArray* a, const uint8_t* d, int n_tb, uint64_t arrival,
           uint64_t* token) {
    for (int k = 0; k < n_tb; k += a->n_tb_upd) {
        const uint8_t* src = d + (size_t)k * a->tb_bytes;
        if (a->run_tb > 0 && src != a->run_ptr + (size_t)a->run_tb * a->tb_bytes)
            MIXER_TRY(launch_run(S, a));  // not contiguous with the held run
        if (a->run_tb == 0) a->run_ptr = src;
        a->run_tb += a->n_tb_upd;
        if (++a->upd_in_integ == a->upd_per_integ) {
            MIXER_TRY(launch_run(S, a));
            MIXER_TRY(finish_integration(S, a, arrival));
        }
    }
    const uint64_t t = a->tok_next++;
    // hold the run while the array's previous launch is still running, and while its tokens
    // fit the event ring (an older token counts as recycled, i.e. complete)
    const bool busy = S->coalesce && a->run_tb > 0 &&
                      a->tok_next - 1 - a->tok_launched < (uint64_t)SrvArray::NTOK / 2 &&
                      hipEventQuery(a->ev_launch) == hipErrorNotReady;
    if (!busy) MIXER_TRY(launch_run(S, a));
    if (token) *token = t;
    std::lock_guard<std::mutex> lk(S->m);
    S->chunks++;
    S->bytes_in += (uint64_t)n_tb * a->tb_bytes;
    return MIXER_OK;
}

// Launch every held run (submitting thread).
int flush_all(mixer_server* S) {
    std::vector<SrvArray*> arr;
    {
        std::lock_guard<std::mutex> lk(S->m);
        for (SrvArray* a : S->arrays)
            if (a) arr.push_back(a);
    }
    for (SrvArray* a : arr) MIXER_TRY(launch_run(S, a));
    return MIXER_OK;
}

int drain(mixer_server* S) {
    MIXER_TRY(flush_all(S));
    for (size_t i = 0; i < S->streams.size(); ++i) MIXER_HIP(hipStreamSynchronize(S->streams[i]));
    MIXER_HIP(hipStreamSynchronize(S->s_d2h));
    std::unique_lock<std::mutex> lk(S->m);
    S->cv_done.wait(lk, [&] { return S->outstanding == 0; });
    return MIXER_OK;
}

double pct(std::vector<float> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const size_t k = std::min(v.size() - 1, (size_t)(p / 100.0 * (double)(v.size() - 1) + 0.5));
    return v[k];
}

void destroy_server(mixer_server* S) {
    if (!S) return;
    (void)hipSetDevice(S->device);
    if (S->th.joinable()) {
        (void)drain(S);
        {
            std::lock_guard<std::mutex> lk(S->m);
            S->stop = true;
        }
        S->cv.notify_all();
        S->th.join();
    }
    for (SrvArray* a : S->arrays) free_array(a);
    for (auto e : S->ev_pool) (void)hipEventDestroy(e);
    for (auto s : S->streams) (void)hipStreamDestroy(s);
    if (S->s_d2h) (void)hipStreamDestroy(S->s_d2h);
    delete S;
}

}  // namespace

extern "C" {

uint64_t mixer_server_footprint(const mixer_array_cfg* c) {
    if (mixer_validate_cfg(c)) return 0;
    return footprint_of(*c);
}

int mixer_server_create(int device, uint64_t hbm_budget, int n_streams, mixer_server** out) {
    if (!out || n_streams < 0 || n_streams > 16) return MIXER_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return MIXER_ERR_NODEV;
    mixer_server* S = new mixer_server();
    S->device = device;
    S->budget = hbm_budget;
    int st = hip_status(hipSetDevice(device));
    for (int i = 0; !st && i < (n_streams ? n_streams : 1); ++i) {
        hipStream_t s = nullptr;
        st = hip_status(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        if (!st) S->streams.push_back(s);
    }
    if (!st) st = hip_status(hipStreamCreateWithFlags(&S->s_d2h, hipStreamNonBlocking));
    if (st) {
        destroy_server(S);
        return st;
    }
    S->th = std::thread(completion_loop, S);
    *out = S;
    return MIXER_OK;
}

int mixer_server_destroy(mixer_server* S) {
    destroy_server(S);
    return MIXER_OK;
}

int mixer_server_admit(mixer_server* S, const mixer_array_cfg* cfg, int* array_id) {
    if (!S || !cfg || !array_id) return MIXER_ERR_ARG;
    MIXER_TRY(mixer_validate_cfg(cfg));
    const uint64_t need = footprint_of(*cfg);
    {
        std::lock_guard<std::mutex> lk(S->m);
        if (S->used + need > S->budget) {
            S->rejected++;
            return MIXER_ERR_BUDGET;
        }
        S->used += need;  // reserved while building
    }
    MIXER_HIP(hipSetDevice(S->device));
    SrvArray* a = new SrvArray();
    int id;
    {
        std::lock_guard<std::mutex> lk(S->m);
        id = (int)S->arrays.size();
    }
    a->s = S->streams[S->rr++ % S->streams.size()];
    const int st = build_array(S, *cfg, id, a);
    std::lock_guard<std::mutex> lk(S->m);
    if (st) {
        S->used -= need;
        free_array(a);
        return st;
    }
    S->arrays.push_back(a);
    S->admitted++;
    *array_id = id;
    return MIXER_OK;
}

int mixer_server_release(mixer_server* S, int id) {
    if (!S) return MIXER_ERR_ARG;
    SrvArray* a = get_array(S, id);
    if (!a) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(S->device));
    MIXER_TRY(drain(S));  // no queued dump may still refer to it
    std::lock_guard<std::mutex> lk(S->m);
    S->used -= a->footprint;
    S->arrays[id] = nullptr;
    free_array(a);
    return MIXER_OK;
}

int mixer_server_input(mixer_server* S, int id, uint8_t** d_ring, int* n_chunks, size_t* chunk_bytes) {
    SrvArray* a = S ? get_array(S, id) : nullptr;
    if (!a) return MIXER_ERR_ARG;
    if (d_ring) *d_ring = a->d_in;
    if (n_chunks) *n_chunks = a->in_chunks;
    if (chunk_bytes) *chunk_bytes = chunk_bytes_of(a->cfg);
    return MIXER_OK;
}

int mixer_server_submit(mixer_server* S, int id, const uint8_t* d, int n_tb, uint64_t arrival_ns,
                        uint64_t* token) {
    SrvArray* a = S ? get_array(S, id) : nullptr;
    if (!a || !d || n_tb <= 0 || n_tb % a->n_tb_upd) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(S->device));
    return submit(S, a, d, n_tb, arrival_ns, token);
}

int mixer_server_token_wait(mixer_server* S, int id, uint64_t t) {
    SrvArray* a = S ? get_array(S, id) : nullptr;
    if (!a || t == 0 || t >= a->tok_next) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(S->device));
    if (t > a->tok_launched) MIXER_TRY(launch_run(S, a));  // still held back: launch it now
    const int k = (int)(t % SrvArray::NTOK);
    if (a->tok_owner[k] != t) return MIXER_OK;  // re-recorded by a later token once t completed
    return hip_status(hipEventSynchronize(a->tok[k]));
}

int mixer_server_set_coalescing(mixer_server* S, int enable) {
    if (!S) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(S->device));
    MIXER_TRY(flush_all(S));
    S->coalesce = enable != 0;
    return MIXER_OK;
}

int mixer_server_drain(mixer_server* S) {
    if (!S) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(S->device));
    return drain(S);
}

int mixer_server_get_stats(mixer_server* S, mixer_server_stats* st) {
    if (!S || !st) return MIXER_ERR_ARG;
    std::vector<float> lat;
    {
        std::lock_guard<std::mutex> lk(S->m);
        st->hbm_budget = S->budget;
        st->hbm_used = S->used;
        st->admitted = S->admitted;
        st->rejected = S->rejected;
        st->chunks = S->chunks;
        st->bytes_in = S->bytes_in;
        st->integrations = S->integrations;
        st->launches = S->launches;
        lat = S->lat_ms;
    }
    st->lat_n = lat.size();
    double sum = 0;
    for (float x : lat) sum += x;
    st->lat_mean_ms = lat.empty() ? 0.0 : sum / lat.size();
    st->lat_p50_ms = pct(lat, 50);
    st->lat_p99_ms = pct(lat, 99);
    st->lat_max_ms = lat.empty() ? 0.0 : *std::max_element(lat.begin(), lat.end());
    return MIXER_OK;
}

void mixer_server_reset_latency(mixer_server* S) {
    if (!S) return;
    std::lock_guard<std::mutex> lk(S->m);
    S->lat_ms.clear();
}

int mixer_server_read_image(mixer_server* S, int id, float* img, float* vmax, float* vmin,
                            uint64_t* integ_index) {
    SrvArray* a = S ? get_array(S, id) : nullptr;
    if (!a) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(S->device));
    MIXER_HIP(hipStreamSynchronize(a->s));
    MIXER_HIP(hipStreamSynchronize(S->s_d2h));
    const uint64_t n = a->n_integ;
    if (integ_index) *integ_index = n;
    if (n == 0) return MIXER_OK;
    const float* h = a->h_img[(n - 1) % 2];
    const size_t C = a->cfg.n_chan;
    if (img) memcpy(img, h, a->img_floats * sizeof(float));
    if (vmax) memcpy(vmax, h + a->img_floats, C * sizeof(float));
    if (vmin) memcpy(vmin, h + a->img_floats + C, C * sizeof(float));
    return MIXER_OK;
}

int mixer_server_dump(mixer_server* S, int id, float* host_v, uint64_t* integ_index) {
    SrvArray* a = S ? get_array(S, id) : nullptr;
    if (!a || !host_v) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(S->device));
    MIXER_HIP(hipStreamSynchronize(a->s));
    const uint64_t n = a->n_integ;
    if (integ_index) *integ_index = n;
    if (n == 0) return MIXER_OK;
    return hip_status(hipMemcpy(host_v, a->d_v[(n - 1) % 2], a->v_bytes, hipMemcpyDeviceToHost));
}

int mixer_serve_run(mixer_server* S, const mixer_serve_spec* spec, mixer_serve_report* rep) {
    if (!S || !spec || !rep || spec->duration_s <= 0 || spec->aggregate_gbps <= 0) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(S->device));
    std::vector<SrvArray*> arr;
    std::vector<int> ids;
    {
        std::lock_guard<std::mutex> lk(S->m);
        for (size_t i = 0; i < S->arrays.size(); ++i)
            if (S->arrays[i]) { arr.push_back(S->arrays[i]); ids.push_back((int)i); }
    }
    if (arr.empty()) return MIXER_ERR_STATE;
    MIXER_TRY(drain(S));
    // shared sample clock: fs samples/s on every array, sum of bytes per sample = aggregate rate
    double bytes_per_sample = 0;
    for (SrvArray* a : arr) bytes_per_sample += (double)a->cfg.n_elem * mixer_payload_bytes(&a->cfg) / a->cfg.t_pkt;
    const double fs = spec->aggregate_gbps * 1e9 / bytes_per_sample;
    std::mt19937_64 rng(spec->seed);
    using Ev = std::pair<uint64_t, int>;  // (arrival ns, array index)
    std::priority_queue<Ev, std::vector<Ev>, std::greater<Ev>> pq;
    std::vector<std::exponential_distribution<double>> gap;
    const uint64_t t0 = now_ns() + 2000000;  // start 2 ms from now
    for (size_t i = 0; i < arr.size(); ++i) {
        const double mean_s = arr[i]->cfg.k_update / fs;  // one chunk = k_update samples
        gap.emplace_back(1.0 / mean_s);
        pq.push({t0 + (uint64_t)(gap[i](rng) * 1e9), (int)i});
    }
    uint64_t chunks0, integ0, bytes0;
    {
        std::lock_guard<std::mutex> lk(S->m);
        S->lat_ms.clear();
        chunks0 = S->chunks;
        integ0 = S->integrations;
        bytes0 = S->bytes_in;
    }
    const uint64_t t_end = t0 + (uint64_t)(spec->duration_s * 1e9);
    uint64_t late = 0;
    int st = MIXER_OK;
    while (!pq.empty()) {
        const Ev ev = pq.top();
        pq.pop();
        if (ev.first >= t_end) continue;  // past the window: this array is finished
        SrvArray* a = arr[ev.second];
        for (;;) {  // wait for the arrival time: sleep while far, then spin
            const uint64_t t = now_ns();
            if (t >= ev.first) break;
            if (ev.first - t > 300000) std::this_thread::sleep_for(std::chrono::nanoseconds(ev.first - t - 200000));
        }
        if (now_ns() > ev.first + 1000000) late++;
        const size_t cb = chunk_bytes_of(a->cfg);
        const uint8_t* src = a->d_in + (size_t)(a->in_next++ % a->in_chunks) * cb;
        if ((st = submit(S, a, src, a->n_tb_upd, ev.first, nullptr))) break;
        pq.push({ev.first + (uint64_t)(gap[ev.second](rng) * 1e9), ev.second});
    }
    const int st2 = drain(S);
    const uint64_t t_fin = now_ns();
    if (st) return st;
    MIXER_TRY(st2);
    std::vector<float> lat;
    {
        std::lock_guard<std::mutex> lk(S->m);
        lat = S->lat_ms;
        rep->chunks = S->chunks - chunks0;
        rep->integrations = S->integrations - integ0;
        const double wall = (double)(t_fin - t0) * 1e-9;
        rep->wall_s = wall;
        rep->achieved_gbps = (double)(S->bytes_in - bytes0) / wall / 1e9;
    }
    rep->late_submits = late;
    rep->offered_gbps = spec->aggregate_gbps;
    rep->lat_p50_ms = pct(lat, 50);
    rep->lat_p99_ms = pct(lat, 99);
    rep->lat_max_ms = lat.empty() ? 0.0 : *std::max_element(lat.begin(), lat.end());
    return MIXER_OK;
}

}  // extern "C"
