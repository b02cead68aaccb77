// batch.cpp -- batched correlation of many independent arrays (mixer_batch_*): one corner-turn
// launch, one persistent correlator launch over a cost-ordered work list, one integrate +
// gridding launch and one image launch per run, captured once as a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.h"
#include "kernels.h"
#include "mixer.h"

using namespace mixer;

struct mixer_batch {
    int device = 0;
    std::vector<mixer_array_cfg> cfg;
    mixer_batch_tuning tun{};
    std::vector<BatchJobDev> jobs;
    std::vector<int> slot_cap;  // accumulator capacity (time slots) per job
    BatchJobDev* d_jobs = nullptr;
    bool jobs_dirty = false;    // host job table changed but not yet uploaded to d_jobs
    int4* d_items = nullptr;
    int items_cap = 0, n_items = 0, n_wgs = 0;
    int* d_counter = nullptr;  // one per XCD queue: [0, 8) fold items, [8, 16) BLOCKED items
    BatchQueues queues{}, queues_blk{};
    int n_fold_items = 0, n_blk_items = 0;
    uint32_t* d_wrap = nullptr;
    int ct_total = 0, int_total = 0, img_total = 0, max_ng = 8;
    std::vector<void*> allocs;
    hipStream_t s = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool have_inputs = false;
    uint64_t payload = 0, exec_ops = 0, dev_bytes = 0;
};

namespace {

int balloc(mixer_batch* b, void** p, size_t bytes) {
    MIXER_HIP(hipMalloc(p, bytes));
    b->allocs.push_back(*p);
    b->dev_bytes += bytes;
    return MIXER_OK;
}

int items_per_slot(const mixer_array_cfg& c, int pack) {
    const int nb = c.n_elem / 32;
    if (nb > 8) return (int)sblocks_for(nb).size() * c.n_chan;
    const int g = pack == 1 ? 1 : batch_fold_groups(nb);
    return (c.n_chan + g - 1) / g;
}

int slots_for(const mixer_array_cfg& c, int pack, int tb_target) {
    const int n_tb = c.k_integ / c.t_pkt;
    if (items_per_slot(c, pack) >= 512) return 1;
    int s = (n_tb + tb_target - 1) / tb_target;
    return std::max(1, std::min({s, 8, n_tb}));
}

void destroy_graph(mixer_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// staging 0 (automatic): every array straight from the heaps. With the per-XCD queues the
// super-blocks of a channel share their element blocks in one L2, and the 64-array batch runs
// 108 ms against 121 ms with corner-turned tiles for the arrays above 256 elements (whose
// corner-turn launch alone took 25 ms).
bool job_fused(const mixer_batch* b, int nb) {
    (void)nb;
    return b->tun.staging != 2;
}

int enqueue(mixer_batch* b) {
    MIXER_HIP(launch_batch_reset(b->s, b->d_counter));
    if (b->ct_total > 0)
        MIXER_HIP(launch_batch_corner_turn(b->s, b->d_jobs, (int)b->jobs.size(), b->ct_total));
    if (b->n_blk_items)
        MIXER_HIP(launch_batch_xcorr_blk(b->s, b->d_jobs, b->d_items, b->queues_blk, b->d_counter + 8,
                                         b->n_wgs));
    if (b->n_fold_items)
        MIXER_HIP(launch_batch_xcorr(b->s, b->d_jobs, b->d_items, b->queues, b->d_counter, b->n_wgs,
                                     b->tun.mfma == 16 ? 16 : 32));
    MIXER_HIP(launch_batch_integrate(b->s, b->d_jobs, (int)b->jobs.size(), b->int_total, b->max_ng));
    MIXER_HIP(launch_batch_image(b->s, b->d_jobs, (int)b->jobs.size(), b->img_total, b->max_ng));
    return MIXER_OK;
}

// (Re)build the work list, the time split of each job and the graph for the current tuning.
int plan(mixer_batch* b) {
    MIXER_HIP(hipStreamSynchronize(b->s));
    destroy_graph(b);
    const int pack = b->tun.pack, tbt = b->tun.tb_target > 0 ? b->tun.tb_target : 128;
    // groups of items that read the same data: the super-blocks of one (job, channel, slot), or
    // the fold items of 8 adjacent channels (two 128-byte heap lines)
    struct Group { std::vector<int4> w; double cost = 0; };
    std::vector<Group> groups, groups_blk;  // fold jobs / BLOCKED jobs (separate launches)
    b->exec_ops = 0;
    for (size_t j = 0; j < b->jobs.size(); ++j) {
        const mixer_array_cfg& c = b->cfg[j];
        BatchJobDev& J = b->jobs[j];
        const int nb = J.nb, n_tb = J.n_tb;
        J.n_slots = std::min(slots_for(c, pack, tbt), b->slot_cap[j]);
        J.ch_per_item = pack == 1 ? 1 : batch_fold_groups(nb);
        const int NT = nb * (nb + 1) / 2;
        b->exec_ops += (uint64_t)NT * 1024 * 4 * 2 * (uint64_t)n_tb * 32 * c.n_chan;
        for (int sl = 0; sl < J.n_slots; ++sl) {
            const double kb = (double)n_tb / J.n_slots;
            if (nb <= 8) {
                const int g = pack == 1 ? 1 : batch_fold_groups(nb);
                for (int c0 = 0; c0 < c.n_chan; c0 += g) {
                    if (c0 == 0 || c0 / 8 != (c0 - g) / 8) groups.emplace_back();
                    groups.back().w.push_back(make_int4((int)j, c0, 0, sl));
                    groups.back().cost += kb * NT * std::min(g, c.n_chan - c0);
                }
            } else {
                const auto sb = sblocks_for(nb);
                for (int ch = 0; ch < c.n_chan; ++ch) {
                    groups_blk.emplace_back();
                    for (size_t q = 0; q < sb.size(); ++q) {
                        const int tiles = sb[q].second < 0 ? 36 : sb[q].second + 8 <= sb[q].first ? 32 : 10;
                        groups_blk.back().w.push_back(make_int4((int)j, ch, (int)q, sl));
                        groups_blk.back().cost += kb * tiles;
                    }
                }
            }
        }
    }
    // largest groups first, each to the least loaded XCD queue; the BLOCKED queues, then the
    // fold queues, in one item list
    std::vector<int4> h;
    auto build = [&](std::vector<Group>& gs, BatchQueues& bq) {
        std::stable_sort(gs.begin(), gs.end(), [](const Group& x, const Group& y) { return x.cost > y.cost; });
        std::vector<int4> queue[8];
        double load[8] = {0};
        for (const Group& g : gs) {
            const int x = (int)(std::min_element(load, load + 8) - load);
            load[x] += g.cost;
            queue[x].insert(queue[x].end(), g.w.begin(), g.w.end());
        }
        const size_t h0 = h.size();
        for (int x = 0; x < 8; ++x) {
            bq.start[x] = (int)h.size();
            bq.size[x] = (int)queue[x].size();
            h.insert(h.end(), queue[x].begin(), queue[x].end());
        }
        return (int)(h.size() - h0);
    };
    b->n_blk_items = build(groups_blk, b->queues_blk);
    b->n_fold_items = build(groups, b->queues);
    b->n_items = (int)h.size();
    if (b->n_items > b->items_cap) {
        if (b->d_items) {
            (void)hipFree(b->d_items);
            b->allocs.erase(std::find(b->allocs.begin(), b->allocs.end(), (void*)b->d_items));
        }
        b->items_cap = b->n_items;
        MIXER_TRY(balloc(b, (void**)&b->d_items, sizeof(int4) * b->items_cap));
    }
    MIXER_HIP(hipMemcpy(b->d_items, h.data(), sizeof(int4) * h.size(), hipMemcpyHostToDevice));
    MIXER_HIP(hipMemcpy(b->d_jobs, b->jobs.data(), sizeof(BatchJobDev) * b->jobs.size(),
                        hipMemcpyHostToDevice));
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess) cus = prop.multiProcessorCount;
    b->n_wgs = std::max(1, std::min(b->n_items, cus));
    if (b->tun.use_graph) {
        MIXER_HIP(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
        const int st = enqueue(b);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->s, &g);
        if (st) { if (g) (void)hipGraphDestroy(g); return st; }
        MIXER_HIP(e);
        b->graph = g;
        MIXER_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    return MIXER_OK;
}

void destroy_batch(mixer_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->s) (void)hipStreamSynchronize(b->s);
    destroy_graph(b);
    for (void* p : b->allocs) (void)hipFree(p);
    if (b->s) (void)hipStreamDestroy(b->s);
    delete b;
}

int create_batch(const mixer_array_cfg* cfgs, int n_jobs, int device, const mixer_batch_tuning* t,
                 mixer_batch* b) {
    b->device = device;
    MIXER_HIP(hipSetDevice(device));
    if (t && t->mfma != 0 && t->mfma != 16 && t->mfma != 32) return MIXER_ERR_ARG;
    if (t) b->tun = *t;
    else mixer_batch_default_tuning(&b->tun);
    MIXER_HIP(hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking));
    MIXER_TRY(balloc(b, (void**)&b->d_counter, 16 * sizeof(int)));
    MIXER_TRY(balloc(b, (void**)&b->d_wrap, sizeof(uint32_t)));
    MIXER_HIP(hipMemset(b->d_wrap, 0, sizeof(uint32_t)));
    MIXER_TRY(balloc(b, (void**)&b->d_jobs, sizeof(BatchJobDev) * n_jobs));
    int ct = 0, in = 0, im = 0;
    for (int j = 0; j < n_jobs; ++j) {
        mixer_array_cfg c = cfgs[j];
        c.k_update = c.k_integ;
        MIXER_TRY(mixer_validate_cfg(&c));
        b->cfg.push_back(c);
        BatchJobDev J;
        memset(&J, 0, sizeof(J));
        const int nb = c.n_elem / 32, NT = nb * (nb + 1) / 2, n_tb = c.k_integ / c.t_pkt;
        J.n_elem = c.n_elem; J.n_chan = c.n_chan; J.n_tb = n_tb; J.grid_x = c.grid_x;
        J.ng = c.ng; J.nb = nb; J.wrap = b->d_wrap;
        // accumulator capacity: the most time slots any tuning candidate (tb_target >= 64) uses
        const int cap = std::min(slots_for(c, 0, 64), 8);
        b->slot_cap.push_back(cap);
        J.fused = job_fused(b, nb) ? 1 : 0;
        if (!J.fused) MIXER_TRY(balloc(b, (void**)&J.tiles, mixer_tiles_bytes(&c, n_tb)));
        MIXER_TRY(balloc(b, (void**)&J.acc, (size_t)cap * c.n_chan * NT * 2048 * sizeof(int32_t)));
        MIXER_TRY(balloc(b, (void**)&J.v, (size_t)c.n_chan * 2 * c.n_elem * c.n_elem * sizeof(float)));
        const size_t gbytes = (size_t)c.n_chan * c.ng * c.ng * 16;
        MIXER_TRY(balloc(b, (void**)&J.G, gbytes));
        MIXER_HIP(hipMemset(J.G, 0, gbytes));
        MIXER_TRY(balloc(b, (void**)&J.img, ((size_t)c.n_chan * c.ng * c.ng + 2 * c.n_chan) * sizeof(float)));
        MIXER_TRY(get_twiddles(c.ng, &J.tw));
        XcorrPlan p;
        MIXER_TRY(get_plan(c, &p));
        if (nb > 8) {
            J.sblocks = (const int2*)p.d_tasks;
            J.n_sb = p.n_tasks;
        }
        J.segs = (const int4*)p.d_segs;
        J.n_segs = p.n_segs;
        J.ct_wg0 = ct;  // corner-turn workgroups only for jobs correlated from tiles
        if (!J.fused) ct += n_tb * nb * ((c.n_chan + 7) / 8);
        J.int_wg0 = in; in += p.n_segs * c.n_chan;
        J.img_wg0 = im; im += c.n_chan;
        b->max_ng = std::max(b->max_ng, c.ng);
        b->payload += (uint64_t)n_tb * c.n_elem * mixer_payload_bytes(&c);
        b->jobs.push_back(J);
    }
    b->ct_total = ct; b->int_total = in; b->img_total = im;
    MIXER_HIP(set_batch_lds_limits(b->max_ng));
    MIXER_HIP(hipDeviceSynchronize());
    return plan(b);
}

}  // namespace

extern "C" {

void mixer_batch_default_tuning(mixer_batch_tuning* t) {
    if (!t) return;
    t->pack = 0;
    t->tb_target = 128;
    t->use_graph = 1;
    t->staging = 0;
    t->mfma = 0;
}

int mixer_batch_create(const mixer_array_cfg* cfgs, int n_jobs, int device,
                       const mixer_batch_tuning* tuning, mixer_batch** out) {
    if (!cfgs || n_jobs <= 0 || !out) return MIXER_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return MIXER_ERR_NODEV;
    mixer_batch* b = new mixer_batch();
    const int st = create_batch(cfgs, n_jobs, device, tuning, b);
    if (st) {
        destroy_batch(b);
        return st;
    }
    *out = b;
    return MIXER_OK;
}

int mixer_batch_destroy(mixer_batch* b) {
    destroy_batch(b);
    return MIXER_OK;
}

int mixer_batch_run(mixer_batch* b, const uint8_t* const* d_heaps) {
    if (!b || !d_heaps) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(b->device));
    // validate every pointer before touching the job table, so a refused call leaves the host
    // table equal to the device copy
    for (size_t j = 0; j < b->jobs.size(); ++j)
        if (!d_heaps[j]) return MIXER_ERR_ARG;
    bool changed = b->jobs_dirty;
    for (size_t j = 0; j < b->jobs.size(); ++j)
        if (b->jobs[j].heaps != d_heaps[j]) { b->jobs[j].heaps = d_heaps[j]; changed = true; }
    if (changed) {
        // the table is read by the graph's kernels at run time; update it in stream order. It
        // stays dirty until the upload has succeeded.
        b->jobs_dirty = true;
        MIXER_HIP(hipStreamSynchronize(b->s));
        MIXER_HIP(hipMemcpy(b->d_jobs, b->jobs.data(), sizeof(BatchJobDev) * b->jobs.size(),
                            hipMemcpyHostToDevice));
        b->jobs_dirty = false;
    }
    b->have_inputs = true;
    if (b->exec) return hip_status(hipGraphLaunch(b->exec, b->s));
    return enqueue(b);
}

int mixer_batch_sync(mixer_batch* b) {
    if (!b) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(b->device));
    return hip_status(hipStreamSynchronize(b->s));
}

void* mixer_batch_stream(mixer_batch* b) { return b ? (void*)b->s : nullptr; }

int mixer_batch_get_info(mixer_batch* b, mixer_batch_info* info) {
    if (!b || !info) return MIXER_ERR_ARG;
    info->n_jobs = (int)b->jobs.size();
    info->n_items = b->n_items;
    info->n_wgs = b->n_wgs;
    info->graph = b->exec ? 1 : 0;
    info->payload_bytes = b->payload;
    info->exec_ops = b->exec_ops;
    info->device_bytes = b->dev_bytes;
    return MIXER_OK;
}

int mixer_batch_dump(mixer_batch* b, int job, float* host_v) {
    if (!b || !host_v || job < 0 || job >= (int)b->jobs.size()) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(b->device));
    MIXER_HIP(hipStreamSynchronize(b->s));
    const mixer_array_cfg& c = b->cfg[job];
    return hip_status(hipMemcpy(host_v, b->jobs[job].v,
                                (size_t)c.n_chan * 2 * c.n_elem * c.n_elem * sizeof(float),
                                hipMemcpyDeviceToHost));
}

int mixer_batch_read_image(mixer_batch* b, int job, float* img, float* vmax, float* vmin) {
    if (!b || job < 0 || job >= (int)b->jobs.size()) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(b->device));
    MIXER_HIP(hipStreamSynchronize(b->s));
    const mixer_array_cfg& c = b->cfg[job];
    const size_t n = (size_t)c.n_chan * c.ng * c.ng;
    const float* d = b->jobs[job].img;
    if (img) MIXER_HIP(hipMemcpy(img, d, n * 4, hipMemcpyDeviceToHost));
    if (vmax) MIXER_HIP(hipMemcpy(vmax, d + n, c.n_chan * 4, hipMemcpyDeviceToHost));
    if (vmin) MIXER_HIP(hipMemcpy(vmin, d + n + c.n_chan, c.n_chan * 4, hipMemcpyDeviceToHost));
    return MIXER_OK;
}

int mixer_batch_acc(mixer_batch* b, int job, const int32_t** d_acc, int* n_slots) {
    if (!b || job < 0 || job >= (int)b->jobs.size()) return MIXER_ERR_ARG;
    if (d_acc) *d_acc = b->jobs[job].acc;
    if (n_slots) *n_slots = b->jobs[job].n_slots;
    return MIXER_OK;
}

// Times every candidate as the median over 3 blocks of reps back-to-back runs (after one warm-up
// run), as a caller that replays the batch sees it (consecutive runs overlap at their ends, which
// isolated runs do not show), and moves off the starting tuning only for a candidate at least 2 %
// faster than the best so far, so run-to-run noise cannot leave the batch slower than it started.
int mixer_batch_autotune(mixer_batch* b, int reps, mixer_batch_tuning* best) {
    if (!b || !best || reps <= 0) return MIXER_ERR_ARG;
    if (!b->have_inputs) return MIXER_ERR_STATE;
    MIXER_HIP(hipSetDevice(b->device));
    hipEvent_t e0, e1;
    MIXER_HIP(hipEventCreate(&e0));
    MIXER_HIP(hipEventCreate(&e1));
    const mixer_batch_tuning start = b->tun;
    auto median_ms = [&](float* med) -> int {
        MIXER_TRY(plan(b));
        auto run = [&]() { return b->exec ? hip_status(hipGraphLaunch(b->exec, b->s)) : enqueue(b); };
        MIXER_TRY(run());  // warm-up
        float t[3];
        for (int k = 0; k < 3; ++k) {
            MIXER_HIP(hipEventRecord(e0, b->s));
            for (int r = 0; r < reps; ++r) MIXER_TRY(run());
            MIXER_HIP(hipEventRecord(e1, b->s));
            MIXER_HIP(hipEventSynchronize(e1));
            MIXER_HIP(hipEventElapsedTime(&t[k], e0, e1));
            t[k] /= (float)reps;
        }
        std::sort(t, t + 3);
        *med = t[1];
        return MIXER_OK;
    };
    mixer_batch_tuning win = start;
    float best_ms = 0;
    int st = median_ms(&best_ms);
    const int start_mfma = start.mfma == 16 ? 16 : 32;
    for (int mfma : {32, 16}) {
        for (int pack : {0, 1}) {
            for (int tbt : {64, 128, 256, 512}) {
                if (st) break;
                if (pack == start.pack && tbt == start.tb_target && mfma == start_mfma) continue;
                b->tun.pack = pack;
                b->tun.tb_target = tbt;
                b->tun.use_graph = start.use_graph;
                b->tun.mfma = mfma;
                float ms = 0;
                if ((st = median_ms(&ms))) break;
                if (ms < 0.98f * best_ms) { best_ms = ms; win = b->tun; }
            }
        }
    }
    // The GPU's clock drifts over a search of seconds (it warms up), so candidates timed late are
    // not measured under the same conditions as the start. Confirm the winner against the start
    // timed alternately, two rounds each, and keep it only if it is still 2 % faster.
    if (!st && memcmp(&win, &start, sizeof(win)) != 0) {
        float t_start = 0, t_win = 0;
        for (int r = 0; r < 2 && !st; ++r) {
            float ms = 0;
            b->tun = start;
            if ((st = median_ms(&ms))) break;
            t_start += ms;
            b->tun = win;
            if ((st = median_ms(&ms))) break;
            t_win += ms;
        }
        if (!st && !(t_win < 0.98f * t_start)) win = start;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->tun = st ? start : win;
    MIXER_TRY(plan(b));
    if (st) return st;
    *best = win;
    return MIXER_OK;
}

int mixer_batch_set_tuning(mixer_batch* b, const mixer_batch_tuning* t) {
    if (!b || !t || (t->mfma != 0 && t->mfma != 16 && t->mfma != 32)) return MIXER_ERR_ARG;
    MIXER_HIP(hipSetDevice(b->device));
    MIXER_HIP(hipStreamSynchronize(b->s));
    const mixer_batch_tuning old = b->tun;
    b->tun = *t;
    b->tun.staging = old.staging;  // the staging mode is fixed at creation
    const int st = plan(b);
    if (st) {
        b->tun = old;
        (void)plan(b);
    }
    return st;
}

int mixer_batch_get_tuning(mixer_batch* b, mixer_batch_tuning* t) {
    if (!b || !t) return MIXER_ERR_ARG;
    *t = b->tun;
    return MIXER_OK;
}

int mixer_fill_random(void* stream, void* d_ptr, size_t bytes, uint64_t seed) {
    if (!d_ptr || bytes % 8) return MIXER_ERR_ARG;
    return hip_status(launch_fill_random((hipStream_t)stream, (uint8_t*)d_ptr, bytes, seed));
}

}  // extern "C"
