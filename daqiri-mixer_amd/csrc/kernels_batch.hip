// kernels_batch.hip -- batched launches over many independent arrays ("jobs") of different
// sizes: one corner-turn launch, two persistent correlator launches pulling work items from
// per-XCD queues (arrays of more than 8 element blocks on the blocked wave programs, the others
// on a runtime-generic fold program), one integrate (+ gridding) launch and one image launch.
// The batch engine captures the sequence once as a hipGraph and replays it.
#include "corr_impl.h"
#include "kernels_body.h"

namespace mixer {

// Job j owning global workgroup b of a batched launch, from the per-job first-workgroup
// offsets (jobs are few, a linear scan over the table is cheap).
template <typename F>
__device__ __forceinline__ int job_of(const BatchJobDev* jobs, int n_jobs, int b, F wg0) {
    int j = 0;
    while (j + 1 < n_jobs && b >= wg0(jobs[j + 1])) ++j;
    return j;
}

__global__ void k_batch_reset(int* counters) {
    if (threadIdx.x < 16) counters[threadIdx.x] = 0;
}

constexpr int kBatchCCH = 8;  // corner-turn channels per workgroup in batched launches

__global__ __launch_bounds__(256) void k_batch_corner_turn(const BatchJobDev* __restrict__ jobs,
                                                           int n_jobs) {
    const int b = blockIdx.x;
    const int j = job_of(jobs, n_jobs, b, [](const BatchJobDev& J) { return J.ct_wg0; });
    const BatchJobDev& J = jobs[j];
    const int local = b - J.ct_wg0;
    const int tb = local % J.n_tb, r = local / J.n_tb;
    const int eb = r % J.nb, cc = r / J.nb;
    corner_turn_body<kBatchCCH>(J.heaps, J.tiles, J.wrap, J.n_elem, J.n_chan, J.n_tb, tb, eb * 32,
                                cc * kBatchCCH, J.n_chan);
}

// ------------------------------------------------------------------------------------------
// Runtime-generic correlator wave program for the persistent kernel: one code path for every
// array size. A wave owns up to kMaxT tiles: row A (LDS block ra, global block ra_g) with
// columns [a0, a0+na) and row B with columns [b0, b0+nb) (LDS slots; global *_g), accumulator
// slot s < na for row A and na <= s < na+nb for row B, so accumulator indices stay static.
// The group's nth threads (local ids ltid; ltid < 0 for waves outside any group, which only
// meet the barriers) stage nblk element blocks (global block gblk(b)) per 32-sample block,
// addressed through m (tiles or heaps), through registers into two LDS buffers (one barrier
// per block, like corr_wave with V = 1).
// ------------------------------------------------------------------------------------------
constexpr int kMaxT = 5;
// 16-byte chunks staged per thread per block: a fold group of NB <= 8 blocks has NB * 128 chunks
// over 64 * fold_waves_rt(NB) threads, at most 2 each
constexpr int kMaxStage = 2;

struct WaveTiles {
    int ra, ra_g, a0, a0_g, na;
    int rb, rb_g, b0, b0_g, nb;
};

__device__ __forceinline__ int lds_unit_rt(int ne, int p, int e, int g) {
    return p * 2 * ne + 2 * e + (g ^ (((e >> 2) ^ (e >> 3)) & 1));
}
__device__ __forceinline__ Frag load_frag_rt(const uint8_t* L, int ne, int B, int lrow, int kg) {
    Frag f;
    f.re = *(const v4i*)(L + (lds_unit_rt(ne, 0, B * 32 + lrow, kg) << 4));
    f.im = *(const v4i*)(L + (lds_unit_rt(ne, 1, B * 32 + lrow, kg) << 4));
    f.ng = neg8x16(f.im);
    return f;
}

template <typename GB>
__device__ __forceinline__ void corr_wave_rt(const int8_t* __restrict__ src, const XSrc& m, int nblk,
                                             GB gblk, uint8_t* lds0, uint8_t* lds1, int n_kb,
                                             const WaveTiles& wt, bool active, int32_t* __restrict__ acc,
                                             int ltid, int nth) {
    const int lane = threadIdx.x & 63, lrow = lane & 31, kg = lane >> 5;
    const int ne = nblk * 32, chunks = nblk * 128;
    v16i ar[kMaxT], ai[kMaxT];
#pragma unroll
    for (int s = 0; s < kMaxT; ++s) { ar[s] = (v16i){0}; ai[s] = (v16i){0}; }
    v4i st[kMaxStage];
    auto gload = [&](const int8_t* sl) {
#pragma unroll
        for (int i = 0; i < kMaxStage; ++i) {
            const int u = ltid + i * nth;
            if (ltid >= 0 && u < chunks) {
                const int b = u >> 7, w = u & 127, p = w >> 6, eg = w & 63;
                st[i] = gld((const v4i*)(sl + m.off(p, gblk(b) * 32 + (eg >> 1), eg & 1)));
            }
        }
    };
    auto sstore = [&](uint8_t* L) {
#pragma unroll
        for (int i = 0; i < kMaxStage; ++i) {
            const int u = ltid + i * nth;
            if (ltid >= 0 && u < chunks) {
                const int b = u >> 7, w = u & 127, p = w >> 6, eg = w & 63;
                *(v4i*)(L + (lds_unit_rt(ne, p, b * 32 + (eg >> 1), eg & 1) << 4)) = st[i];
            }
        }
    };
    gload(src);
    sstore(lds0);
    __syncthreads();
    for (int kb = 0; kb < n_kb; ++kb) {
        const bool more = kb + 1 < n_kb;
        if (more) gload(src + (size_t)(kb + 1) * m.tb);
        if (active) {
            const uint8_t* L = (kb & 1) ? lds1 : lds0;
            const Frag fa = load_frag_rt(L, ne, wt.ra, lrow, kg);
            const Frag fb = wt.nb ? load_frag_rt(L, ne, wt.rb, lrow, kg) : fa;
#pragma unroll
            for (int s = 0; s < kMaxT; ++s) {
                if (s < wt.na) {
                    const Frag b = load_frag_rt(L, ne, wt.a0 + s, lrow, kg);
                    tile_step(fa, b, ar[s], ai[s]);
                } else if (s < wt.na + wt.nb) {
                    const Frag b = load_frag_rt(L, ne, wt.b0 + s - wt.na, lrow, kg);
                    tile_step(fb, b, ar[s], ai[s]);
                }
            }
        }
        if (more) sstore((kb & 1) ? lds0 : lds1);
        __syncthreads();
    }
    if (!active) return;
#pragma unroll
    for (int s = 0; s < kMaxT; ++s) {
        if (s < wt.na + wt.nb) {
            const int I = s < wt.na ? wt.ra_g : wt.rb_g;
            const int J = s < wt.na ? wt.a0_g + s : wt.b0_g + s - wt.na;
            int32_t* t = acc + (size_t)(I * (I + 1) / 2 + J) * 2048;
            store_tile(ar[s], t, lrow, kg, 0);
            store_tile(ai[s], t + 1024, lrow, kg, 0);
        }
    }
}

// The same program in the 16x16x64 tile shape: two-block (64-sample) stages, one barrier per
// stage, an odd last block paired with zeros; the accumulator tiles as four 16x16 sub-tiles.
__device__ __forceinline__ Frag load_frag16_rt(const uint8_t* L, int ne, int B, int h, int lane) {
    const int e = B * 32 + 16 * h + (lane & 15), g = lane >> 4;
    const uint8_t* blk = L + (g >> 1) * (ne * 64);
    Frag f;
    f.re = *(const v4i*)(blk + (lds_unit_rt(ne, 0, e, g & 1) << 4));
    f.im = *(const v4i*)(blk + (lds_unit_rt(ne, 1, e, g & 1) << 4));
    f.ng = neg8x16(f.im);
    return f;
}

template <typename GB>
__device__ __forceinline__ void corr_wave_rt16(const int8_t* __restrict__ src, const XSrc& m, int nblk,
                                               GB gblk, uint8_t* lds0, uint8_t* lds1, int n_kb,
                                               const WaveTiles& wt, bool active,
                                               int32_t* __restrict__ acc, int ltid, int nth) {
    const int lane = threadIdx.x & 63;
    const int ne = nblk * 32;
    v16i ar[kMaxT], ai[kMaxT];  // four 16x16 sub-tiles per tile
#pragma unroll
    for (int s = 0; s < kMaxT; ++s) { ar[s] = (v16i){0}; ai[s] = (v16i){0}; }
    // Staging by direct global -> LDS loads (no staging registers): a stage is 2 blocks x 2 planes
    // x nblk element blocks of 1 KiB windows; lane l of a window fills its 16-byte unit l (element
    // l / 2, sample half (l & 1) ^ swizzle), so the layout is lds_unit_rt's. A block past the end
    // is written as zeros.
    const int el = lane >> 1, hsw = ((lane & 1) ^ ((el >> 2) ^ (el >> 3))) & 1;
    const int nwin = 2 * nblk, wg = ltid >> 6, nw = nth >> 6;
    auto dma = [&](uint8_t* L, int kb_first) {
        if (ltid < 0) return;
        for (int win = wg; win < 2 * nwin; win += nw) {
            const int kbl = win / nwin, r = win % nwin, p = r / nblk, b = r % nblk;
            const int kb = kb_first + kbl;
            uint8_t* l = L + kbl * (ne * 64) + p * (ne * 32) + b * 1024;
            if (kb < n_kb) {
                const int8_t* g = src + (size_t)kb * m.tb + m.off(p, gblk(b) * 32 + el, hsw);
                __builtin_amdgcn_global_load_lds((const void*)g, MIXER_LDS_PTR(l), 16, 0, 0);
            } else {
                *(v4i*)(l + lane * 16) = (v4i){0, 0, 0, 0};
            }
        }
    };
    const int n_st = (n_kb + 1) / 2;
    dma(lds0, 0);
    __syncthreads();  // waits for this wave's loads into LDS before the barrier
    for (int sg = 0; sg < n_st; ++sg) {
        uint8_t* cur = (sg & 1) ? lds1 : lds0;
        uint8_t* nxt = (sg & 1) ? lds0 : lds1;
        if (sg + 1 < n_st) dma(nxt, 2 * (sg + 1));
        if (active) {
            // row A's tiles, then row B's: one row's fragments live at a time, and each column
            // fragment half is read just before its MFMAs
            {
                const Frag fa[2] = {load_frag16_rt(cur, ne, wt.ra, 0, lane), load_frag16_rt(cur, ne, wt.ra, 1, lane)};
#pragma unroll
                for (int s = 0; s < kMaxT; ++s) {
                    if (s < wt.na) {
                        tile_half16v<0>(fa, load_frag16_rt(cur, ne, wt.a0 + s, 0, lane), ar[s], ai[s]);
                        tile_half16v<1>(fa, load_frag16_rt(cur, ne, wt.a0 + s, 1, lane), ar[s], ai[s]);
                    }
                    // keep the next tile's fragment reads below these MFMAs (register pressure)
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            if (wt.nb) {
                const Frag fb[2] = {load_frag16_rt(cur, ne, wt.rb, 0, lane), load_frag16_rt(cur, ne, wt.rb, 1, lane)};
#pragma unroll
                for (int s = 0; s < kMaxT; ++s) {
                    if (s >= wt.na && s < wt.na + wt.nb) {
                        tile_half16v<0>(fb, load_frag16_rt(cur, ne, wt.b0 + s - wt.na, 0, lane), ar[s], ai[s]);
                        tile_half16v<1>(fb, load_frag16_rt(cur, ne, wt.b0 + s - wt.na, 1, lane), ar[s], ai[s]);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
        __syncthreads();
    }
    if (!active) return;
#pragma unroll
    for (int s = 0; s < kMaxT; ++s) {
        if (s < wt.na + wt.nb) {
            const int I = s < wt.na ? wt.ra_g : wt.rb_g;
            const int J = s < wt.na ? wt.a0_g + s : wt.b0_g + s - wt.na;
            int32_t* t = acc + (size_t)(I * (I + 1) / 2 + J) * 2048;
            store_tile16v(ar[s], t, lane, 0);
            store_tile16v(ai[s], t + 1024, lane, 0);
        }
    }
}

// Waves per channel of the folded layout in the persistent kernel: NB = 1 is one tile on one
// wave; NB >= 2 splits each row pair (p, NB-1-p) over two waves (at most 5 tiles per wave).
__device__ __host__ inline int fold_waves_rt(int nb) { return nb == 1 ? 1 : 2 * ((nb + 1) / 2); }

int batch_fold_groups(int nb) { return nb > 8 ? 1 : 8 / fold_waves_rt(nb); }

// The split-fold tiles of wave w of an nb-block channel (same layout as FoldLayout<NB, 2>).
__device__ __forceinline__ WaveTiles fold_tiles_rt(int nb, int w, bool* active) {
    WaveTiles t{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    *active = w < fold_waves_rt(nb);
    if (!*active) return t;
    if (nb == 1) {
        t.na = 1;
        return t;
    }
    const int P = (nb + 1) / 2, p = w % P, R1 = p, R2 = nb - 1 - p;
    const int T = (R1 == R2) ? R1 + 1 : R1 + R2 + 2, S = (T + 1) / 2;
    if (w < P) {
        t.ra = t.ra_g = R2; t.a0 = t.a0_g = 0; t.na = S;
    } else if (R1 == R2) {
        t.ra = t.ra_g = R2; t.a0 = t.a0_g = S; t.na = R2 + 1 - S;
    } else {
        t.ra = t.ra_g = R1; t.a0 = t.a0_g = 0; t.na = R1 + 1;
        if (S <= R2) { t.rb = t.rb_g = R2; t.b0 = t.b0_g = S; t.nb = R2 + 1 - S; }
    }
    *active = t.na + t.nb > 0;
    return t;
}

// LDS of the persistent fold correlator: two stage buffers per channel group, at most 32 KiB
// (8 blocks: 2 x 16 KiB; packed folds e.g. 4 groups x 2 x 4 KiB at 64 elements).
constexpr int kBatchLds = 32 * 1024;

// Take the next item index of this workgroup's XCD queue, or of the others' once it is empty;
// -1 when every queue is empty (the same value for the whole workgroup). probe (thread 0):
// queues found empty so far; an empty queue stays empty.
__device__ __forceinline__ int next_item(const BatchQueues& q, int* counters, int home, int& probe,
                                         int& item_sh) {
    if (threadIdx.x == 0) {
        int it = -1;
        for (; probe < 8; ++probe) {
            const int x = (home + probe) & 7;
            const int k = atomicAdd(&counters[x], 1);
            if (k < q.size[x]) {
                it = q.start[x] + k;
                break;
            }
        }
        item_sh = it;
    }
    __syncthreads();
    const int it = item_sh;
    __syncthreads();
    return it;
}

// Persistent correlator: each workgroup takes the next work item (job, first channel,
// super-block, time slot) from the queue of its XCD (dispatch places workgroup b on XCD b % 8)
// and, once that queue is empty, from the other XCDs' queues in turn, until all are empty. A
// queue holds whole channels, largest first, so the super-blocks that share a channel's element
// blocks run together on one XCD and share its L2; the queues end with the short items.
// SHAPE 32: v_mfma_i32_32x32x32_i8 tiles, one-block stages; 16: v_mfma_i32_16x16x64_i8, two-block
// stages (twice the LDS).
template <int SHAPE>
__global__ __launch_bounds__(512) void k_batch_xcorr(const BatchJobDev* __restrict__ jobs,
                                                     const int4* __restrict__ items, BatchQueues q,
                                                     int* __restrict__ counters) {
    constexpr int KB = SHAPE == 16 ? 2 : 1;  // blocks per stage
    __shared__ __attribute__((aligned(16))) uint8_t lds[KB * kBatchLds];
    __shared__ int item_sh;
    const int w8 = threadIdx.x >> 6;
    const int home = blockIdx.x & 7;
    int probe = 0;  // thread 0: queues found empty so far (an empty queue stays empty)
    for (;;) {
        const int it = next_item(q, counters, home, probe, item_sh);
        if (it < 0) break;  // every wave reaches this with the same value
        const int4 w = items[it];
        const BatchJobDev& J = jobs[w.x];
        const int nb = J.nb, slot = w.w;
        const int kb0 = (int)((long)J.n_tb * slot / J.n_slots);
        const int kb1 = (int)((long)J.n_tb * (slot + 1) / J.n_slots);
        const int NT = nb * (nb + 1) / 2;
        if (kb1 > kb0) {  // uniform
            // Fold items only (arrays of at most 8 element blocks; the larger ones run on
            // k_batch_xcorr_blk): one call site for every array size, a channel group per item.
            const XSrc m = J.fused ? xsrc_heaps(J.heaps, J.n_elem, J.n_chan)
                                 : xsrc_tiles(J.tiles, J.n_elem, J.n_tb);
            bool active = false;
            const int wpc = fold_waves_rt(nb), G = min(8 / wpc, J.ch_per_item);
            const int g = w8 / wpc;
            const int nblk = nb, buf = KB * nb * 32 * 64, nth = 64 * wpc;
            const int ltid = g < G ? (int)(threadIdx.x % nth) : -1;
            int c = w.y + (g < G ? g : 0);
            const bool have = g < G && c < J.n_chan;
            if (!have) c = w.y;  // idle groups stage a valid slab and store nothing
            const WaveTiles wt = fold_tiles_rt(nb, w8 % wpc, &active);
            uint8_t* l0 = lds + (g < G ? g : 0) * 2 * buf;
            const int row0 = 0, col0 = 4;  // identity block map
            auto gblk = [=](int b) {
                const int g = b < 4 ? row0 + b : col0 + (b - 4);
                return g < nb ? g : nb - 1;  // columns past the array only feed idle waves
            };
            const int8_t* src = m.base + (size_t)c * m.chan + (size_t)kb0 * m.tb;
            int32_t* a = J.acc + ((size_t)slot * J.n_chan + c) * NT * 2048;
            if constexpr (SHAPE == 16)
                corr_wave_rt16(src, m, nblk, gblk, l0, l0 + buf, kb1 - kb0, wt, active && have, a,
                               ltid, nth);
            else
                corr_wave_rt(src, m, nblk, gblk, l0, l0 + buf, kb1 - kb0, wt, active && have, a,
                             ltid, nth);
        }
        __syncthreads();  // LDS reuse by the next item
    }
}

// Persistent correlator for the arrays of more than 8 element blocks: the items are BLOCKED
// super-blocks (diagonal 8-block triangles, full 4 x 8 super-blocks, the tail triangle) run by
// the same wave programs as k_xcorr_blk (two register sets, loads two blocks ahead).
__global__ __launch_bounds__(512) void k_batch_xcorr_blk(const BatchJobDev* __restrict__ jobs,
                                                         const int4* __restrict__ items,
                                                         BatchQueues q,
                                                         int* __restrict__ counters) {
    __shared__ __attribute__((aligned(16))) uint8_t lds0[BlkLds<3>::BYTES];
    __shared__ __attribute__((aligned(16))) uint8_t lds1[BlkLds<3>::BYTES];
    __shared__ int item_sh;
    const int home = blockIdx.x & 7;
    int probe = 0;
    for (;;) {
        const int it = next_item(q, counters, home, probe, item_sh);
        if (it < 0) break;
        const int4 w = items[it];
        const BatchJobDev& J = jobs[w.x];
        const int nb = J.nb, slot = w.w, c = w.y;
        const int kb0 = (int)((long)J.n_tb * slot / J.n_slots);
        const int kb1 = (int)((long)J.n_tb * (slot + 1) / J.n_slots);
        if (kb1 > kb0) {  // uniform
            const XSrc m = J.fused ? xsrc_heaps(J.heaps, J.n_elem, J.n_chan)
                                   : xsrc_tiles(J.tiles, J.n_elem, J.n_tb);
            const int2 sb = J.sblocks[w.z];
            const int NT = nb * (nb + 1) / 2;
            const int8_t* src = m.base + (size_t)c * m.chan + (size_t)kb0 * m.tb;
            int32_t* a = J.acc + ((size_t)slot * J.n_chan + c) * NT * 2048;
            unsigned saw = 0;
            blk_item_n<3>(src, m, AccOut{a, 0}, lds0, lds1, kb1 - kb0, sb.x, sb.y, nb, threadIdx.x, false,
                          saw);
        }
        __syncthreads();  // LDS reuse by the next item
    }
}

__global__ __launch_bounds__(256) void k_batch_integrate(const BatchJobDev* __restrict__ jobs,
                                                         int n_jobs) {
    const int b = blockIdx.x;
    const int j = job_of(jobs, n_jobs, b, [](const BatchJobDev& J) { return J.int_wg0; });
    const BatchJobDev& J = jobs[j];
    const int local = b - J.int_wg0;
    integrate_seg(J.acc, J.v, nullptr, J.G, J.n_elem, J.n_chan, J.n_slots, J.grid_x, J.ng,
                  J.segs[local % J.n_segs], local / J.n_segs);
}

__global__ __launch_bounds__(256) void k_batch_image(const BatchJobDev* __restrict__ jobs,
                                                     int n_jobs) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int b = blockIdx.x;
    const int j = job_of(jobs, n_jobs, b, [](const BatchJobDev& J) { return J.img_wg0; });
    const BatchJobDev& J = jobs[j];
    const size_t nimg = (size_t)J.n_chan * J.ng * J.ng;
    image_body(J.G, nullptr, J.tw, J.img, J.img + nimg, J.img + nimg + J.n_chan, J.ng, b - J.img_wg0, smem);
}

hipError_t launch_batch_reset(hipStream_t s, int* counters) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, counters);
    return hipGetLastError();
}

hipError_t launch_batch_corner_turn(hipStream_t s, const BatchJobDev* jobs, int n_jobs, int n_wgs) {
    hipLaunchKernelGGL(k_batch_corner_turn, dim3(n_wgs), dim3(256), 0, s, jobs, n_jobs);
    return hipGetLastError();
}

hipError_t launch_batch_xcorr(hipStream_t s, const BatchJobDev* jobs, const int4* items,
                              const BatchQueues& q, int* counters, int n_wgs, int mfma) {
    if (mfma == 16)
        hipLaunchKernelGGL(k_batch_xcorr<16>, dim3(n_wgs), dim3(512), 0, s, jobs, items, q, counters);
    else
        hipLaunchKernelGGL(k_batch_xcorr<32>, dim3(n_wgs), dim3(512), 0, s, jobs, items, q, counters);
    return hipGetLastError();
}

hipError_t launch_batch_xcorr_blk(hipStream_t s, const BatchJobDev* jobs, const int4* items,
                                  const BatchQueues& q, int* counters, int n_wgs) {
    hipLaunchKernelGGL(k_batch_xcorr_blk, dim3(n_wgs), dim3(512), 0, s, jobs, items, q, counters);
    return hipGetLastError();
}

hipError_t launch_batch_integrate(hipStream_t s, const BatchJobDev* jobs, int n_jobs, int n_wgs,
                                  int max_ng) {
    hipLaunchKernelGGL(k_batch_integrate, dim3(n_wgs), dim3(256), 0, s, jobs, n_jobs);
    return hipGetLastError();
}

hipError_t launch_batch_image(hipStream_t s, const BatchJobDev* jobs, int n_jobs, int n_wgs,
                              int max_ng) {
    hipLaunchKernelGGL(k_batch_image, dim3(n_wgs), dim3(256), (size_t)max_ng * max_ng * 16, s, jobs,
                       n_jobs);
    return hipGetLastError();
}

hipError_t set_batch_lds_limits(int max_ng) {
    return hipFuncSetAttribute((const void*)k_batch_image, hipFuncAttributeMaxDynamicSharedMemorySize,
                               max_ng * max_ng * 16);
}

}  // namespace mixer
