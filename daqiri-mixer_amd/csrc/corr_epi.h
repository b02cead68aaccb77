// corr_epi.h -- the BLOCKED correlator with the integrate fused into its epilogue.
//
// When one launch with one time slot covers a whole integration, every lower-triangle tile is
// final when its wave leaves the time loop. Instead of storing the int32 accumulator tile (and
// the integrate kernel reading it back: 1.1 GB each way at 1024 x 256), the wave writes what
// k_integrate would write from it (replaces k_flush_cube, src/rx_host_corr.cu:86-95, and the
// gridding of k_grid_cube, :79-85):
//   - V block (I, J) in f32 straight from the MFMA registers (store and/or add), one rounding of
//     the exact int32 sum;
//   - for I > J the conjugate block (J, I), through a [32][33] LDS transpose of one plane at a
//     time (coalesced 16-byte row stores);
//   - the tile's baselines into the exact int64 image grid G, which holds one orientation of each
//     baseline (the image kernel adds the conjugate mirror): tiles below the diagonal add into
//     int64 LDS cells of the workgroup's baseline rows, flushed once per workgroup; diagonal
//     tiles (I == J, which hold both orientations) add their lower baselines and
//     autocorrelations straight into G.
// The cell layout follows integrate_seg: with R = 32 / gs, sub-block (ri, rj) and diagonal dd of
// tile (I, J) are the baseline cell dy = R (I - J) + ri - rj, dx = dd.
#pragma once
#include "corr_impl.h"

namespace mixer {

constexpr int kEpiT = 32 * 33;     // ints of one wave's transpose region (4224 bytes)
constexpr int kEpiCells = 744;     // >= (12 R - 1)(2 gs - 1) for every gs dividing 32 (713)

// Fence + wave barrier: the wave's LDS writes are visible to all its lanes.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct VOut {
    VOutArgs a;
    int c;                         // channel
    int* T;                        // this wave's transpose region (LDS)
    unsigned long long* cells;     // [kEpiCells][2] int64 grid cells (LDS, workgroup)
    int dy0, ncell;                // cell q: dy = dy0 + q / (2 gs - 1), dx = q % (2 gs - 1) - (gs - 1)
    int ltid, nth;                 // thread id and count of the workgroup

    __device__ __forceinline__ void begin() const {
        __syncthreads();  // every wave has left the time loop: the staging LDS is free
        if (a.G)
            for (int q = ltid; q < 2 * ncell; q += nth) cells[q] = 0ull;
        __syncthreads();
    }

    __device__ __forceinline__ void end() const {
        __syncthreads();
        if (!a.G) return;
        const int gs = a.gs, W = 2 * gs - 1, ng = a.ng;
        unsigned long long* Gc = a.G + (size_t)c * 2 * ng * ng;
        for (int q = ltid; q < ncell; q += nth) {
            const unsigned long long re = cells[2 * q], im = cells[2 * q + 1];
            if (!re && !im) continue;
            grid_add_lower(Gc, ng, dy0 + q / W, q % W - (gs - 1), re, im, true);
        }
    }

    // Plane p of tile (I, J) from the wave's transpose region: the conjugate block and the grid.
    __device__ __forceinline__ void from_T(int I, int J, int p, int lane) const {
        const int N = a.N, d = I - J;
        if (d > 0 && (a.v_store || a.v_add)) {
            // block (J, I) = conj(block (I, J))^T: lane writes row J*32 + b, columns I*32 + a0..+15
            const int b = lane >> 1, a0 = (lane & 1) * 16;
            const float sg = p ? -1.f : 1.f;
            const size_t off = (((size_t)c * 2 + p) * N + (size_t)J * 32 + b) * N + (size_t)I * 32 + a0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r0 = a0 + 4 * q;
                const v4f f = {sg * (float)T[(r0 + 0) * 33 + b], sg * (float)T[(r0 + 1) * 33 + b],
                               sg * (float)T[(r0 + 2) * 33 + b], sg * (float)T[(r0 + 3) * 33 + b]};
                if (a.v_store) gst((v4f*)(a.v_store + off + 4 * q), f);
                if (a.v_add) gst((v4f*)(a.v_add + off + 4 * q), gld((const v4f*)(a.v_add + off + 4 * q)) + f);
            }
        }
        if (!a.G) return;
        const int gs = a.gs, R = 32 / gs, W = 2 * gs - 1, ng = a.ng, nseg = R * R * W;
        unsigned long long* Gc = a.G + (size_t)c * 2 * ng * ng;
        for (int sgi = lane; sgi < nseg; sgi += 64) {
            const int sub = sgi / W, dd = sgi % W - (gs - 1);
            const int ri = sub / R, rj = sub % R;
            // a diagonal tile holds both orientations: its lower baselines and autocorrelations
            if (d == 0 && (ri < rj || (ri == rj && dd < 0))) continue;
            const int i0 = dd > 0 ? dd : 0, len = gs - (dd > 0 ? dd : -dd);
            // the diagonal from (ri gs + i0, rj gs + i0 - dd), 34 ints per step; eight reads per
            // round are issued back to back (the workgroup is alone on its CU here, so a chain of
            // dependent LDS round trips is not hidden by other waves)
            const int* t = T + (ri * gs + i0) * 33 + rj * gs + i0 - dd;
            long long sum = 0;
            for (int k0 = 0; k0 < len; k0 += 8) {
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (k0 + k < len) sum += t[(k0 + k) * 34];
            }
            if (!sum) continue;
            const int dy = R * d + ri - rj;
            if (d > 0) {
                const int q = (dy - dy0) * W + dd + gs - 1;
                atomicAdd(&cells[2 * q + p], (unsigned long long)sum);
            } else {
                const int y = dy < 0 ? dy + ng : dy, x = dd < 0 ? dd + ng : dd;
                gadd(&Gc[2 * (y * ng + x) + p], (unsigned long long)sum);  // never aliases to q = 0
            }
        }
    }

    __device__ __forceinline__ void operator()(const v16i& vr, const v16i& vi, int I, int J, int lrow,
                                               int kg) const {
        const int N = a.N, lane = threadIdx.x & 63;
        if (a.v_store || a.v_add) {  // block (I, J): register r is row 8(r/4) + 4 kg + r%4, column lrow
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const v16i& x = p ? vi : vr;
                const size_t base = (((size_t)c * 2 + p) * N + (size_t)I * 32) * N + (size_t)J * 32 + lrow;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const size_t off = base + (size_t)(8 * (r >> 2) + 4 * kg + (r & 3)) * N;
                    const float f = (float)x[r];
                    if (a.v_store) gst(a.v_store + off, f);
                    if (a.v_add) gst(a.v_add + off, gld(a.v_add + off) + f);
                }
            }
        }
        if (I == J && !a.G) return;
        if (I > J && !a.G && !(a.v_store || a.v_add)) return;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const v16i& x = p ? vi : vr;
#pragma unroll
            for (int r = 0; r < 16; ++r) T[(8 * (r >> 2) + 4 * kg + (r & 3)) * 33 + lrow] = x[r];
            wave_lds_sync();
            from_T(I, J, p, lane);
            wave_lds_sync();  // before the next plane overwrites T
        }
    }
};

// The epilogue of work item (I0, J0) (J0 < 0: the 8-block diagonal triangle at I0) of channel c on
// 512 threads, with the staging buffers lds0 and lds1 (each >= 24 KiB) reused after the time loop:
// waves 0-4 transpose in lds0, waves 5-7 in lds1, then the cells.
__device__ __forceinline__ VOut make_vout(const VOutArgs& vo, int c, int I0, int J0, uint8_t* lds0,
                                          uint8_t* lds1) {
    const int w = threadIdx.x >> 6;
    const int R = 32 / vo.gs;
    // baseline rows of the tiles below the diagonal: d = I - J in [dmin, dmax], d >= 1
    int dmin, dmax;
    if (J0 < 0) { dmin = 1; dmax = 7; }
    else { dmin = I0 - J0 - 7 > 1 ? I0 - J0 - 7 : 1; dmax = I0 + 3 - J0; }
    VOut out;
    out.a = vo;
    out.c = c;
    out.T = (int*)(w < 5 ? lds0 + w * kEpiT * 4 : lds1 + (w - 5) * kEpiT * 4);
    out.cells = (unsigned long long*)(lds1 + 3 * kEpiT * 4);
    out.dy0 = R * dmin - (R - 1);
    out.ncell = (R * (dmax - dmin) + 2 * R - 1) * (2 * vo.gs - 1);
    out.ltid = threadIdx.x;
    out.nth = 512;
    return out;
}

// k_xcorr_blk with the fused integrate: one workgroup per (super-block, channel) of one time slot
// covering the whole integration (n_tb blocks). The wave programs take the array's block count nb
// at run time (as in the batched kernel), so one instance serves every size.
template <int V>
__global__ __launch_bounds__(512)
void k_xcorr_blk_v(XSrc xs, VOutArgs vo, uint32_t* __restrict__ wrap, const int2* __restrict__ sblocks,
                   int n_sb, int nb, int C, int n_tb) {
    static_assert(BlkLds<V>::BYTES >= 5 * kEpiT * 4 && BlkLds<V>::BYTES >= 3 * kEpiT * 4 + kEpiCells * 16,
                  "epilogue LDS must fit the staging buffers");
    __shared__ __attribute__((aligned(16))) uint8_t lds0[BlkLds<V>::BYTES];
    __shared__ __attribute__((aligned(16))) uint8_t lds1[BlkLds<V>::BYTES];
    const int total = n_sb * C;
    const int logical = xcd_logical(blockIdx.x, total);
    if (logical >= total) return;
    const int c = logical / n_sb;
    const int2 sb = sblocks[logical % n_sb];
    const VOut out = make_vout(vo, c, sb.x, sb.y, lds0, lds1);
    const int8_t* src = xs.base + (size_t)c * xs.chan;
    unsigned saw = 0;
    blk_item_n<V>(src, xs, out, lds0, lds1, n_tb, sb.x, sb.y, nb, threadIdx.x, wrap != nullptr, saw);
    count_wrap(wrap, saw);
}

template <int V>
inline hipError_t launch_blk_v_t(hipStream_t s, const XcorrPlan& p, const XSrc& src, const VOutArgs& vo,
                                 uint32_t* wrap, int C, int n_tb) {
    const int wgs = p.n_tasks * C;
    hipLaunchKernelGGL((k_xcorr_blk_v<V>), dim3((wgs + 7) / 8 * 8), dim3(512), 0, s, src, vo, wrap,
                       (const int2*)p.d_tasks, p.n_tasks, p.n_elem / 32, C, n_tb);
    return hipGetLastError();
}

}  // namespace mixer
