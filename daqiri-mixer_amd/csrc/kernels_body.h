// kernels_body.h -- device bodies of the corner-turn, integrate (+ gridding) and image kernels,
// shared by the single-array launches (kernels.hip) and the batched multi-array launches
// (kernels_batch.hip).
#pragma once
#include "device_common.h"

namespace mixer {

// ------------------------------------------------------------------------------------------
// Corner-turn. One workgroup = one time block x 32 elements x CCH channels (CCH*2 KiB of LDS).
// Load: 64-lane runs of 16 B from heap planes (a wave reads 64*16 B contiguous: one (element,
// plane) run of 32 channels, or two of 16, ...), all issued before the first LDS write.
// Store: each wave writes whole 1 KiB rows = 32 elements x 32 samples of one (channel, plane).
// LDS rows are XOR-swizzled in 16-byte units so both phases are bank-conflict free.
// ------------------------------------------------------------------------------------------
template <int CCH>
__device__ __forceinline__ void corner_turn_body(const uint8_t* __restrict__ heaps,
                                                 int8_t* __restrict__ tiles,
                                                 uint32_t* __restrict__ wrap, int N, int C,
                                                 int n_tb, int tb, int e0, int c0, int c_end) {
    constexpr int NLD = CCH / 2;  // 16-byte loads (and stores) per thread
    __shared__ __attribute__((aligned(16))) uint8_t lds[CCH * 2048];
    const int tid = threadIdx.x;
    const size_t payload = (size_t)C * 64;
    unsigned saw = 0;
    v4i v[NLD];
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
        const int u = tid + i * 256;           // chunk: (element, plane) run, channel, half
        const int q = u / (2 * CCH), w = u % (2 * CCH);
        const int e = q >> 1, p = q & 1, c = w >> 1, h = w & 1;
        v[i] = (v4i){0, 0, 0, 0};
        if (c0 + c < c_end)
            v[i] = gld((const v4i*)(heaps + ((size_t)tb * N + e0 + e) * payload + (size_t)p * C * 32 +
                                    (size_t)(c0 + c) * 32 + h * 16));
    }
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
        const int u = tid + i * 256;
        const int q = u / (2 * CCH), w = u % (2 * CCH);
        const int e = q >> 1, p = q & 1, c = w >> 1, h = w & 1;
        if (p) saw |= has80(v[i].x) | has80(v[i].y) | has80(v[i].z) | has80(v[i].w);
        const int s = (e * 2 + h) ^ ((c * 2) & 15);
        *(v4i*)(lds + (((p * CCH + c) * 64 + s) << 4)) = v[i];
    }
    __syncthreads();
    const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int row = wv; row < 2 * CCH; row += 4) {
        const int p = row / CCH, cc = row % CCH;
        if (c0 + cc >= c_end) continue;
        const int s = lane ^ ((cc * 2) & 15);
        const v4i x = *(const v4i*)(lds + (((p * CCH + cc) * 64 + s) << 4));
        gst((v4i*)(tiles + ((((size_t)(c0 + cc) * n_tb + tb) * 2 + p) * N + e0) * 32 + lane * 16), x);
    }
    if (wrap) {
        __shared__ unsigned any;
        if (tid == 0) any = 0;
        __syncthreads();
        if (saw) atomicOr(&any, 1u);
        __syncthreads();
        if (tid == 0 && any) atomicAdd(wrap, 1u);
    }
}

// ------------------------------------------------------------------------------------------
// Integrate + grid one segment of a tile diagonal: tiles (J + d, J), J in [j0, j1), of channel
// c (grid: x = segment, y = channel). For each tile the accumulator slots are summed in int32
// (exact) and converted once to f32; the tile is written as V block (I, J) and, for d > 0, as
// the conjugate-transposed block (J, I) through an LDS transpose, so every accumulator tile is
// read once (v_store overwrites, v_add accumulates; either may be null).
// Gridding (G non-null): grid_x (gs) divides 32, so a gs x gs sub-block (ri, rj) of tile (I, J)
// pairs one grid row of elements with one other and each of its diagonals dd is one baseline
// cell dy = R d + ri - rj (R = 32 / gs), dx = dd. All tiles of a segment share d, so they touch
// the same (2R - 1)(2gs - 1) <= 105 cells: the diagonals are summed in chunks of <= 8 elements
// (all 256 threads busy) into an int64 LDS grid of those cells, which is added once to G[c]
// (int64 [ng*ng][2], exact) at the end, conjugated at (-dy, -dx) as well for d > 0 (the mirror
// of a tile below the diagonal is the upper triangle). With both V pointers null it only grids.
// ------------------------------------------------------------------------------------------
constexpr int kSegCells = 128;  // >= max over gs of (2 * 32 / gs - 1) * (2 * gs - 1) = 105

__device__ __forceinline__ void grid_seg_tile(const int (*T)[32][33], unsigned long long* Gs, int gs,
                                              bool diag) {
    const int R = 32 / gs, W = 2 * gs - 1, nseg = R * R * W;
    const int len = gs > 8 ? 8 : gs, nck = gs / len;
    for (int w = threadIdx.x; w < nseg * nck; w += 256) {
        const int sg = w / nck, k = w % nck;
        const int sub = sg / W, dd = sg % W - (gs - 1);
        const int ri = sub / R, rj = sub % R;
        // a diagonal tile holds both orientations of its baselines: grid the lower ones (row >
        // column) and the autocorrelations (dd = 0 on the diagonal sub-blocks) only
        if (diag && (ri < rj || (ri == rj && dd < 0))) continue;
        const int i0 = dd > 0 ? dd : 0, i1 = dd > 0 ? gs : gs + dd;
        const int c0 = i0 + k * len, c1 = c0 + len < i1 ? c0 + len : i1;
        if (c0 >= c1) continue;
        long long sr = 0, si = 0;
        for (int ip = c0; ip < c1; ++ip) {
            sr += T[0][ri * gs + ip][rj * gs + ip - dd];
            si += T[1][ri * gs + ip][rj * gs + ip - dd];
        }
        const int cell = (ri - rj + R - 1) * W + (dd + gs - 1);
        atomicAdd(&Gs[2 * cell], (unsigned long long)sr);
        atomicAdd(&Gs[2 * cell + 1], (unsigned long long)si);
    }
}

__device__ __forceinline__ void integrate_seg(const int32_t* __restrict__ acc,
                                              float* __restrict__ v_store,
                                              float* __restrict__ v_add,
                                              unsigned long long* __restrict__ G, int N, int C,
                                              int n_slots, int grid_x, int ng, int4 seg, int c) {
    __shared__ int T[2][32][33];
    __shared__ unsigned long long Gs[2 * kSegCells];
    const int d = seg.x, j0 = seg.y, j1 = seg.z;
    const int nb = N / 32, NT = nb * (nb + 1) / 2;
    const int tid = threadIdx.x, i = tid >> 3, jq = (tid & 7) * 4;
    const size_t slot_stride = (size_t)C * NT * 2048;
    const int32_t* ac = acc + (size_t)c * NT * 2048;
    const bool write_v = v_store || v_add;
    const int gs = grid_x, R = 32 / gs, W = 2 * gs - 1, ncell = (2 * R - 1) * W;
    if (G) Gs[tid] = 0ull;  // 256 = 2 * kSegCells entries
    // the slot-summed tile (J + d, J): this thread's row i, columns jq..jq+3, both planes
    auto load = [&](int J, v4i* x) {
        const int I = J + d;
        const int32_t* t = ac + ((size_t)I * (I + 1) / 2 + J) * 2048 + i * 32 + jq;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            x[p] = gld((const v4i*)(t + p * 1024));
            for (int sl = 1; sl < n_slots; ++sl) x[p] += gld((const v4i*)(t + sl * slot_stride + p * 1024));
        }
    };
    auto put = [&](int row, int col, int p, float4 f) {  // V element (row, col..col+3) of plane p
        const size_t off = ((size_t)c * 2 + p) * N * N + (size_t)row * N + col;
        const v4f fv = {f.x, f.y, f.z, f.w};
        if (v_store) gst((v4f*)(v_store + off), fv);
        if (v_add) {
            gst((v4f*)(v_add + off), gld((const v4f*)(v_add + off)) + fv);
        }
    };
    v4i nxt[2];
    load(j0, nxt);
    for (int J = j0; J < j1; ++J) {
        const int I = J + d;
        const v4i cur[2] = {nxt[0], nxt[1]};
        if (J + 1 < j1) load(J + 1, nxt);  // in flight while this tile is stored and gridded
        if (write_v) {
#pragma unroll
            for (int p = 0; p < 2; ++p)
                put(I * 32 + i, J * 32 + jq, p,
                    make_float4((float)cur[p].x, (float)cur[p].y, (float)cur[p].z, (float)cur[p].w));
        }
        if (G || (write_v && d > 0)) {
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                T[p][i][jq + 0] = cur[p].x;
                T[p][i][jq + 1] = cur[p].y;
                T[p][i][jq + 2] = cur[p].z;
                T[p][i][jq + 3] = cur[p].w;
            }
            __syncthreads();
            if (G) grid_seg_tile(T, Gs, gs, d == 0);
            if (write_v && d > 0) {  // block (J, I) = conj(block (I, J))^T
                put(J * 32 + i, I * 32 + jq, 0,
                    make_float4((float)T[0][jq][i], (float)T[0][jq + 1][i], (float)T[0][jq + 2][i],
                                (float)T[0][jq + 3][i]));
                put(J * 32 + i, I * 32 + jq, 1,
                    make_float4((float)-T[1][jq][i], (float)-T[1][jq + 1][i], (float)-T[1][jq + 2][i],
                                (float)-T[1][jq + 3][i]));
            }
            __syncthreads();
        }
    }
    if (G) {
        unsigned long long* Gc = G + (size_t)c * 2 * ng * ng;
        for (int q = tid; q < ncell; q += 256) {
            const unsigned long long re = Gs[2 * q], im = Gs[2 * q + 1];
            if (!re && !im) continue;
            const int dy = R * d + q / W - (R - 1), dx = q % W - (gs - 1);
            grid_add_lower(Gc, ng, dy, dx, re, im, d > 0);
        }
    }
}

// ------------------------------------------------------------------------------------------
// Image: one workgroup per channel: the exact int64 grid G[c] (from k_integrate) is converted
// once to f32, transformed by a forward unnormalised ng x ng FFT (Stockham radix-2 along x
// then y, ping-pong in LDS, twiddles tw[q] = exp(-2 pi i q / ng) from an FP64 host table), and
// published as fftshift + real part + per-channel max/min. G[c] is zeroed for the next
// integration.
// ------------------------------------------------------------------------------------------
// Frame grid (optional): the exact int64 grid summed over the integrations since the last frame
// (the reference's monitor images dV, the change since its previous frame,
// src/rx_host_corr.cu:218-222). One workgroup owns a channel's cells, so a plain add suffices.
__device__ __forceinline__ void frame_add(unsigned long long* Gfc, int q, unsigned long long re,
                                          unsigned long long im) {
    gst(&Gfc[2 * q], gld(&Gfc[2 * q]) + re);
    gst(&Gfc[2 * q + 1], gld(&Gfc[2 * q + 1]) + im);
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// ng = 32 in registers: a 32-point transform runs on 8 lanes with 4 points each (lane r of a line
// holds points r, r + 8, r + 16, r + 24). A 4-point DFT in registers, the twiddles W32^(r k1),
// then an 8-point DFT across the 8 lanes in three shuffle butterflies leave lane r with
// X[k1 + 4 bitrev3(r)], k1 = 0..3 (X[k] = sum_n x[n] W32^(nk), n = r + 8m, k = k1 + 4 k2). The
// row and column passes meet in one LDS transpose (rows padded to 33 points), so the image takes
// one barrier where the radix-2 Stockham passes take ten.
__device__ __forceinline__ float2 tw32(const float2* __restrict__ tw, int q) {  // W32^q, q < 32
    const float2 t = tw[q & 15];  // the table holds W32^0..W32^15
    return q < 16 ? t : make_float2(-t.x, -t.y);
}

__device__ __forceinline__ void fft32_line(float2 v[4], int r, const float2* __restrict__ tw) {
    // 4-point DFT over m, W4 = -i; (-i) (x + iy) = y - ix
    const float2 s02 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
    const float2 d02 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
    const float2 s13 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
    const float2 d13 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
    v[0] = make_float2(s02.x + s13.x, s02.y + s13.y);
    v[1] = make_float2(d02.x + d13.y, d02.y - d13.x);
    v[2] = make_float2(s02.x - s13.x, s02.y - s13.y);
    v[3] = make_float2(d02.x - d13.y, d02.y + d13.x);
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1) v[k1] = cmul(v[k1], tw32(tw, r * k1));
    // 8-point decimation-in-frequency DFT across the lanes r ^ h (h = 4, 2, 1): the lower lane of
    // a pair keeps a + b, the upper (a - b) W_{2h}^(r mod h), W_{2h}^j = W32^(16 j / h)
#pragma unroll
    for (int h = 4; h >= 1; h >>= 1) {
        const bool upper = (r & h) != 0;
        const float2 w = tw32(tw, (16 / h) * (r & (h - 1)));
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            const float2 p = make_float2(__shfl_xor(v[k1].x, h), __shfl_xor(v[k1].y, h));
            v[k1] = upper ? cmul(make_float2(p.x - v[k1].x, p.y - v[k1].y), w)
                          : make_float2(v[k1].x + p.x, v[k1].y + p.y);
        }
    }
}

__device__ __forceinline__ void image32_body(unsigned long long* __restrict__ G,
                                             unsigned long long* __restrict__ Gf,
                                             const float2* __restrict__ tw, float* __restrict__ img,
                                             float* __restrict__ vmax, float* __restrict__ vmin,
                                             int c, unsigned char* smem) {
    constexpr int ng = 32, half = 16, TS = 33;  // transpose row stride (float2)
    float2* T = (float2*)smem;                  // [32][33]
    __shared__ float red_mx[4], red_mn[4];
    const int tid = threadIdx.x, lane = tid & 63;
    const int r = lane & 7, line = (tid >> 6) * 8 + (lane >> 3);  // 32 lines of 8 lanes
    const int br = ((r & 1) << 2) | (r & 2) | ((r >> 2) & 1);      // bitrev3(r)
    unsigned long long* Gc = G + (size_t)c * 2 * ng * ng;
    float2 v[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {  // row y = line, column x = r + 8 m
        const int x = r + 8 * m, q = line * ng + x;
        const int qm = ((ng - line) & (ng - 1)) * ng + ((ng - x) & (ng - 1));  // -q
        // F(q) = G(q) + conj(G(-q)), F(0) = G(0); the mirror loads are unconditional so that all
        // eight go out together (-0 = 0 reads the cell itself)
        const unsigned long long gr = gld(&Gc[2 * q]), gi = gld(&Gc[2 * q + 1]);
        const unsigned long long mr = gld(&Gc[2 * qm]), mi = gld(&Gc[2 * qm + 1]);
        const unsigned long long fr = q ? gr + mr : gr, fi = q ? gi - mi : gi;
        v[m] = make_float2((float)(long long)fr, (float)(long long)fi);
        if (Gf) frame_add(Gf + (size_t)c * 2 * ng * ng, q, gr, gi);
    }
    fft32_line(v, r, tw);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) T[line * TS + k1 + 4 * br] = v[k1];
    __syncthreads();  // every cell and its mirror are read: clear the grid for the next integration
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int q = line * ng + r + 8 * m;
        gst(&Gc[2 * q], 0ull);
        gst(&Gc[2 * q + 1], 0ull);
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) v[m] = T[(r + 8 * m) * TS + line];  // column kx = line
    fft32_line(v, r, tw);
    float mx = -1e30f, mn = 1e30f;
    float* im = img + (size_t)c * ng * ng;
    const int x = (line + half) & (ng - 1);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {  // F[ky][kx] at ky = k1 + 4 br, fftshifted
        const int y = (k1 + 4 * br + half) & (ng - 1);
        const float val = v[k1].x;
        gst(&im[y * ng + x], val);
        mx = fmaxf(mx, val);
        mn = fminf(mn, val);
    }
    for (int off = 32; off > 0; off >>= 1) {
        mx = fmaxf(mx, __shfl_xor(mx, off));
        mn = fminf(mn, __shfl_xor(mn, off));
    }
    if (lane == 0) { red_mx[tid >> 6] = mx; red_mn[tid >> 6] = mn; }
    __syncthreads();
    if (tid == 0) {
        float a = red_mx[0], b = red_mn[0];
        for (int k = 1; k < 4; ++k) { a = fmaxf(a, red_mx[k]); b = fminf(b, red_mn[k]); }
        gst(&vmax[c], a);
        gst(&vmin[c], b);
    }
}

__device__ __forceinline__ void image_body(unsigned long long* __restrict__ G,
                                           unsigned long long* __restrict__ Gf,
                                           const float2* __restrict__ tw, float* __restrict__ img,
                                           float* __restrict__ vmax, float* __restrict__ vmin,
                                           int ng, int c, unsigned char* smem) {
    if (ng == 32) {
        image32_body(G, Gf, tw, img, vmax, vmin, c, smem);
        return;
    }
    const int nn = ng * ng;
    float2* F0 = (float2*)smem;  // [nn]
    float2* F1 = F0 + nn;        // [nn]
    __shared__ float red_mx[4], red_mn[4];
    const int tid = threadIdx.x;
    unsigned long long* Gc = G + (size_t)c * 2 * nn;
    for (int q = tid; q < nn; q += 256) {
        const int y = q / ng, x = q % ng;
        const int qm = ((ng - y) & (ng - 1)) * ng + ((ng - x) & (ng - 1));  // -q
        // F(q) = G(q) + conj(G(-q)), F(0) = G(0)
        const unsigned long long gr = gld(&Gc[2 * q]), gi = gld(&Gc[2 * q + 1]);
        const unsigned long long mr = gld(&Gc[2 * qm]), mi = gld(&Gc[2 * qm + 1]);
        const unsigned long long fr = q ? gr + mr : gr, fi = q ? gi - mi : gi;
        F0[q] = make_float2((float)(long long)fr, (float)(long long)fi);
        if (Gf) frame_add(Gf + (size_t)c * 2 * nn, q, gr, gi);
    }
    __syncthreads();  // every cell and its mirror are read: clear the grid for the next integration
    for (int q = tid; q < nn; q += 256) {
        gst(&Gc[2 * q], 0ull);
        gst(&Gc[2 * q + 1], 0ull);
    }
    float2* src = F0;
    float2* dst = F1;
    const int half = ng >> 1;
    for (int dim = 0; dim < 2; ++dim) {
        const int sx = dim == 0 ? 1 : ng;   // stride between points of one transform
        const int sl = dim == 0 ? ng : 1;   // stride between transforms
        for (int Ns = 1; Ns < ng; Ns <<= 1) {
            for (int q = tid; q < nn / 2; q += 256) {
                const int line = q / half, j = q % half;
                const int o = line * sl;
                const float2 a = src[o + j * sx];
                const float2 b = cmul(src[o + (j + half) * sx], tw[(j % Ns) * (ng / (2 * Ns))]);
                const int d = (j / Ns) * 2 * Ns + (j % Ns);
                dst[o + d * sx] = make_float2(a.x + b.x, a.y + b.y);
                dst[o + (d + Ns) * sx] = make_float2(a.x - b.x, a.y - b.y);
            }
            __syncthreads();
            float2* tmp = src; src = dst; dst = tmp;
        }
    }
    float mx = -1e30f, mn = 1e30f;
    float* im = img + (size_t)c * nn;
    for (int q = tid; q < nn; q += 256) {
        const int y = q / ng, x = q % ng;
        const float v = src[((y + half) % ng) * ng + (x + half) % ng].x;
        gst(&im[q], v);
        mx = fmaxf(mx, v);
        mn = fminf(mn, v);
    }
    for (int off = 32; off > 0; off >>= 1) {
        mx = fmaxf(mx, __shfl_xor(mx, off));
        mn = fminf(mn, __shfl_xor(mn, off));
    }
    if ((tid & 63) == 0) { red_mx[tid >> 6] = mx; red_mn[tid >> 6] = mn; }
    __syncthreads();
    if (tid == 0) {
        float a = red_mx[0], b = red_mn[0];
        for (int k = 1; k < 4; ++k) { a = fmaxf(a, red_mx[k]); b = fminf(b, red_mn[k]); }
        gst(&vmax[c], a);
        gst(&vmin[c], b);
    }
}

}  // namespace mixer
