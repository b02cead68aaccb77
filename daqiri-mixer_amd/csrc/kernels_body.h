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
            v[i] = *(const v4i*)(heaps + ((size_t)tb * N + e0 + e) * payload + (size_t)p * C * 32 +
                                 (size_t)(c0 + c) * 32 + h * 16);
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
        *(v4i*)(tiles + ((((size_t)(c0 + cc) * n_tb + tb) * 2 + p) * N + e0) * 32 + lane * 16) = x;
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
// Gridding of one 32x32 accumulator tile (I, J), I >= J, held in LDS as T[plane][i][j].
// grid_x divides 32, so a grid_x x grid_x sub-block of the tile spans one grid row of elements
// on each side and every diagonal of it is one baseline cell (dy, dx) = ((qa-qb), (pa-pb)) mod
// ng: the diagonal is summed first, then added with one int64 LDS atomic (and once more,
// conjugated, at (-dy, -dx) for tiles below the diagonal, whose mirror is the upper triangle).
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void grid_tile(const int (*T)[32][33], unsigned long long* G, int I,
                                          int J, int gs, int ng) {
    const int nsub = 32 / gs, per = 2 * gs - 1, nseg = nsub * nsub * per;
    // diagonals are summed in chunks of at most 8 elements (4 per diagonal at grid_x 32), so
    // all 256 threads share the work; each chunk adds its partial sum to the cell
    const int len = gs > 8 ? 8 : gs, nck = gs / len;
    for (int w = threadIdx.x; w < nseg * nck; w += 256) {
        const int sg = w / nck, k = w % nck;
        const int sub = sg / per, d = sg % per - (gs - 1);
        const int ri = sub / nsub, rj = sub % nsub;
        const int i0 = d > 0 ? d : 0, i1 = d > 0 ? gs : gs + d;
        const int c0 = i0 + k * len, c1 = c0 + len < i1 ? c0 + len : i1;
        if (c0 >= c1) continue;
        long long sr = 0, si = 0;
        for (int ip = c0; ip < c1; ++ip) {
            sr += T[0][ri * gs + ip][rj * gs + ip - d];
            si += T[1][ri * gs + ip][rj * gs + ip - d];
        }
        const int a = I * 32 + ri * gs + i0, b = J * 32 + rj * gs + i0 - d;
        int dy = a / gs - b / gs, dx = a % gs - b % gs;
        dy += (dy < 0) ? ng : 0;
        dx += (dx < 0) ? ng : 0;
        const int cell = dy * ng + dx;
        atomicAdd(&G[2 * cell], (unsigned long long)sr);
        atomicAdd(&G[2 * cell + 1], (unsigned long long)si);
        if (I != J) {
            const int cell2 = (dy ? ng - dy : 0) * ng + (dx ? ng - dx : 0);
            atomicAdd(&G[2 * cell2], (unsigned long long)sr);
            atomicAdd(&G[2 * cell2 + 1], (unsigned long long)(-si));
        }
    }
}

// ------------------------------------------------------------------------------------------
// Integrate: grid x = row block I, y = channel. Sums the accumulator slots in int32 (exact),
// converts once to f32 and writes rows [32I, 32I+32) of both planes of the full square
// (v_store overwrites, v_add accumulates; either may be null). Tiles above the diagonal come
// from tile (J, I) transposed and conjugated through LDS. With G non-null it also grids the
// lower-triangle tiles it reads into an LDS partial image grid and adds that to G[c]
// (int64 [ng*ng][2], exact) with one global atomic per non-zero cell; with both V pointers
// null it only grids.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void integrate_body(const int32_t* __restrict__ acc,
                                               float* __restrict__ v_store,
                                               float* __restrict__ v_add,
                                               unsigned long long* __restrict__ G, int N, int C,
                                               int n_slots, int grid_x, int ng, int I, int c,
                                               unsigned char* smem) {
    unsigned long long* Gp = (unsigned long long*)smem;  // [ng*ng][2] when G is non-null
    __shared__ int T[2][32][33];
    const int nb = N / 32, NT = nb * (nb + 1) / 2;
    const int tid = threadIdx.x, i = tid >> 3, j0 = (tid & 7) * 4;
    const size_t slot_stride = (size_t)C * NT * 2048;
    const int32_t* ac = acc + (size_t)c * NT * 2048;
    const bool write_v = v_store || v_add;
    const int nn = ng * ng;
    if (G) {
        for (int q = tid; q < 2 * nn; q += 256) Gp[q] = 0ull;
    }
    const int jend = write_v ? nb : I + 1;
    // the slot-summed tile of column block J, this thread's 2 x 4 values
    auto load = [&](int J, v4i* x) {
        const bool lower = J <= I;
        const int32_t* t = ac + (size_t)(lower ? I * (I + 1) / 2 + J : J * (J + 1) / 2 + I) * 2048 +
                           i * 32 + j0;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            x[p] = *(const v4i*)(t + p * 1024);
            for (int sl = 1; sl < n_slots; ++sl) x[p] += *(const v4i*)(t + sl * slot_stride + p * 1024);
        }
    };
    v4i nxt[2];
    load(0, nxt);
    for (int J = 0; J < jend; ++J) {
        int vals[2][4];
        const bool lower = J <= I;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            vals[p][0] = nxt[p].x; vals[p][1] = nxt[p].y; vals[p][2] = nxt[p].z; vals[p][3] = nxt[p].w;
        }
        if (J + 1 < jend) load(J + 1, nxt);  // in flight while this tile is gridded and stored
        if (!lower || G) {
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) T[p][i][j0 + q] = vals[p][q];
            __syncthreads();
            if (lower) {
                grid_tile(T, Gp, I, J, grid_x, ng);
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    vals[0][q] = T[0][j0 + q][i];
                    vals[1][q] = -T[1][j0 + q][i];
                }
            }
            __syncthreads();
        }
        if (write_v) {
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                float4 f = make_float4((float)vals[p][0], (float)vals[p][1], (float)vals[p][2],
                                       (float)vals[p][3]);
                const size_t off = ((size_t)c * 2 + p) * N * N + (size_t)(I * 32 + i) * N + J * 32 + j0;
                if (v_store) *(float4*)(v_store + off) = f;
                if (v_add) {
                    float4 o = *(float4*)(v_add + off);
                    o.x += f.x; o.y += f.y; o.z += f.z; o.w += f.w;
                    *(float4*)(v_add + off) = o;
                }
            }
        }
    }
    if (G) {
        unsigned long long* Gc = G + (size_t)c * 2 * nn;
        for (int q = tid; q < 2 * nn; q += 256)
            if (Gp[q]) atomicAdd(&Gc[q], Gp[q]);
    }
}

// ------------------------------------------------------------------------------------------
// Image: one workgroup per channel: the exact int64 grid G[c] (from k_integrate) is converted
// once to f32, transformed by a forward unnormalised ng x ng FFT (Stockham radix-2 along x
// then y, ping-pong in LDS, twiddles tw[q] = exp(-2 pi i q / ng) from an FP64 host table), and
// published as fftshift + real part + per-channel max/min. G[c] is zeroed for the next
// integration.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ void image_body(unsigned long long* __restrict__ G,
                                           const float2* __restrict__ tw, float* __restrict__ img,
                                           float* __restrict__ vmax, float* __restrict__ vmin,
                                           int ng, int c, unsigned char* smem) {
    const int nn = ng * ng;
    float2* F0 = (float2*)smem;  // [nn]
    float2* F1 = F0 + nn;        // [nn]
    __shared__ float red_mx[4], red_mn[4];
    const int tid = threadIdx.x;
    unsigned long long* Gc = G + (size_t)c * 2 * nn;
    for (int q = tid; q < nn; q += 256) {
        F0[q] = make_float2((float)(long long)Gc[2 * q], (float)(long long)Gc[2 * q + 1]);
        Gc[2 * q] = 0ull;
        Gc[2 * q + 1] = 0ull;
    }
    __syncthreads();
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
        im[q] = v;
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
        vmax[c] = a;
        vmin[c] = b;
    }
}

}  // namespace mixer
