// corr_quad.h -- the QUAD correlator: arrays of NB % 4 == 0 element blocks (N = 128, 256, 384,
// ..., 1024) staged straight from the heaps.
//
// Staging from the heaps, a one-channel workgroup fetches 32-byte runs: every 128-byte line of
// a heap plane holds 4 channels, so each of its reads uses a quarter of a line and the gather is
// bound by the number of L2 requests. Here a workgroup owns 4 adjacent channels (one line per
// element and plane) and one item of the triangle, so each staged request is a whole line and the
// register budget (5 tiles per wave) is unchanged. Items (a per-size list, see quad_items_for):
//   (d0, -1)  4-block diagonal triangle at blocks d0..d0+3 (10 tiles per channel): wave (ch, s)
//             runs the fold pair (s, 3 - s) of channel ch (5 tiles)
//   (r0, c0)  rectangle rows r0, r0+1 x columns c0..c0+3 below the diagonal (8 tiles per
//             channel): wave (ch, s) runs both rows of channel ch against columns c0+2s, c0+2s+1
//             (4 tiles from 4 fragments; one row against 4 columns reads 5)
// Per channel the items cover the lower triangle exactly once (N = 256: triangles 0-3, 4-7 and
// rectangles rows 4-5, 6-7 x columns 0-3).
#pragma once
#include "corr_impl.h"

namespace mixer {

template <int NE>
struct QuadLds {
    static constexpr int PS = NE * 32 + 128;   // plane stride: +128 B staggers the im plane's banks
    static constexpr int CS = 2 * PS + 32;     // channel stride: +32 B staggers the channels' banks
    static constexpr int BYTES = 4 * CS;
};

// Staging of one 32-sample block of NBLK element blocks for 4 channels: thread t stages plane
// t / 256 (waves 0-3 the real plane, waves 4-7 the imaginary one), and its chunk i (16 B) is
// k = t % 256 + 256 i: line g = k / 8 (element e(g): lines 2j, 2j+1 of each group of 8 are elements
// j and j + 4), channel (k / 2) % 4, half k % 2. 8 consecutive threads load one contiguous
// 128-byte line and a wave's load covers 8 whole lines; in LDS the lines of elements j and j + 4
// fall in disjoint halves of the 256-byte bank row (channel stride = 32 B mod 256), so the 16
// lanes of each store phase hit distinct banks. Only the imaginary waves test for -128, one wave
// per SIMD, with every lane busy.
template <int NBLK>
struct QuadStager {
    static constexpr int NE = NBLK * 32;
    static constexpr int CHUNKS = NE * 16;
    static constexpr int PER = CHUNKS / 512;
    v4i r[PER];
    __device__ __forceinline__ static void chunk(int i, int& el, int& p, int& ch, int& h) {
        const int t = threadIdx.x, k = (t & 255) + 256 * i, g = k >> 3;
        p = t >> 8;
        el = (g & ~7) | ((g & 1) << 2) | ((g >> 1) & 3);
        ch = (k >> 1) & 3;
        h = k & 1;
    }
    // Buffer loads: the block's base in a scalar descriptor, one 32-bit offset register per load
    // (the same offsets for every block, so both register sets share them), where a global load
    // would keep a 64-bit address per load.
    template <typename GB>
    __device__ __forceinline__ void load(const int8_t* sl, const XSrc& m, GB gblk, int q4, int C) {
        const __amdgpu_buffer_rsrc_t rs = block_rsrc(sl);
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            int el, p, ch, h;
            chunk(i, el, p, ch, h);
            const int c = q4 + ch < C ? q4 + ch : C - 1;  // ragged last quad: re-read a valid channel
            const uint32_t o = m.off32(p, gblk(el >> 5) * 32 + (el & 31), h) + (uint32_t)c * (uint32_t)m.chan;
            r[i] = bload(rs, o);
        }
    }
    __device__ __forceinline__ void store(uint8_t* L) const {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            int el, p, ch, h;
            chunk(i, el, p, ch, h);
            const int unit = 2 * el + (h ^ (((el >> 2) ^ (el >> 3)) & 1));
            *(v4i*)(L + ch * QuadLds<NE>::CS + p * QuadLds<NE>::PS + unit * 16) = r[i];
        }
    }
    // Nonzero if an imaginary byte staged by this thread is -128 (three operations per register).
    __device__ __forceinline__ unsigned saw80() const {
        unsigned s = 0;
        if (threadIdx.x < 256) return 0;  // real-plane waves (uniform per wave)
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            s = acc80(s, r[i].x);
            s = acc80(s, r[i].y);
            s = acc80(s, r[i].z);
            s = acc80(s, r[i].w);
        }
        return s & 0x80808080u;
    }
};

// Fragment of LDS block slot B of one channel region (the per-channel layout of lds_unit with
// the QuadLds plane stride).
template <int NE>
__device__ __forceinline__ Frag quad_frag(const uint8_t* Lc, int B, int lrow, int kg) {
    const int e = B * 32 + lrow;
    const int unit = 2 * e + (kg ^ (((e >> 2) ^ (e >> 3)) & 1));
    Frag f;
    f.re = *(const v4i*)(Lc + unit * 16);
    f.im = *(const v4i*)(Lc + QuadLds<NE>::PS + unit * 16);
    f.ng = neg8x16(f.im);
    return f;
}

// MFMAs of one block: row slot RA against column slots [A0, A1), optional row slot RB against
// [B0, B1) (RB < 0: none).
template <int NE, int RA, int A0, int A1, int RB, int B0, int B1>
__device__ __forceinline__ void quad_mfma(const uint8_t* Lc, int lrow, int kg, v16i* ar, v16i* ai,
                                          v16i* br, v16i* bi) {
    constexpr int JLO = (RB >= 0 && B0 < A0) ? B0 : A0;
    constexpr int JHI = (RB >= 0 && B1 > A1) ? B1 : A1;
    const Frag fa = quad_frag<NE>(Lc, RA, lrow, kg);
    Frag fb = fa;
    if constexpr (RB >= 0) fb = quad_frag<NE>(Lc, RB, lrow, kg);
    static_for<JLO, JHI>([&](auto Jc) {
        constexpr int J = decltype(Jc)::value;
        constexpr bool inA = J >= A0 && J < A1;
        constexpr bool inB = RB >= 0 && J >= B0 && J < B1;
        if constexpr (inA || inB) {
            Frag b;
            if constexpr (J == RA) b = fa;
            else if constexpr (J == RB) b = fb;
            else b = quad_frag<NE>(Lc, J, lrow, kg);
            if constexpr (inA) tile_step(fa, b, ar[J - A0], ai[J - A0]);
            if constexpr (inB) tile_step(fb, b, br[J - B0], bi[J - B0]);
        }
    });
}

// One wave's program over the item's time range. TR maps a row/column LDS slot to the global
// element block; the tiles are (TR(RA), TR(J)) for J in [A0, A1) and (TR(RB), TR(J)) for J in
// [B0, B1). V = 1: loads one block ahead; V = 3: two blocks ahead (two register sets).
template <int V, int NBLK, int RA, int A0, int A1, int RB, int B0, int B1, typename GB>
__device__ __forceinline__ void quad_wave(const int8_t* __restrict__ src, const XSrc& m, GB gblk,
                                          int q4, int C, uint8_t* lds0, uint8_t* lds1, int n_kb,
                                          int ch, bool valid, int32_t* __restrict__ acc,
                                          int accumulate, bool chk, unsigned& saw) {
    constexpr int NE = NBLK * 32;
    constexpr int NA = A1 - A0;
    constexpr int NBT = RB >= 0 ? B1 - B0 : 1;
    const int lane = threadIdx.x & 63, lrow = lane & 31, kg = lane >> 5;
    v16i ar[NA], ai[NA], br[NBT], bi[NBT];
#pragma unroll
    for (int j = 0; j < NA; ++j) { ar[j] = (v16i){0}; ai[j] = (v16i){0}; }
#pragma unroll
    for (int j = 0; j < NBT; ++j) { br[j] = (v16i){0}; bi[j] = (v16i){0}; }
    const int coff = ch * QuadLds<NE>::CS;
    using St = QuadStager<NBLK>;
    auto compute = [&](const uint8_t* L) {
        quad_mfma<NE, RA, A0, A1, RB, B0, B1>(L + coff, lrow, kg, ar, ai, br, bi);
    };
    if constexpr (V == 1) {
        St st;
        st.load(src, m, gblk, q4, C);
        st.store(lds0);
        if (chk) saw |= st.saw80();
        __syncthreads();
        for (int kb = 0; kb < n_kb; ++kb) {
            const bool more = kb + 1 < n_kb;
            if (more) st.load(src + (size_t)(kb + 1) * m.tb, m, gblk, q4, C);
            compute((kb & 1) ? lds1 : lds0);
            if (more) {
                st.store((kb & 1) ? lds0 : lds1);
                if (chk) saw |= st.saw80();  // overlaps the stores' LDS latency
            }
            __syncthreads();
        }
    } else {
        // block j in register set j & 1, both parities written out so the sets keep fixed
        // registers. Loads and stores are unconditional (past the end the last block is loaded
        // and staged again, never computed), so the wait before each store covers only the
        // older set's loads: a load skipped on some path makes the compiler wait for both.
        St s0, s1;
        s0.load(src, m, gblk, q4, C);
        s1.load(src + (size_t)(n_kb > 1 ? 1 : 0) * m.tb, m, gblk, q4, C);
        s0.store(lds0);
        if (chk) saw |= s0.saw80();
        __syncthreads();
        const int8_t* nx = src + 2 * m.tb;
        const int8_t* last = src + (size_t)(n_kb - 1) * m.tb;
        int kb = 0;
        for (; kb + 2 <= n_kb; kb += 2) {
            s0.load(kb + 2 < n_kb ? nx : last, m, gblk, q4, C);
            compute(lds0);
            s1.store(lds1);
            if (chk) saw |= s1.saw80();  // after the stores: overlaps their LDS latency
            __syncthreads();
            s1.load(kb + 3 < n_kb ? nx + m.tb : last, m, gblk, q4, C);
            compute(lds1);
            s0.store(lds0);
            if (chk) saw |= s0.saw80();
            __syncthreads();
            nx += 2 * m.tb;
        }
        if (kb < n_kb) {  // odd tail: block kb in lds0
            compute(lds0);
            __syncthreads();
        }
    }
    if (!valid) return;
    const int lr = lrow;
#pragma unroll
    for (int j = 0; j < NA; ++j) {
        const int I = gblk(RA), J = gblk(A0 + j);
        int32_t* t = acc + ((size_t)I * (I + 1) / 2 + J) * 2048;
        store_tile(ar[j], t, lr, kg, accumulate);
        store_tile(ai[j], t + 1024, lr, kg, accumulate);
    }
    if constexpr (RB >= 0) {
#pragma unroll
        for (int j = 0; j < NBT; ++j) {
            const int I = gblk(RB), J = gblk(B0 + j);
            int32_t* t = acc + ((size_t)I * (I + 1) / 2 + J) * 2048;
            store_tile(br[j], t, lr, kg, accumulate);
            store_tile(bi[j], t + 1024, lr, kg, accumulate);
        }
    }
}

// grid: n_items x ceil(C / 4) quads x n_slots (XCD-mapped: the items of a quad adjacent and
// consecutive quads of a slot on one XCD, so they share the quad's lines in L2), 512 threads.
template <int V>
__global__ __launch_bounds__(512)
void k_xcorr_quad(XSrc xs, int32_t* __restrict__ acc, uint32_t* __restrict__ wrap,
                  const int2* __restrict__ items, int n_items, int nb, int C, int n_tb,
                  int n_slots, int accumulate) {
    const int NT = nb * (nb + 1) / 2;
    __shared__ __attribute__((aligned(16))) uint8_t lds0[QuadLds<192>::BYTES];
    __shared__ __attribute__((aligned(16))) uint8_t lds1[QuadLds<192>::BYTES];
    const int nq = (C + 3) / 4, total = n_items * nq * n_slots;
    const int logical = xcd_logical(blockIdx.x, total);
    if (logical >= total) return;  // uniform over the workgroup
    const int item = logical % n_items, q = (logical / n_items) % nq, slot = logical / n_items / nq;
    const int kb0 = (int)((long)n_tb * slot / n_slots);
    const int kb1 = (int)((long)n_tb * (slot + 1) / n_slots);
    if (kb1 <= kb0) return;
    const int w = threadIdx.x >> 6, ch = w >> 1, s = w & 1;
    const int q4 = 4 * q, c = q4 + ch;
    const bool valid = c < C;
    int32_t* a = acc + ((size_t)slot * C + (valid ? c : 0)) * NT * 2048;
    const int8_t* src = xs.base + (size_t)kb0 * xs.tb;
    const int n_kb = kb1 - kb0;
    const bool chk = wrap != nullptr;
    unsigned saw = 0;
    const int2 it = items[item];
    if (it.y < 0) {
        const int d0 = it.x;
        auto gblk = [=](int b) { return d0 + b; };
        if (s == 0)
            quad_wave<V, 4, 3, 0, 4, 0, 0, 1>(src, xs, gblk, q4, C, lds0, lds1, n_kb, ch, valid, a,
                                              accumulate, chk, saw);
        else
            quad_wave<V, 4, 2, 0, 3, 1, 0, 2>(src, xs, gblk, q4, C, lds0, lds1, n_kb, ch, valid, a,
                                              accumulate, chk, saw);
    } else {
        const int r0 = it.x, c0 = it.y;  // LDS slots 0, 1: rows r0, r0 + 1; 2..5: columns c0..c0+3
        auto gblk = [=](int b) { return b < 2 ? r0 + b : c0 + b - 2; };
        if (s == 0)
            quad_wave<V, 6, 0, 2, 4, 1, 2, 4>(src, xs, gblk, q4, C, lds0, lds1, n_kb, ch, valid, a,
                                              accumulate, chk, saw);
        else
            quad_wave<V, 6, 0, 4, 6, 1, 4, 6>(src, xs, gblk, q4, C, lds0, lds1, n_kb, ch, valid, a,
                                              accumulate, chk, saw);
    }
    count_wrap(wrap, saw);
}

template <int V>
inline hipError_t launch_quad_t(hipStream_t s, const XcorrPlan& p, const XSrc& src, int32_t* acc,
                                uint32_t* wrap, int C, int n_tb, int accumulate) {
    const int wgs = p.n_quad * ((C + 3) / 4) * p.n_slots;
    hipLaunchKernelGGL((k_xcorr_quad<V>), dim3((wgs + 7) / 8 * 8), dim3(512), 0, s, src, acc, wrap,
                       (const int2*)p.d_quad, p.n_quad, p.n_elem / 32, C, n_tb, p.n_slots,
                       accumulate);
    return hipGetLastError();
}

}  // namespace mixer
