// corr_impl.h -- the correlator: lower-triangle complex int8 GEMM V += X X^H on
// v_mfma_i32_32x32x32_i8 with int32 accumulation into a tile-packed accumulator (replaces
// cgemm_blk_t<ACC>, src/cgemm.cuh:28-69).
//   k_xcorr_fold<NB,V>  arrays of NB <= 8 blocks of 32 elements (N <= 256)
//   k_xcorr_blk<NB,V>   arrays of NB = 12..32 blocks (N = 384..1024, N % 128 == 0)
// V selects the time-loop pipeline (the plan's "prefetch" field):
//   V = 1   one 32-sample block per barrier, staged global -> registers -> LDS one block ahead;
//   V = 3   the same with two register sets, loads issued two blocks ahead (more bytes in
//           flight per CU for the latency-bound gather straight from the heaps);
//   V = 5   three register sets, loads issued three blocks ahead;
//   V = 2,4 K = V blocks per barrier (BLOCKED: K = min(V, 3)), staged by direct
//           global -> LDS loads (global_load_lds_dwordx4) one stage ahead into two separately
//           declared LDS buffers, so the barrier and LDS-latency bubble at each stage boundary
//           is paid once per K blocks and no staging registers are needed.
//   V = 6   the 16x16x64 tile shape (v_mfma_i32_16x16x64_i8, FOLD only): one instruction
//           contracts two 32-sample blocks, so a stage is two blocks (one barrier per 64
//           samples), staged through registers one stage ahead (an odd last block is paired with
//           zeros).
#pragma once
#include "device_common.h"
#include "kernels.h"

namespace mixer {

#define MIXER_LDS_PTR(p) ((__attribute__((address_space(3))) void*)(p))

template <int V, int NBLK>
struct StageOf {
    static constexpr int KB = (V == 1 || V == 3 || V == 5) ? 1 : V == 6 ? 2 : (NBLK > 8 ? (V < 3 ? V : 3) : V);
};

// The MFMAs of one two-block stage in the 16x16x64 shape, tiles (RA, [A0,A1)) u (RB, [B0,B1)).
template <int NE, int RA, int A0, int A1, int RB, int B0, int B1, int NA, int NBT>
__device__ __forceinline__ void block_mfma16(const uint8_t* L, int lane, v4i (&ar)[NA][4],
                                             v4i (&ai)[NA][4], v4i (&br)[NBT][4], v4i (&bi)[NBT][4]) {
    constexpr int JLO = (RB >= 0 && B0 < A0) ? B0 : A0;
    constexpr int JHI = (RB >= 0 && B1 > A1) ? B1 : A1;
    const Frag fa[2] = {load_frag16<NE>(L, RA, 0, lane), load_frag16<NE>(L, RA, 1, lane)};
    Frag fb[2] = {fa[0], fa[1]};
    if constexpr (RB >= 0) {
        fb[0] = load_frag16<NE>(L, RB, 0, lane);
        fb[1] = load_frag16<NE>(L, RB, 1, lane);
    }
    static_for<JLO, JHI>([&](auto Jc) {
        constexpr int J = decltype(Jc)::value;
        constexpr bool inA = J >= A0 && J < A1;
        constexpr bool inB = RB >= 0 && J >= B0 && J < B1;
        if constexpr (inA || inB) {
            Frag b[2];
            if constexpr (J == RA) { b[0] = fa[0]; b[1] = fa[1]; }
            else if constexpr (J == RB) { b[0] = fb[0]; b[1] = fb[1]; }
            else { b[0] = load_frag16<NE>(L, J, 0, lane); b[1] = load_frag16<NE>(L, J, 1, lane); }
            if constexpr (inA) tile_step16(fa, b, ar[J - A0], ai[J - A0]);
            if constexpr (inB) tile_step16(fb, b, br[J - B0], bi[J - B0]);
        }
    });
}

// The MFMAs of one 32-sample block for tiles (RA, [A0,A1)) u (RB, [B0,B1)) from the LDS slab L
// (NE elements, swizzled as lds_unit<NE>).
template <int NE, int RA, int A0, int A1, int RB, int B0, int B1>
__device__ __forceinline__ void block_mfma(const uint8_t* L, int lrow, int kg, v16i* ar, v16i* ai,
                                           v16i* br, v16i* bi) {
    constexpr int JLO = (RB >= 0 && B0 < A0) ? B0 : A0;
    constexpr int JHI = (RB >= 0 && B1 > A1) ? B1 : A1;
    const Frag fa = load_frag<NE>(L, RA, lrow, kg);
    Frag fb = fa;
    if constexpr (RB >= 0) fb = load_frag<NE>(L, RB, lrow, kg);
    static_for<JLO, JHI>([&](auto Jc) {
        constexpr int J = decltype(Jc)::value;
        constexpr bool inA = J >= A0 && J < A1;
        constexpr bool inB = RB >= 0 && J >= B0 && J < B1;
        if constexpr (inA || inB) {
            Frag b;
            if constexpr (J == RA) b = fa;
            else if constexpr (J == RB) b = fb;
            else b = load_frag<NE>(L, J, lrow, kg);
            if constexpr (inA) tile_step(fa, b, ar[J - A0], ai[J - A0]);
            if constexpr (inB) tile_step(fb, b, br[J - B0], bi[J - B0]);
        }
    });
}

// ------------------------------------------------------------------------------------------
// One wave's share of the correlator over the whole time range (n_kb 32-sample blocks from
// src, addressed through m). The staged LDS slab holds NBLK element
// blocks (NE = 32*NBLK elements) per time block; GB maps a staged block slot to its element
// block in the global slab; NTH threads with local ids ltid share the staging. Every wave of
// the group runs this loop and meets the others at exactly one barrier per stage (plus one).
// ------------------------------------------------------------------------------------------
template <int V, int NE, int NBLK, int NTH, int RA, int A0, int A1, int RB, int B0, int B1, typename GB>
__device__ __forceinline__ void corr_wave(const int8_t* __restrict__ src, const XSrc& m,
                                          uint8_t* lds0, uint8_t* lds1, int n_kb, GB gblk,
                                          v16i* ar, v16i* ai, v16i* br, v16i* bi, int ltid,
                                          bool chk, unsigned& saw) {
    const int lane = threadIdx.x & 63, lrow = lane & 31, kg = lane >> 5;
    if constexpr (V == 3) {
        // two register sets: set s0 holds even blocks, s1 odd blocks; the loop body is written
        // out for both parities so both sets keep fixed registers. Every load and store is
        // unconditional (past the end the last block is loaded and staged again, and never
        // computed): a load skipped on some path would make the compiler's wait before the
        // older set's LDS store cover the newer set's loads too, one block of lead instead of two.
        auto ident = [](int b) { return b; };
        using St = Stager<NE, NBLK, NTH>;
        St s0, s1;
        auto gload = [&](St& st, const int8_t* sl) {
#pragma unroll
            for (int i = 0; i < St::PER; ++i) {
                const int u = ltid + i * NTH;
                if (St::CHUNKS % NTH == 0 || u < St::CHUNKS) {
                    const int b = u >> 7, w = u & 127, p = w >> 6, eg = w & 63;
                    st.r[i] = gld((const v4i*)(sl + m.off(p, gblk(b) * 32 + (eg >> 1), eg & 1)));
                }
            }
        };
        auto check = [&](const St& st) {
#pragma unroll
            for (int i = 0; i < St::PER; ++i) {
                const int u = ltid + i * NTH;
                if ((St::CHUNKS % NTH == 0 || u < St::CHUNKS) && ((u >> 6) & 1))
                    saw |= has80(st.r[i].x) | has80(st.r[i].y) | has80(st.r[i].z) | has80(st.r[i].w);
            }
        };
        gload(s0, src);
        gload(s1, src + (n_kb > 1 ? m.tb : 0));
        if (chk) check(s0);
        s0.store(lds0, ident, ltid);
        __syncthreads();
        const int8_t* nx = src + 2 * m.tb;
        const int8_t* last = src + (size_t)(n_kb - 1) * m.tb;
        int kb = 0;
        for (; kb + 2 <= n_kb; kb += 2) {
            // even step: block kb in lds0, s1 = kb+1, load kb+2 into s0
            gload(s0, kb + 2 < n_kb ? nx : last);
            block_mfma<NE, RA, A0, A1, RB, B0, B1>(lds0, lrow, kg, ar, ai, br, bi);
            if (chk) check(s1);
            s1.store(lds1, ident, ltid);
            __syncthreads();
            // odd step: block kb+1 in lds1, s0 = kb+2, load kb+3 into s1
            gload(s1, kb + 3 < n_kb ? nx + m.tb : last);
            block_mfma<NE, RA, A0, A1, RB, B0, B1>(lds1, lrow, kg, ar, ai, br, bi);
            if (chk) check(s0);
            s0.store(lds0, ident, ltid);
            __syncthreads();
            nx += 2 * m.tb;
        }
        if (kb < n_kb) {  // odd tail: block kb in lds0
            block_mfma<NE, RA, A0, A1, RB, B0, B1>(lds0, lrow, kg, ar, ai, br, bi);
            __syncthreads();
        }
    } else if constexpr (V == 5) {
        // three register sets, loads issued three blocks ahead: step j of a six-step round
        // computes block kb+j from LDS buffer j % 2, loads block kb+j+3 into set j % 3 (whose
        // block kb+j was staged one step earlier) and stages set (j+1) % 3 (block kb+j+1, loaded
        // two steps earlier) into the other buffer. Six steps (three sets x two buffers) keep
        // every set in fixed registers; loads past the end re-read the last block (never
        // computed), as in V = 3.
        auto ident = [](int b) { return b; };
        using St = Stager<NE, NBLK, NTH>;
        St s0, s1, s2;
        auto gload = [&](St& st, const int8_t* sl) {
#pragma unroll
            for (int i = 0; i < St::PER; ++i) {
                const int u = ltid + i * NTH;
                if (St::CHUNKS % NTH == 0 || u < St::CHUNKS) {
                    const int b = u >> 7, w = u & 127, p = w >> 6, eg = w & 63;
                    st.r[i] = gld((const v4i*)(sl + m.off(p, gblk(b) * 32 + (eg >> 1), eg & 1)));
                }
            }
        };
        auto check = [&](const St& st) {
#pragma unroll
            for (int i = 0; i < St::PER; ++i) {
                const int u = ltid + i * NTH;
                if ((St::CHUNKS % NTH == 0 || u < St::CHUNKS) && ((u >> 6) & 1))
                    saw |= has80(st.r[i].x) | has80(st.r[i].y) | has80(st.r[i].z) | has80(st.r[i].w);
            }
        };
        const int8_t* last = src + (size_t)(n_kb - 1) * m.tb;
        auto at = [&](int kb) { return kb < n_kb ? src + (size_t)kb * m.tb : last; };
        auto set = [&](auto Sc) -> St& {
            constexpr int S = decltype(Sc)::value;
            if constexpr (S == 0) return s0;
            else if constexpr (S == 1) return s1;
            else return s2;
        };
        auto step = [&](auto Jc, int kb) {
            constexpr int J = decltype(Jc)::value;
            St& ld = set(std::integral_constant<int, J % 3>{});
            St& sv = set(std::integral_constant<int, (J + 1) % 3>{});
            gload(ld, at(kb + J + 3));
            block_mfma<NE, RA, A0, A1, RB, B0, B1>((J & 1) ? lds1 : lds0, lrow, kg, ar, ai, br, bi);
            if (chk) check(sv);
            sv.store((J & 1) ? lds0 : lds1, ident, ltid);
            __syncthreads();
        };
        gload(s0, src);
        gload(s1, at(1));
        gload(s2, at(2));
        if (chk) check(s0);
        s0.store(lds0, ident, ltid);
        __syncthreads();
        int kb = 0;
        for (; kb + 6 <= n_kb; kb += 6) static_for<0, 6>([&](auto Jc) { step(Jc, kb); });
        static_for<0, 5>([&](auto Jc) {  // the last n_kb % 6 blocks
            if (kb + decltype(Jc)::value < n_kb) step(Jc, kb);
        });
    } else if constexpr (V == 1) {
        auto ident = [](int b) { return b; };
        using St = Stager<NE, NBLK, NTH>;
        St st;
        auto gload = [&](const int8_t* sl) {
#pragma unroll
            for (int i = 0; i < St::PER; ++i) {
                const int u = ltid + i * NTH;
                if (St::CHUNKS % NTH == 0 || u < St::CHUNKS) {
                    const int b = u >> 7, w = u & 127, p = w >> 6, eg = w & 63;
                    st.r[i] = gld((const v4i*)(sl + m.off(p, gblk(b) * 32 + (eg >> 1), eg & 1)));
                }
            }
        };
        // imaginary -128 (its negation wraps): checked on the staged registers
        auto check = [&]() {
#pragma unroll
            for (int i = 0; i < St::PER; ++i) {
                const int u = ltid + i * NTH;
                if ((St::CHUNKS % NTH == 0 || u < St::CHUNKS) && ((u >> 6) & 1))
                    saw |= has80(st.r[i].x) | has80(st.r[i].y) | has80(st.r[i].z) | has80(st.r[i].w);
            }
        };
        gload(src);
        if (chk) check();
        st.store(lds0, ident, ltid);
        __syncthreads();
        for (int kb = 0; kb < n_kb; ++kb) {
            const bool more = kb + 1 < n_kb;
            if (more) gload(src + (size_t)(kb + 1) * m.tb);
            block_mfma<NE, RA, A0, A1, RB, B0, B1>((kb & 1) ? lds1 : lds0, lrow, kg, ar, ai, br, bi);
            if (more) {
                if (chk) check();
                st.store((kb & 1) ? lds0 : lds1, ident, ltid);
            }
            __syncthreads();
        }
    } else {
        constexpr int KB = StageOf<V, NBLK>::KB;
        constexpr int NW = NTH / 64, NWIN = NBLK * 2;  // 1 KiB windows per time block
        const int w = ltid >> 6, el = lane >> 1;
        // lane l fills unit l of a window: element el, k-half (l & 1) ^ swizzle(el)
        const int hsw = ((lane & 1) ^ ((el >> 2) ^ (el >> 3))) & 1;
        const int n_st = (n_kb + KB - 1) / KB;
        auto dma = [&](uint8_t* L, int stg) {
            const int kb0 = stg * KB;
            const int kbs = n_kb - kb0 < KB ? n_kb - kb0 : KB;
            for (int win = w; win < kbs * NWIN; win += NW) {
                const int kbl = win / NWIN, r = win % NWIN, p = r / NBLK, b = r % NBLK;
                const int8_t* g = src + (size_t)(kb0 + kbl) * m.tb + m.off(p, gblk(b) * 32 + el, hsw);
                uint8_t* l = L + kbl * (NE * 64) + p * (NE * 32) + b * 1024;
                __builtin_amdgcn_global_load_lds((const void*)g, MIXER_LDS_PTR(l), 16, 0, 0);
            }
        };
        auto compute = [&](const uint8_t* L, int stg) {
            const int kbs = n_kb - stg * KB < KB ? n_kb - stg * KB : KB;
            static_for<0, KB>([&](auto kc) {
                constexpr int kbl = decltype(kc)::value;
                if (kbl < kbs) {
                    block_mfma<NE, RA, A0, A1, RB, B0, B1>(L + kbl * (NE * 64), lrow, kg, ar, ai, br, bi);
                    if (chk) {  // imaginary -128: scan the staged imaginary plane (NE * 32 bytes)
                        const uint8_t* im = L + kbl * (NE * 64) + NE * 32;
                        for (int u = ltid; u < NE * 2; u += NTH) {
                            const v4i v = *(const v4i*)(im + u * 16);
                            saw |= has80(v.x) | has80(v.y) | has80(v.z) | has80(v.w);
                        }
                    }
                }
            });
        };
        dma(lds0, 0);
        __syncthreads();  // waits for this wave's DMA (vmcnt) before the barrier
        for (int stg = 0; stg < n_st; stg += 2) {
            if (stg + 1 < n_st) dma(lds1, stg + 1);
            compute(lds0, stg);
            __syncthreads();
            if (stg + 1 >= n_st) break;
            if (stg + 2 < n_st) dma(lds0, stg + 2);
            compute(lds1, stg + 1);
            __syncthreads();
        }
    }
}

// Where a wave puts a finished 32x32 tile (I, J) of its channel (I >= J, element blocks): the
// default writes the tile-packed int32 accumulator (tile t = I(I+1)/2 + J, 2048 ints; store or
// add); the fused integrate epilogue (corr_epi.h) writes V and grids it instead.
struct AccOut {
    int32_t* acc;
    int accumulate;
    __device__ __forceinline__ void operator()(const v16i& vr, const v16i& vi, int I, int J, int lrow,
                                               int kg) const {
        int32_t* t = acc + ((size_t)I * (I + 1) / 2 + J) * 2048;
        store_tile(vr, t, lrow, kg, accumulate);
        store_tile(vi, t + 1024, lrow, kg, accumulate);
    }
    // workgroup-wide steps around a wave program's tile outputs (every wave of the group calls
    // both, after its time loop): nothing to do for the accumulator
    __device__ __forceinline__ void begin() const {}
    __device__ __forceinline__ void end() const {}
};

// FOLD wave program: tiles (RA, [A0,A1)) u (RB, [B0,B1)) of the lower triangle of NB blocks
// starting at element block d0 (0 for a whole array; the diagonal triangles of the BLOCKED
// decomposition use d0 = 8k); acc points at the channel's tile-packed accumulator (tile
// t = I(I+1)/2 + J, 2048 ints).
// FOLD wave program in the 16x16x64 shape (V = 6): two-block stages staged through registers
// one stage ahead into two LDS stage buffers, one barrier per stage.
template <int NB, int NTH, int RA, int A0, int A1, int RB, int B0, int B1>
__device__ __forceinline__ void fold_prog16(const int8_t* __restrict__ src, const XSrc& m,
                                            int32_t* __restrict__ acc, uint8_t* lds0, uint8_t* lds1,
                                            int n_kb, int accumulate, int ltid, bool chk,
                                            unsigned& saw, int d0) {
    constexpr int NE = NB * 32;
    constexpr int NA = A1 - A0;
    constexpr int NBT = RB >= 0 ? B1 - B0 : 1;
    const int lane = threadIdx.x & 63;
    v4i ar[NA][4], ai[NA][4], br[NBT][4], bi[NBT][4];
#pragma unroll
    for (int j = 0; j < NA; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) { ar[j][q] = (v4i){0, 0, 0, 0}; ai[j][q] = (v4i){0, 0, 0, 0}; }
#pragma unroll
    for (int j = 0; j < NBT; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) { br[j][q] = (v4i){0, 0, 0, 0}; bi[j][q] = (v4i){0, 0, 0, 0}; }
    using St = Stager<NE, NB, NTH>;
    St s0, s1;  // the two blocks of the stage in flight
    auto gload = [&](St& st, const int8_t* sl) {
#pragma unroll
        for (int i = 0; i < St::PER; ++i) {
            const int u = ltid + i * NTH;
            if (St::CHUNKS % NTH == 0 || u < St::CHUNKS) {
                const int b = u >> 7, w = u & 127, p = w >> 6, eg = w & 63;
                st.r[i] = gld((const v4i*)(sl + m.off(p, d0 * 32 + b * 32 + (eg >> 1), eg & 1)));
            }
        }
    };
    auto check = [&](const St& st) {
#pragma unroll
        for (int i = 0; i < St::PER; ++i) {
            const int u = ltid + i * NTH;
            if ((St::CHUNKS % NTH == 0 || u < St::CHUNKS) && ((u >> 6) & 1))
                saw |= has80(st.r[i].x) | has80(st.r[i].y) | has80(st.r[i].z) | has80(st.r[i].w);
        }
    };
    // block slab L of the stage; zero: the stage's second block is past the end (zeros)
    auto sstore = [&](const St& st, uint8_t* L, bool zero) {
#pragma unroll
        for (int i = 0; i < St::PER; ++i) {
            const int u = ltid + i * NTH;
            if (St::CHUNKS % NTH == 0 || u < St::CHUNKS) {
                const int b = u >> 7, w = u & 127, p = w >> 6, eg = w & 63;
                const int e = b * 32 + (eg >> 1);
                *(v4i*)(L + (lds_unit<NE>(p, e, eg & 1) << 4)) = zero ? (v4i){0, 0, 0, 0} : st.r[i];
            }
        }
    };
    const int n_st = (n_kb + 1) / 2;
    const int8_t* lastp = src + (size_t)(n_kb - 1) * m.tb;
    gload(s0, src);
    gload(s1, n_kb > 1 ? src + m.tb : lastp);
    if (chk) { check(s0); if (n_kb > 1) check(s1); }
    sstore(s0, lds0, false);
    sstore(s1, lds0 + NE * 64, n_kb < 2);
    __syncthreads();
    for (int st = 0; st < n_st; ++st) {
        uint8_t* cur = (st & 1) ? lds1 : lds0;
        uint8_t* nxt = (st & 1) ? lds0 : lds1;
        const int kb = 2 * (st + 1);
        const bool more = st + 1 < n_st;
        if (more) {
            gload(s0, src + (size_t)kb * m.tb);
            gload(s1, kb + 1 < n_kb ? src + (size_t)(kb + 1) * m.tb : lastp);
        }
        block_mfma16<NE, RA, A0, A1, RB, B0, B1, NA, NBT>(cur, lane, ar, ai, br, bi);
        if (more) {
            if (chk) { check(s0); if (kb + 1 < n_kb) check(s1); }
            sstore(s0, nxt, false);
            sstore(s1, nxt + NE * 64, kb + 1 >= n_kb);
        }
        __syncthreads();
    }
    const int IA = d0 + RA, IB = d0 + RB;
#pragma unroll
    for (int j = 0; j < NA; ++j) {
        int32_t* t = acc + ((size_t)IA * (IA + 1) / 2 + d0 + A0 + j) * 2048;
        store_tile16(ar[j], t, lane, accumulate);
        store_tile16(ai[j], t + 1024, lane, accumulate);
    }
    if constexpr (RB >= 0) {
#pragma unroll
        for (int j = 0; j < NBT; ++j) {
            int32_t* t = acc + ((size_t)IB * (IB + 1) / 2 + d0 + B0 + j) * 2048;
            store_tile16(br[j], t, lane, accumulate);
            store_tile16(bi[j], t + 1024, lane, accumulate);
        }
    }
}

template <int V, int NB, int NTH, int RA, int A0, int A1, int RB, int B0, int B1, typename OUT>
__device__ __forceinline__ void fold_prog32(const int8_t* __restrict__ src, const XSrc& m,
                                            const OUT& out, uint8_t* lds0, uint8_t* lds1, int n_kb,
                                            int ltid, bool chk, unsigned& saw, int d0) {
    constexpr int NA = A1 - A0;
    constexpr int NBT = RB >= 0 ? B1 - B0 : 1;
    const int lane = threadIdx.x & 63, lrow = lane & 31, kg = lane >> 5;
    v16i ar[NA], ai[NA], br[NBT], bi[NBT];
#pragma unroll
    for (int j = 0; j < NA; ++j) { ar[j] = (v16i){0}; ai[j] = (v16i){0}; }
#pragma unroll
    for (int j = 0; j < NBT; ++j) { br[j] = (v16i){0}; bi[j] = (v16i){0}; }
    corr_wave<V, NB * 32, NB, NTH, RA, A0, A1, RB, B0, B1>(src, m, lds0, lds1, n_kb,
                                                           [=](int b) { return d0 + b; }, ar, ai,
                                                           br, bi, ltid, chk, saw);
    const int IA = d0 + RA, IB = d0 + RB;
    out.begin();
#pragma unroll
    for (int j = 0; j < NA; ++j) out(ar[j], ai[j], IA, d0 + A0 + j, lrow, kg);
    if constexpr (RB >= 0) {
#pragma unroll
        for (int j = 0; j < NBT; ++j) out(br[j], bi[j], IB, d0 + B0 + j, lrow, kg);
    }
    out.end();
}

template <int V, int NB, int NTH, int RA, int A0, int A1, int RB, int B0, int B1, typename OUT>
__device__ __forceinline__ void fold_prog(const int8_t* __restrict__ src, const XSrc& m,
                                          const OUT& out, uint8_t* lds0, uint8_t* lds1, int n_kb,
                                          int ltid, bool chk, unsigned& saw, int d0) {
    if constexpr (V == 6) {
        static_assert(std::is_same<OUT, AccOut>::value, "the 16x16x64 programs store accumulators");
        fold_prog16<NB, NTH, RA, A0, A1, RB, B0, B1>(src, m, out.acc, lds0, lds1, n_kb,
                                                     out.accumulate, ltid, chk, saw, d0);
    } else {
        fold_prog32<V, NB, NTH, RA, A0, A1, RB, B0, B1>(src, m, out, lds0, lds1, n_kb, ltid, chk,
                                                        saw, d0);
    }
}

// Wave layout of the FOLD correlator. Row pair p = (p, NB-1-p) holds T_p = NB+1 tiles (the
// middle row of an odd NB holds (NB+1)/2). NB < SPLIT_FROM: one wave per pair (up to 6 tiles,
// which needs one wave per SIMD); NB >= SPLIT_FROM (6 by default, 2 in the batched persistent
// kernel that runs two waves per SIMD): two waves per pair, split at column S = ceil(T_p/2) of the far row: wave A = (R2, [0,S)), wave B =
// (R1, [0,R1]) u (R2, [S,R2]); waves 0..P-1 are the A halves, P..2P-1 the B halves, so with
// waves placed on SIMD (w mod 4) each SIMD of an 8-block array gets 5 + 4 = 9 tiles.
template <int NB, int SPLIT_FROM = 6>
struct FoldLayout {
    static constexpr int P = (NB + 1) / 2;
    static constexpr bool SPLIT = NB >= SPLIT_FROM;
    static constexpr int WAVES = SPLIT ? 2 * P : P;
    static constexpr int NTH = 64 * WAVES;
};

template <int V, int NB, int W, int SPLIT_FROM = 6, typename OUT>
__device__ __forceinline__ void fold_run(const int8_t* src, const XSrc& m, const OUT& out,
                                         uint8_t* lds0, uint8_t* lds1, int n_kb, int ltid, bool chk,
                                         unsigned& saw, int d0) {
    using Lay = FoldLayout<NB, SPLIT_FROM>;
    constexpr int NTH = Lay::NTH;
    if constexpr (!Lay::SPLIT) {
        constexpr int R1 = W, R2 = NB - 1 - W;
        if constexpr (R1 == R2)
            fold_prog<V, NB, NTH, R1, 0, R1 + 1, -1, 0, 0>(src, m, out, lds0, lds1, n_kb, ltid, chk, saw, d0);
        else
            fold_prog<V, NB, NTH, R1, 0, R1 + 1, R2, 0, R2 + 1>(src, m, out, lds0, lds1, n_kb, ltid, chk, saw, d0);
    } else {
        constexpr int p = W % Lay::P;
        constexpr int R1 = p, R2 = NB - 1 - p;
        constexpr int T = (R1 == R2) ? R1 + 1 : R1 + R2 + 2;
        constexpr int S = (T + 1) / 2;
        if constexpr (W < Lay::P) {
            fold_prog<V, NB, NTH, R2, 0, S, -1, 0, 0>(src, m, out, lds0, lds1, n_kb, ltid, chk, saw, d0);
        } else if constexpr (R1 == R2) {
            fold_prog<V, NB, NTH, R2, S, R2 + 1, -1, 0, 0>(src, m, out, lds0, lds1, n_kb, ltid, chk, saw, d0);
        } else if constexpr (S > R2) {
            fold_prog<V, NB, NTH, R1, 0, R1 + 1, -1, 0, 0>(src, m, out, lds0, lds1, n_kb, ltid, chk, saw, d0);
        } else {
            fold_prog<V, NB, NTH, R1, 0, R1 + 1, R2, S, R2 + 1>(src, m, out, lds0, lds1, n_kb, ltid, chk, saw, d0);
        }
    }
}

// LDS bytes of one stage buffer of the FOLD correlator
template <int NB, int V>
struct FoldLds {
    static constexpr int BYTES = StageOf<V, NB>::KB * NB * 32 * 64;
};

// One channel (or time slot of it) of the FOLD correlator on the FoldLayout<NB>::NTH threads
// with local ids ltid in [0, NTH) (the group's wave w = ltid / 64 runs wave program w).
template <int NB, int V, int SPLIT_FROM = 6, typename OUT>
__device__ __forceinline__ void fold_item(const int8_t* src, const XSrc& m, const OUT& out,
                                          uint8_t* lds0, uint8_t* lds1, int n_kb, int ltid, bool chk,
                                          unsigned& saw, int d0 = 0) {
    const int w = ltid >> 6;
    static_for<0, FoldLayout<NB, SPLIT_FROM>::WAVES>([&](auto Wc) {
        constexpr int W = decltype(Wc)::value;
        if (w == W)
            fold_run<V, NB, W, SPLIT_FROM>(src, m, out, lds0, lds1, n_kb, ltid, chk, saw, d0);
    });
}

// One increment of the -128 counter per workgroup that staged an imaginary -128.
__device__ __forceinline__ void count_wrap(uint32_t* wrap, unsigned saw) {
    if (!wrap) return;
    const int any = __syncthreads_or(saw ? 1 : 0);
    if (threadIdx.x == 0 && any) atomicAdd(wrap, 1u);
}

// Workgroup b of a grid padded to 8 * ceil(total / 8): dispatch is round-robin over the 8 XCDs,
// so b -> (b % 8) * per_xcd + b / 8 gives each XCD a contiguous range of logical work items
// (a bijection onto [0, 8 * per_xcd); items >= total are idle).
__device__ __forceinline__ int xcd_logical(int b, int total) {
    const int per_xcd = (total + 7) / 8;
    return (b % 8) * per_xcd + b / 8;
}

// FOLD correlator (arrays of NB <= 8 blocks of 32 elements): one workgroup per (time slot,
// channel), consecutive channels of a slot on one XCD (4 channels share each 128-byte line of
// a heap plane, so their workgroups share the L2 when staging straight from the heaps).
// The whole channel slab (NB blocks) is staged per stage into one of two LDS buffers.
template <int NB, int V>
__global__ __launch_bounds__(FoldLayout<NB>::NTH)
void k_xcorr_fold(XSrc xs, int32_t* __restrict__ acc, uint32_t* __restrict__ wrap, int C,
                  int n_tb, int n_slots, int accumulate) {
    constexpr int NT = NB * (NB + 1) / 2;
    __shared__ __attribute__((aligned(16))) uint8_t lds0[FoldLds<NB, V>::BYTES];
    __shared__ __attribute__((aligned(16))) uint8_t lds1[FoldLds<NB, V>::BYTES];
    const int logical = xcd_logical(blockIdx.x, C * n_slots);
    if (logical >= C * n_slots) return;  // uniform over the workgroup
    const int c = logical % C, slot = logical / C;
    const int kb0 = (int)((long)n_tb * slot / n_slots);
    const int kb1 = (int)((long)n_tb * (slot + 1) / n_slots);
    if (kb1 <= kb0) return;
    int32_t* a = acc + ((size_t)slot * C + c) * NT * 2048;
    const int8_t* src = xs.base + (size_t)c * xs.chan + (size_t)kb0 * xs.tb;
    unsigned saw = 0;
    fold_item<NB, V>(src, xs, AccOut{a, accumulate}, lds0, lds1, kb1 - kb0, threadIdx.x,
                     wrap != nullptr, saw);
    count_wrap(wrap, saw);
}

// ------------------------------------------------------------------------------------------
// BLOCKED correlator (arrays of NB = 12..32 blocks, NB % 4 == 0). A workgroup owns a
// super-block of 4 tile rows x 8 tile columns of the lower triangle (rows I0..I0+3, I0 % 4 == 0;
// columns J0..J0+7, J0 % 8 == 0); wave w (8 waves) owns the 2x2 tile block rows
// I0 + 2(w/4) + {0,1}, columns J0 + 2(w%4) + {0,1}. With even row and column starts a wave
// block is FULL (4 tiles), DIAG (3 tiles) or EMPTY, fixed per wave, so every path has a
// compile-time tile set. The 12 row/column element blocks are staged per stage into LDS
// (slots 0..3 rows, 4..11 columns). Super-blocks of one channel are mapped to one XCD
// (dispatch is round-robin over the 8 XCDs) so the channel's slab is read from HBM about once
// per XCD L2.
// ------------------------------------------------------------------------------------------
template <int V, int PAT, int RS, int CS, typename GB, typename OUT>
__device__ __forceinline__ void blk_prog(const int8_t* __restrict__ src, const XSrc& m,
                                         const OUT& out, uint8_t* lds0, uint8_t* lds1, int n_kb,
                                         GB gblk, int Ib, int Jb, int ltid, bool chk, unsigned& saw) {
    // PAT 0 FULL: rows (RS, RS+1) x columns (CS, CS+1); 1 DIAG: (0,0), (1,0), (1,1); 2 EMPTY
    const int lane = threadIdx.x & 63, lrow = lane & 31, kg = lane >> 5;
    v16i ar[2], ai[2], br[2], bi[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) { ar[j] = (v16i){0}; ai[j] = (v16i){0}; br[j] = (v16i){0}; bi[j] = (v16i){0}; }
    if constexpr (PAT == 0)
        corr_wave<V, 12 * 32, 12, 512, RS, CS, CS + 2, RS + 1, CS, CS + 2>(src, m, lds0, lds1, n_kb, gblk, ar, ai, br, bi, ltid, chk, saw);
    else if constexpr (PAT == 1)
        corr_wave<V, 12 * 32, 12, 512, RS, CS, CS + 1, RS + 1, CS, CS + 2>(src, m, lds0, lds1, n_kb, gblk, ar, ai, br, bi, ltid, chk, saw);
    else
        corr_wave<V, 12 * 32, 12, 512, RS, CS, CS, -1, 0, 0>(src, m, lds0, lds1, n_kb, gblk, ar, ai, br, bi, ltid, chk, saw);
    out.begin();
    if constexpr (PAT != 2) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (PAT == 1 && i == 0 && j == 1) continue;
                out(i == 0 ? ar[j] : br[j], i == 0 ? ai[j] : bi[j], Ib + i, Jb + j, lrow, kg);
            }
    }
    out.end();
}

template <int V>
struct BlkLds {
    static constexpr int BYTES = StageOf<V, 12>::KB * 12 * 32 * 64;
};

// One work item (I0, J0) of the BLOCKED correlator on 512 threads with local ids ltid:
// J0 < 0 is the diagonal triangle of element blocks I0..I0+7, run with the 8-wave FOLD layout
// (36 tiles, no idle wave); J0 >= 0 is the 4x8-tile super-block at rows I0.., columns J0.. .
// nb (the array's element blocks) only clamps the staged column blocks, so the batched kernel
// runs the same wave programs for every array size.
template <int V, typename OUT>
__device__ __forceinline__ void blk_item_n(const int8_t* src, const XSrc& m, const OUT& out,
                                           uint8_t* lds0, uint8_t* lds1, int n_kb, int I0, int J0,
                                           int nb, int ltid, bool chk, unsigned& saw) {
    static_assert(FoldLds<8, V>::BYTES <= BlkLds<V>::BYTES, "fold stage must fit the blocked LDS");
    if (J0 < 0) {
        fold_item<8, V>(src, m, out, lds0, lds1, n_kb, ltid, chk, saw, I0);
        return;
    }
    auto gblk = [=](int s) {
        const int g = s < 4 ? I0 + s : J0 + (s - 4);
        return g < nb ? g : nb - 1;  // columns past the array only feed EMPTY waves
    };
    const int w = ltid >> 6;
    static_for<0, 8>([&](auto Wc) {
        constexpr int W = decltype(Wc)::value;
        constexpr int RS = 2 * (W / 4), CS = 4 + 2 * (W % 4);
        if (w == W) {
            const int Ib = I0 + RS, Jb = J0 + CS - 4;
            if (Jb + 1 < Ib)
                blk_prog<V, 0, RS, CS>(src, m, out, lds0, lds1, n_kb, gblk, Ib, Jb, ltid, chk, saw);
            else if (Jb == Ib)
                blk_prog<V, 1, RS, CS>(src, m, out, lds0, lds1, n_kb, gblk, Ib, Jb, ltid, chk, saw);
            else
                blk_prog<V, 2, RS, CS>(src, m, out, lds0, lds1, n_kb, gblk, Ib, Jb, ltid, chk, saw);
        }
    });
}

template <int NB, int V, typename OUT>
__device__ __forceinline__ void blk_item(const int8_t* src, const XSrc& m, const OUT& out,
                                         uint8_t* lds0, uint8_t* lds1, int n_kb, int I0, int J0,
                                         int ltid, bool chk, unsigned& saw) {
    blk_item_n<V>(src, m, out, lds0, lds1, n_kb, I0, J0, NB, ltid, chk, saw);
}

template <int NB, int V>
__global__ __launch_bounds__(512)
void k_xcorr_blk(XSrc xs, int32_t* __restrict__ acc, uint32_t* __restrict__ wrap,
                 const int2* __restrict__ sblocks, int n_sb, int C, int n_tb, int n_slots,
                 int accumulate) {
    constexpr int NT = NB * (NB + 1) / 2;
    __shared__ __attribute__((aligned(16))) uint8_t lds0[BlkLds<V>::BYTES];
    __shared__ __attribute__((aligned(16))) uint8_t lds1[BlkLds<V>::BYTES];
    const int total = n_sb * C * n_slots;
    const int logical = xcd_logical(blockIdx.x, total);
    if (logical >= total) return;
    const int slot = logical / (n_sb * C);
    const int rem = logical % (n_sb * C);
    const int c = rem / n_sb;
    const int2 sb = sblocks[rem % n_sb];
    const int kb0 = (int)((long)n_tb * slot / n_slots);
    const int kb1 = (int)((long)n_tb * (slot + 1) / n_slots);
    if (kb1 <= kb0) return;
    const int8_t* src = xs.base + (size_t)c * xs.chan + (size_t)kb0 * xs.tb;
    int32_t* a = acc + ((size_t)slot * C + c) * NT * 2048;
    unsigned saw = 0;
    blk_item<NB, V>(src, xs, AccOut{a, accumulate}, lds0, lds1, kb1 - kb0, sb.x, sb.y, threadIdx.x,
                    wrap != nullptr, saw);
    count_wrap(wrap, saw);
}

// ------------------------------------------------------------------------------------------
// Launch helpers, instantiated per variant in corr_{fold,blk}_pf{1,2,4}.hip
// ------------------------------------------------------------------------------------------
template <int NB, int V>
inline hipError_t launch_fold_t(hipStream_t s, const XcorrPlan& p, const XSrc& src,
                                int32_t* acc, uint32_t* wrap, int C, int n_tb, int accumulate) {
    const int wgs = C * p.n_slots;
    hipLaunchKernelGGL((k_xcorr_fold<NB, V>), dim3((wgs + 7) / 8 * 8), dim3(FoldLayout<NB>::NTH),
                       0, s, src, acc, wrap, C, n_tb, p.n_slots, accumulate);
    return hipGetLastError();
}

template <int NB, int V>
inline hipError_t launch_blk_t(hipStream_t s, const XcorrPlan& p, const XSrc& src,
                               int32_t* acc, uint32_t* wrap, int C, int n_tb, int accumulate) {
    const int wgs = p.n_tasks * C * p.n_slots;
    hipLaunchKernelGGL((k_xcorr_blk<NB, V>), dim3((wgs + 7) / 8 * 8), dim3(512), 0, s, src, acc,
                       wrap, (const int2*)p.d_tasks, p.n_tasks, C, n_tb, p.n_slots, accumulate);
    return hipGetLastError();
}

template <int V>
inline hipError_t launch_fold_pf(hipStream_t s, const XcorrPlan& p, const XSrc& src,
                                 int32_t* acc, uint32_t* wrap, int C, int n_tb, int accumulate) {
    switch (p.n_elem / 32) {
    case 1: return launch_fold_t<1, V>(s, p, src, acc, wrap, C, n_tb, accumulate);
    case 2: return launch_fold_t<2, V>(s, p, src, acc, wrap, C, n_tb, accumulate);
    case 3: return launch_fold_t<3, V>(s, p, src, acc, wrap, C, n_tb, accumulate);
    case 4: return launch_fold_t<4, V>(s, p, src, acc, wrap, C, n_tb, accumulate);
    case 5: return launch_fold_t<5, V>(s, p, src, acc, wrap, C, n_tb, accumulate);
    case 6: return launch_fold_t<6, V>(s, p, src, acc, wrap, C, n_tb, accumulate);
    case 7: return launch_fold_t<7, V>(s, p, src, acc, wrap, C, n_tb, accumulate);
    case 8: return launch_fold_t<8, V>(s, p, src, acc, wrap, C, n_tb, accumulate);
    default: return hipErrorInvalidValue;
    }
}

template <int V>
inline hipError_t launch_blk_pf(hipStream_t s, const XcorrPlan& p, const XSrc& src,
                                int32_t* acc, uint32_t* wrap, int C, int n_tb, int accumulate) {
    switch (p.n_elem / 32) {
    case 12: return launch_blk_t<12, V>(s, p, src, acc, wrap, C, n_tb, accumulate);
    case 16: return launch_blk_t<16, V>(s, p, src, acc, wrap, C, n_tb, accumulate);
    case 20: return launch_blk_t<20, V>(s, p, src, acc, wrap, C, n_tb, accumulate);
    case 24: return launch_blk_t<24, V>(s, p, src, acc, wrap, C, n_tb, accumulate);
    case 28: return launch_blk_t<28, V>(s, p, src, acc, wrap, C, n_tb, accumulate);
    case 32: return launch_blk_t<32, V>(s, p, src, acc, wrap, C, n_tb, accumulate);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace mixer
