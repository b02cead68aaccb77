// kernels.hip -- the X-engine hot path as hand-written CDNA4 (gfx950) kernels.
//
//   k_corner_turn   heaps [tb][e][re C*32 | im C*32]  ->  tiles [c][tb][re|im][e][32]
//                   (replaces k_packAB, src/rx_host_corr.cu:56-72)
//   k_xcorr_fold    lower-triangle complex GEMM V += X X^H on v_mfma_i32_32x32x32_i8, int32
//                   accumulation, tile-packed accumulator (replaces cgemm_blk_t<ACC>,
//                   src/cgemm.cuh:84-125)
//   k_xcorr_blk     the same for large arrays (> 8 blocks of 32 elements): 2x4 tile blocks
//   k_integrate     int32 slots -> f32 full-square cube (replaces k_flush_cube,
//                   src/rx_host_corr.cu:86-95)
//   k_image         exact integer gridding + ng x ng FP32 Stockham FFT in LDS + fftshift/Re/
//                   min/max (replaces k_diff/k_grid_cube/cuFFT/host publish,
//                   src/rx_host_corr.cu:74-85,205-206,219-235)
//   k_xcorr_plain   one 32x32 full-square tile per wave, reference math, per-channel launch
//                   (the baseline form)
//
// MFMA lane layout (checked on gfx950 hardware with a single-instruction probe):
//   v_mfma_i32_32x32x32_i8: A lane l holds row l%32, k = 16*(l/32)..+15 (16 bytes);
//   B lane l holds column l%32 with the same k; D lane l register r holds
//   row 8*(r/4) + 4*(l/32) + r%4, column l%32.
// Because A and B share the k mapping, loading both operands from the same element x time
// tile gives sum_k x_i[k] y_j[k] whatever the hardware's k order inside a lane.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "kernels.h"

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

namespace mixer {

// Per-byte two's-complement negation of 4 packed int8 (wraps -128 to -128 like the
// reference's (int8)(-im), src/rx_host_corr.cu:71 and src/cgemm.cuh:108).
__device__ __forceinline__ int neg8x4(int w) {
    unsigned nw = ~(unsigned)w;
    return (int)(((nw & 0x7f7f7f7fu) + 0x01010101u) ^ (nw & 0x80808080u));
}
__device__ __forceinline__ v4i neg8x16(v4i v) {
    v4i r;
    r.x = neg8x4(v.x); r.y = neg8x4(v.y); r.z = neg8x4(v.z); r.w = neg8x4(v.w);
    return r;
}
// Nonzero iff some byte of w equals 0x80 (-128).
__device__ __forceinline__ unsigned has80(int w) {
    unsigned t = (unsigned)w ^ 0x80808080u;
    return (t - 0x01010101u) & ~t & 0x80808080u;
}

__device__ __forceinline__ v16i mfma(v4i a, v4i b, v16i c) {
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------------------------------------
// Corner-turn. One workgroup = one time block x 32 elements x 32 channels (64 KiB of LDS).
// Load: each wave reads whole 1 KiB runs of one heap plane (coalesced, 16 B per lane).
// Store: each wave writes one 1 KiB row = 32 elements x 32 samples of one (channel, plane).
// LDS rows are XOR-swizzled in 16-byte units so both phases are bank-conflict free.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_corner_turn(const uint8_t* __restrict__ heaps,
                                                     int8_t* __restrict__ tiles,
                                                     uint32_t* __restrict__ wrap, int N, int C,
                                                     int n_tb, int c_base, int c_end) {
    __shared__ __attribute__((aligned(16))) uint8_t lds[64 * 1024];
    const int tb = blockIdx.x, e0 = blockIdx.y * 32, c0 = c_base + blockIdx.z * 32;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const size_t payload = (size_t)C * 64;
    const int c = lane >> 1, h = lane & 1;
    const bool cok = (c0 + c) < c_end;
    unsigned saw = 0;
#pragma unroll 4
    for (int q = wv; q < 64; q += 4) {
        const int e = q >> 1, p = q & 1;
        v4i v = {0, 0, 0, 0};
        if (cok)
            v = *(const v4i*)(heaps + ((size_t)tb * N + e0 + e) * payload + (size_t)p * C * 32 +
                              (size_t)(c0 + c) * 32 + h * 16);
        if (p) saw |= has80(v.x) | has80(v.y) | has80(v.z) | has80(v.w);
        const int s = (e * 2 + h) ^ ((c * 2) & 15);
        *(v4i*)(lds + (((p * 32 + c) * 64 + s) << 4)) = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int row = wv; row < 64; row += 4) {
        const int p = row >> 5, cc = row & 31;
        if (c0 + cc >= c_end) continue;
        const int s = lane ^ ((cc * 2) & 15);
        v4i v = *(const v4i*)(lds + (((p * 32 + cc) * 64 + s) << 4));
        *(v4i*)(tiles + ((((size_t)(c0 + cc) * n_tb + tb) * 2 + p) * N + e0) * 32 + lane * 16) = v;
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
// Correlator building blocks. A (channel, 32-sample block) slab of NE elements is staged into
// LDS as [plane][element][32 samples] in 16-byte units, unit(p, e, g) = p*2*NE + 2e + (g ^ bit2(e)):
// the fragment read of 32 consecutive elements x both k-halves then touches 64 distinct units
// spread over all banks, and the staging writes (64 consecutive units per wave) stay distinct.
// ------------------------------------------------------------------------------------------
template <int I, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < E) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, E>(f);
    }
}

template <int NE>
__device__ __forceinline__ int lds_unit(int p, int e, int g) {
    return p * 2 * NE + 2 * e + (g ^ ((e >> 2) & 1));
}

// Fragment of element block B (32 elements x 16 samples per lane) of plane p from LDS.
template <int NE>
__device__ __forceinline__ v4i frag(const uint8_t* lds, int p, int B, int lrow, int kg) {
    return *(const v4i*)(lds + (lds_unit<NE>(p, B * 32 + lrow, kg) << 4));
}

// Copies a list of element blocks of one slab global -> registers -> LDS. NBLK blocks, each
// 2 planes x 32 elements x 32 samples = 128 chunks of 16 B; NT threads.
template <int NE, int NBLK, int NT>
struct Stager {
    static constexpr int CHUNKS = NBLK * 128;
    static constexpr int PER = (CHUNKS + NT - 1) / NT;
    v4i r[PER];
    // blk(i) gives the element block of list entry i
    template <typename BF>
    __device__ __forceinline__ void load(const int8_t* slab, BF blk) {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int u = threadIdx.x + i * NT;
            if (CHUNKS % NT == 0 || u < CHUNKS) {
                const int b = u >> 7, w = u & 127, p = w >> 6, eg = w & 63;
                r[i] = *(const v4i*)(slab + (size_t)p * NE * 32 + blk(b) * 1024 + eg * 16);
            }
        }
    }
    template <typename BF>
    __device__ __forceinline__ void store(uint8_t* lds, BF blk) const {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int u = threadIdx.x + i * NT;
            if (CHUNKS % NT == 0 || u < CHUNKS) {
                const int b = u >> 7, w = u & 127, p = w >> 6, eg = w & 63;
                const int e = blk(b) * 32 + (eg >> 1), g = eg & 1;
                *(v4i*)(lds + (lds_unit<NE>(p, e, g) << 4)) = r[i];
            }
        }
    }
};

// Store one 32x32 int32 accumulator tile from MFMA registers to [32][32] row-major memory.
__device__ __forceinline__ void store_tile(const v16i& x, int32_t* __restrict__ dst, int lrow,
                                           int kg, int accumulate) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = 8 * (r >> 2) + 4 * kg + (r & 3);
        int32_t* p = dst + row * 32 + lrow;
        int v = x[r];
        if (accumulate) v += *p;
        *p = v;
    }
}

// The four MFMAs of one tile and one 32-sample step, in the stacked real form of the
// reference's four real products (cgemm.cuh:115-118):
//   Re += re_I . re_J + wrap(-im_I) . wrap(-im_J)      Im += re_I . wrap(-im_J) + im_I . re_J
struct Frag { v4i re, im, ng; };
__device__ __forceinline__ void tile_step(const Frag& a, const Frag& b, v16i& vr, v16i& vi) {
    vr = mfma(a.re, b.re, vr);
    vi = mfma(a.re, b.ng, vi);
    vr = mfma(a.ng, b.ng, vr);
    vi = mfma(a.im, b.re, vi);
}
template <int NE>
__device__ __forceinline__ Frag load_frag(const uint8_t* lds, int B, int lrow, int kg) {
    Frag f;
    f.re = frag<NE>(lds, 0, B, lrow, kg);
    f.im = frag<NE>(lds, 1, B, lrow, kg);
    f.ng = neg8x16(f.im);
    return f;
}

// ------------------------------------------------------------------------------------------
// One wave's share of the correlator: tiles (RA, J in [A0, A1)) and, when RB >= 0, tiles
// (RB, J in [B0, B1)). The staged LDS slab holds NE elements; block indices are LDS blocks.
// Runs the whole time loop: NTH threads stage step kb+1 (global -> registers) while the MFMAs
// of step kb run from LDS, then write it to the other LDS buffer; one barrier per step.
// GB maps a staged block slot to its element block in the global slab.
// ------------------------------------------------------------------------------------------
template <int NE, int NBLK, int NTH, int RA, int A0, int A1, int RB, int B0, int B1, typename GB>
__device__ __forceinline__ void corr_wave(const int8_t* __restrict__ src, size_t slab,
                                          uint8_t (*lds)[NE * 64], int n_kb, GB gblk, v16i* ar,
                                          v16i* ai, v16i* br, v16i* bi) {
    constexpr int JLO = (RB >= 0 && B0 < A0) ? B0 : A0;
    constexpr int JHI = (RB >= 0 && B1 > A1) ? B1 : A1;
    const int lane = threadIdx.x & 63, lrow = lane & 31, kg = lane >> 5;
    auto ident = [](int b) { return b; };
    Stager<NE, NBLK, NTH> st;
    auto gload = [&](const int8_t* sl) {
#pragma unroll
        for (int i = 0; i < decltype(st)::PER; ++i) {
            const int u = threadIdx.x + i * NTH;
            if (decltype(st)::CHUNKS % NTH == 0 || u < decltype(st)::CHUNKS) {
                const int b = u >> 7, w = u & 127, p = w >> 6, eg = w & 63;
                st.r[i] = *(const v4i*)(sl + (size_t)p * (slab >> 1) + gblk(b) * 1024 + eg * 16);
            }
        }
    };
    gload(src);
    st.store(lds[0], ident);
    __syncthreads();
    int cur = 0;
    for (int kb = 0; kb < n_kb; ++kb) {
        const bool more = kb + 1 < n_kb;
        if (more) gload(src + (size_t)(kb + 1) * slab);
        const uint8_t* L = lds[cur];
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
        if (more) st.store(lds[cur ^ 1], ident);
        __syncthreads();
        cur ^= 1;
    }
}

// FOLD wave program: tiles (RA, [A0,A1)) u (RB, [B0,B1)) of the lower triangle of NB blocks;
// acc points at the channel's tile-packed accumulator (tile t = I(I+1)/2 + J, 2048 ints).
template <int NB, int NTH, int RA, int A0, int A1, int RB, int B0, int B1>
__device__ __forceinline__ void fold_prog(const int8_t* __restrict__ src, int32_t* __restrict__ acc,
                                          uint8_t (*lds)[NB * 32 * 64], int n_kb, int accumulate) {
    constexpr int NA = A1 - A0;
    constexpr int NBT = RB >= 0 ? B1 - B0 : 1;
    const int lane = threadIdx.x & 63, lrow = lane & 31, kg = lane >> 5;
    v16i ar[NA], ai[NA], br[NBT], bi[NBT];
#pragma unroll
    for (int j = 0; j < NA; ++j) { ar[j] = (v16i){0}; ai[j] = (v16i){0}; }
#pragma unroll
    for (int j = 0; j < NBT; ++j) { br[j] = (v16i){0}; bi[j] = (v16i){0}; }
    corr_wave<NB * 32, NB, NTH, RA, A0, A1, RB, B0, B1>(src, (size_t)NB * 32 * 64, lds, n_kb,
                                                        [](int b) { return b; }, ar, ai, br, bi);
#pragma unroll
    for (int j = 0; j < NA; ++j) {
        int32_t* t = acc + (size_t)(RA * (RA + 1) / 2 + A0 + j) * 2048;
        store_tile(ar[j], t, lrow, kg, accumulate);
        store_tile(ai[j], t + 1024, lrow, kg, accumulate);
    }
    if constexpr (RB >= 0) {
#pragma unroll
        for (int j = 0; j < NBT; ++j) {
            int32_t* t = acc + (size_t)(RB * (RB + 1) / 2 + B0 + j) * 2048;
            store_tile(br[j], t, lrow, kg, accumulate);
            store_tile(bi[j], t + 1024, lrow, kg, accumulate);
        }
    }
}

// Wave layout of the FOLD correlator. Row pair p = (p, NB-1-p) holds T_p = NB+1 tiles (the
// middle row of an odd NB holds (NB+1)/2). NB <= 5: one wave per pair. NB >= 6: two waves per
// pair, split at column S = ceil(T_p/2) of the far row: wave A = (R2, [0,S)), wave B =
// (R1, [0,R1]) u (R2, [S,R2]); waves 0..P-1 are the A halves, P..2P-1 the B halves, so with
// waves placed on SIMD (w mod 4) each SIMD of an 8-block array gets 5 + 4 = 9 tiles.
template <int NB>
struct FoldLayout {
    static constexpr int P = (NB + 1) / 2;
    static constexpr bool SPLIT = NB >= 6;
    static constexpr int WAVES = SPLIT ? 2 * P : P;
    static constexpr int NTH = 64 * WAVES;
};

template <int NB, int W>
__device__ __forceinline__ void fold_run(const int8_t* src, int32_t* acc,
                                         uint8_t (*lds)[NB * 32 * 64], int n_kb, int accumulate) {
    using Lay = FoldLayout<NB>;
    constexpr int NTH = Lay::NTH;
    if constexpr (!Lay::SPLIT) {
        constexpr int R1 = W, R2 = NB - 1 - W;
        if constexpr (R1 == R2)
            fold_prog<NB, NTH, R1, 0, R1 + 1, -1, 0, 0>(src, acc, lds, n_kb, accumulate);
        else
            fold_prog<NB, NTH, R1, 0, R1 + 1, R2, 0, R2 + 1>(src, acc, lds, n_kb, accumulate);
    } else {
        constexpr int p = W % Lay::P;
        constexpr int R1 = p, R2 = NB - 1 - p;
        constexpr int T = (R1 == R2) ? R1 + 1 : R1 + R2 + 2;
        constexpr int S = (T + 1) / 2;
        if constexpr (W < Lay::P) {
            fold_prog<NB, NTH, R2, 0, S, -1, 0, 0>(src, acc, lds, n_kb, accumulate);
        } else if constexpr (R1 == R2) {
            fold_prog<NB, NTH, R2, S, R2 + 1, -1, 0, 0>(src, acc, lds, n_kb, accumulate);
        } else if constexpr (S > R2) {
            fold_prog<NB, NTH, R1, 0, R1 + 1, -1, 0, 0>(src, acc, lds, n_kb, accumulate);
        } else {
            fold_prog<NB, NTH, R1, 0, R1 + 1, R2, S, R2 + 1>(src, acc, lds, n_kb, accumulate);
        }
    }
}

// FOLD correlator (arrays of NB <= 8 blocks of 32 elements). grid: x = channel, y = time slot.
// The whole channel slab (NB blocks) is staged per 32-sample step into double-buffered LDS.
// Every wave runs its own copy of the time loop, meeting the others at one barrier per step.
template <int NB>
__global__ __launch_bounds__(FoldLayout<NB>::NTH)
void k_xcorr_fold(const int8_t* __restrict__ tiles, int32_t* __restrict__ acc, int C, int n_tb,
                  int n_slots, int accumulate) {
    constexpr int NT = NB * (NB + 1) / 2;
    __shared__ __attribute__((aligned(16))) uint8_t lds[2][NB * 32 * 64];
    const int c = blockIdx.x, slot = blockIdx.y;
    const int kb0 = (int)((long)n_tb * slot / n_slots);
    const int kb1 = (int)((long)n_tb * (slot + 1) / n_slots);
    if (kb1 <= kb0) return;  // uniform over the workgroup
    int32_t* a = acc + ((size_t)slot * C + c) * NT * 2048;
    const int8_t* src = tiles + ((size_t)c * n_tb + kb0) * (size_t)NB * 32 * 64;
    const int w = threadIdx.x >> 6;
    static_for<0, FoldLayout<NB>::WAVES>([&](auto Wc) {
        constexpr int W = decltype(Wc)::value;
        if (w == W) fold_run<NB, W>(src, a, lds, kb1 - kb0, accumulate);
    });
}

// ------------------------------------------------------------------------------------------
// BLOCKED correlator (arrays of NB = 12..32 blocks, NB % 4 == 0). A workgroup owns a
// super-block of 4 tile rows x 8 tile columns of the lower triangle (rows I0..I0+3, I0 % 4 == 0;
// columns J0..J0+7, J0 % 8 == 0); wave w (8 waves) owns the 2x2 tile block rows
// I0 + 2(w/4) + {0,1}, columns J0 + 2(w%4) + {0,1}. With even row and column starts a wave
// block is FULL (4 tiles), DIAG (3 tiles) or EMPTY, fixed per wave, so every path has a
// compile-time tile set. The 12 row/column element blocks are staged per 32-sample step into
// double-buffered LDS (slots 0..3 rows, 4..11 columns). Super-blocks of one channel are
// mapped to one XCD (dispatch is round-robin over the 8 XCDs) so the channel's slab is read
// from HBM about once per XCD L2.
// ------------------------------------------------------------------------------------------
template <int PAT, int RS, int CS, typename GB>
__device__ __forceinline__ void blk_prog(const int8_t* __restrict__ src, size_t slab,
                                         int32_t* __restrict__ acc, uint8_t (*lds)[12 * 32 * 64],
                                         int n_kb, GB gblk, int Ib, int Jb, int accumulate) {
    // PAT 0 FULL: rows (RS, RS+1) x columns (CS, CS+1); 1 DIAG: (0,0), (1,0), (1,1); 2 EMPTY
    const int lane = threadIdx.x & 63, lrow = lane & 31, kg = lane >> 5;
    v16i ar[2], ai[2], br[2], bi[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) { ar[j] = (v16i){0}; ai[j] = (v16i){0}; br[j] = (v16i){0}; bi[j] = (v16i){0}; }
    if constexpr (PAT == 0)
        corr_wave<12 * 32, 12, 512, RS, CS, CS + 2, RS + 1, CS, CS + 2>(src, slab, lds, n_kb, gblk, ar, ai, br, bi);
    else if constexpr (PAT == 1)
        corr_wave<12 * 32, 12, 512, RS, CS, CS + 1, RS + 1, CS, CS + 2>(src, slab, lds, n_kb, gblk, ar, ai, br, bi);
    else
        corr_wave<12 * 32, 12, 512, RS, CS, CS, -1, 0, 0>(src, slab, lds, n_kb, gblk, ar, ai, br, bi);
    if constexpr (PAT != 2) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (PAT == 1 && i == 0 && j == 1) continue;
                const int I = Ib + i, J = Jb + j;
                int32_t* t = acc + (size_t)(I * (I + 1) / 2 + J) * 2048;
                store_tile(i == 0 ? ar[j] : br[j], t, lrow, kg, accumulate);
                store_tile(i == 0 ? ai[j] : bi[j], t + 1024, lrow, kg, accumulate);
            }
    }
}

template <int NB>
__global__ __launch_bounds__(512)
void k_xcorr_blk(const int8_t* __restrict__ tiles, int32_t* __restrict__ acc,
                 const int2* __restrict__ sblocks, int n_sb, int C, int n_tb, int n_slots,
                 int accumulate) {
    constexpr int NE = NB * 32, NT = NB * (NB + 1) / 2;
    __shared__ __attribute__((aligned(16))) uint8_t lds[2][12 * 32 * 64];
    const int total = n_sb * C * n_slots;
    const int b = blockIdx.x;
    const int per_xcd = (total + 7) / 8;
    // the grid is padded to 8 * per_xcd, so b -> (b % 8) * per_xcd + b / 8 is a bijection
    const int logical = (b % 8) * per_xcd + b / 8;
    if (logical >= total) return;
    const int slot = logical / (n_sb * C);
    const int rem = logical % (n_sb * C);
    const int c = rem / n_sb;
    const int2 sb = sblocks[rem % n_sb];
    const int I0 = sb.x, J0 = sb.y;
    const int kb0 = (int)((long)n_tb * slot / n_slots);
    const int kb1 = (int)((long)n_tb * (slot + 1) / n_slots);
    if (kb1 <= kb0) return;
    const size_t slab = (size_t)NE * 64;
    const int8_t* src = tiles + ((size_t)c * n_tb + kb0) * slab;
    int32_t* a = acc + ((size_t)slot * C + c) * NT * 2048;
    auto gblk = [=](int s) {
        const int g = s < 4 ? I0 + s : J0 + (s - 4);
        return g < NB ? g : NB - 1;  // columns past the array only feed EMPTY waves
    };
    const int w = threadIdx.x >> 6;
    static_for<0, 8>([&](auto Wc) {
        constexpr int W = decltype(Wc)::value;
        constexpr int RS = 2 * (W / 4), CS = 4 + 2 * (W % 4);
        if (w == W) {
            const int Ib = I0 + RS, Jb = J0 + CS - 4;
            if (Jb + 1 < Ib)
                blk_prog<0, RS, CS>(src, slab, a, lds, kb1 - kb0, gblk, Ib, Jb, accumulate);
            else if (Jb == Ib)
                blk_prog<1, RS, CS>(src, slab, a, lds, kb1 - kb0, gblk, Ib, Jb, accumulate);
            else
                blk_prog<2, RS, CS>(src, slab, a, lds, kb1 - kb0, gblk, Ib, Jb, accumulate);
        }
    });
}

// ------------------------------------------------------------------------------------------
// Integrate: grid x = row block I, y = channel. Writes rows [32I, 32I+32) of both planes.
// Tiles above the diagonal come from tile (J, I) transposed and conjugated through LDS.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_integrate(const int32_t* __restrict__ acc,
                                                   float* __restrict__ v_store,
                                                   float* __restrict__ v_add, int N, int C,
                                                   int n_slots) {
    __shared__ int32_t lt[2][32][33];
    const int I = blockIdx.x, c = blockIdx.y, nb = N / 32, NT = nb * (nb + 1) / 2;
    const int tid = threadIdx.x, i = tid >> 3, j0 = (tid & 7) * 4;
    const size_t slot_stride = (size_t)C * NT * 2048;
    const int32_t* ac = acc + (size_t)c * NT * 2048;
    for (int J = 0; J < nb; ++J) {
        int vals[2][4];
        if (J <= I) {
            const int32_t* t = ac + (size_t)(I * (I + 1) / 2 + J) * 2048 + i * 32 + j0;
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                v4i s = *(const v4i*)(t + p * 1024);
                for (int sl = 1; sl < n_slots; ++sl) s += *(const v4i*)(t + sl * slot_stride + p * 1024);
                vals[p][0] = s.x; vals[p][1] = s.y; vals[p][2] = s.z; vals[p][3] = s.w;
            }
        } else {
            const int32_t* t = ac + (size_t)(J * (J + 1) / 2 + I) * 2048 + i * 32 + j0;
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                v4i s = *(const v4i*)(t + p * 1024);
                for (int sl = 1; sl < n_slots; ++sl) s += *(const v4i*)(t + sl * slot_stride + p * 1024);
                lt[p][i][j0 + 0] = s.x; lt[p][i][j0 + 1] = s.y;
                lt[p][i][j0 + 2] = s.z; lt[p][i][j0 + 3] = s.w;
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                vals[0][q] = lt[0][j0 + q][i];
                vals[1][q] = -lt[1][j0 + q][i];
            }
            __syncthreads();
        }
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

// ------------------------------------------------------------------------------------------
// Image: one workgroup per channel. (1) Exact gridding of V into int64 cells in LDS, the
// upper triangle added as conj(lower) (G(-d) = conj G(d)); (2) one conversion to f32;
// (3) forward unnormalised ng x ng FFT, Stockham radix-2 along x then y, ping-pong in LDS,
// twiddles tw[q] = exp(-2 pi i q / ng) from an FP64 host table; (4) fftshift, real part,
// per-channel max/min.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__global__ __launch_bounds__(256) void k_image(const int32_t* __restrict__ acc,
                                               const float2* __restrict__ tw,
                                               float* __restrict__ img, float* __restrict__ vmax,
                                               float* __restrict__ vmin, int N, int C,
                                               int n_slots, int grid_x, int ng) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int nn = ng * ng;
    unsigned long long* G = (unsigned long long*)smem;            // [nn][2]
    float2* F0 = (float2*)(smem + (size_t)nn * 16);               // [nn]
    float2* F1 = F0 + nn;                                         // [nn]
    __shared__ float red_mx[4], red_mn[4];
    const int c = blockIdx.x, tid = threadIdx.x;
    const int nb = N / 32, NT = nb * (nb + 1) / 2;
    const size_t slot_stride = (size_t)C * NT * 2048;
    for (int q = tid; q < 2 * nn; q += 256) G[q] = 0ull;
    __syncthreads();
    const int32_t* ac = acc + (size_t)c * NT * 2048;
    for (int I = 0; I < nb; ++I)
        for (int J = 0; J <= I; ++J) {
            const int32_t* t = ac + (size_t)(I * (I + 1) / 2 + J) * 2048;
            for (int e = tid; e < 1024; e += 256) {
                const int i = e >> 5, j = e & 31;
                int vr = t[e], vi = t[1024 + e];
                for (int sl = 1; sl < n_slots; ++sl) {
                    vr += t[sl * slot_stride + e];
                    vi += t[sl * slot_stride + 1024 + e];
                }
                const int a = I * 32 + i, b = J * 32 + j;
                const int pa = a % grid_x, qa = a / grid_x, pb = b % grid_x, qb = b / grid_x;
                int dy = qa - qb, dx = pa - pb;
                dy += (dy < 0) ? ng : 0;
                dx += (dx < 0) ? ng : 0;
                const int cell = dy * ng + dx;
                atomicAdd(&G[2 * cell], (unsigned long long)(long long)vr);
                atomicAdd(&G[2 * cell + 1], (unsigned long long)(long long)vi);
                if (I != J) {
                    const int cell2 = (dy ? ng - dy : 0) * ng + (dx ? ng - dx : 0);
                    atomicAdd(&G[2 * cell2], (unsigned long long)(long long)vr);
                    atomicAdd(&G[2 * cell2 + 1], (unsigned long long)(-(long long)vi));
                }
            }
        }
    __syncthreads();
    for (int q = tid; q < nn; q += 256)
        F0[q] = make_float2((float)(long long)G[2 * q], (float)(long long)G[2 * q + 1]);
    __syncthreads();
    float2* src = F0;
    float2* dst = F1;
    const int half = ng >> 1;
    // along x (contiguous), then along y (stride ng)
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

// ------------------------------------------------------------------------------------------
// Plain form: one wave per 32x32 output tile of the FULL square of one channel, four MFMAs
// per 32-sample step exactly as cgemm.cuh:115-118, written to V[N][N] (overwrite).
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_xcorr_plain(const int8_t* __restrict__ tiles,
                                                     int32_t* __restrict__ vre,
                                                     int32_t* __restrict__ vim, int N, int n_tb,
                                                     int c) {
    const int nb = N / 32;
    const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile >= nb * nb) return;
    const int I = tile / nb, J = tile % nb;
    const int lane = threadIdx.x & 63, lrow = lane & 31, kg = lane >> 5;
    v16i cr = (v16i){0}, ci = (v16i){0};
    const size_t plane = (size_t)N * 32;
    const int8_t* base = tiles + (size_t)c * n_tb * 2 * plane + lrow * 32 + kg * 16;
    for (int kb = 0; kb < n_tb; ++kb, base += 2 * plane) {
        const v4i ar = *(const v4i*)(base + I * 1024);
        const v4i ai = *(const v4i*)(base + plane + I * 1024);
        const v4i br = *(const v4i*)(base + J * 1024);
        const v4i bi = neg8x16(*(const v4i*)(base + plane + J * 1024));
        cr = mfma(ar, br, cr);
        cr = mfma(neg8x16(ai), bi, cr);
        ci = mfma(ar, bi, ci);
        ci = mfma(ai, br, ci);
    }
    const size_t o = (size_t)c * N * N;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = I * 32 + 8 * (r >> 2) + 4 * kg + (r & 3), col = J * 32 + lrow;
        vre[o + (size_t)row * N + col] = cr[r];
        vim[o + (size_t)row * N + col] = ci[r];
    }
}

// ------------------------------------------------------------------------------------------
// Host-side launchers
// ------------------------------------------------------------------------------------------
hipError_t launch_corner_turn(hipStream_t s, const uint8_t* heaps, int8_t* tiles, uint32_t* wrap,
                              int N, int C, int n_tb, int c_base, int c_count) {
    dim3 grid(n_tb, N / 32, (c_count + 31) / 32);
    hipLaunchKernelGGL(k_corner_turn, grid, dim3(256), 0, s, heaps, tiles, wrap, N, C, n_tb,
                       c_base, c_base + c_count);
    return hipGetLastError();
}

template <int NB>
static hipError_t launch_fold(hipStream_t s, const int8_t* tiles, int32_t* acc, int C, int n_tb,
                              int n_slots, int accumulate) {
    hipLaunchKernelGGL(k_xcorr_fold<NB>, dim3(C, n_slots), dim3(FoldLayout<NB>::NTH), 0, s, tiles,
                       acc, C, n_tb, n_slots, accumulate);
    return hipGetLastError();
}

hipError_t launch_xcorr(hipStream_t s, const XcorrPlan& plan, const int8_t* tiles, int32_t* acc,
                        int C, int n_tb, int accumulate) {
    const int nb = plan.n_elem / 32;
    switch (plan.kind) {
    case XcorrPlan::FOLD:
        switch (nb) {
        case 1: return launch_fold<1>(s, tiles, acc, C, n_tb, plan.n_slots, accumulate);
        case 2: return launch_fold<2>(s, tiles, acc, C, n_tb, plan.n_slots, accumulate);
        case 3: return launch_fold<3>(s, tiles, acc, C, n_tb, plan.n_slots, accumulate);
        case 4: return launch_fold<4>(s, tiles, acc, C, n_tb, plan.n_slots, accumulate);
        case 5: return launch_fold<5>(s, tiles, acc, C, n_tb, plan.n_slots, accumulate);
        case 6: return launch_fold<6>(s, tiles, acc, C, n_tb, plan.n_slots, accumulate);
        case 7: return launch_fold<7>(s, tiles, acc, C, n_tb, plan.n_slots, accumulate);
        case 8: return launch_fold<8>(s, tiles, acc, C, n_tb, plan.n_slots, accumulate);
        default: return hipErrorInvalidValue;
        }
    case XcorrPlan::BLOCKED: {
        const int wgs = plan.n_tasks * C * plan.n_slots;
        const dim3 grid((wgs + 7) / 8 * 8);
        const int2* sb = (const int2*)plan.d_tasks;
        switch (nb) {
        case 12: hipLaunchKernelGGL(k_xcorr_blk<12>, grid, dim3(512), 0, s, tiles, acc, sb, plan.n_tasks, C, n_tb, plan.n_slots, accumulate); break;
        case 16: hipLaunchKernelGGL(k_xcorr_blk<16>, grid, dim3(512), 0, s, tiles, acc, sb, plan.n_tasks, C, n_tb, plan.n_slots, accumulate); break;
        case 20: hipLaunchKernelGGL(k_xcorr_blk<20>, grid, dim3(512), 0, s, tiles, acc, sb, plan.n_tasks, C, n_tb, plan.n_slots, accumulate); break;
        case 24: hipLaunchKernelGGL(k_xcorr_blk<24>, grid, dim3(512), 0, s, tiles, acc, sb, plan.n_tasks, C, n_tb, plan.n_slots, accumulate); break;
        case 28: hipLaunchKernelGGL(k_xcorr_blk<28>, grid, dim3(512), 0, s, tiles, acc, sb, plan.n_tasks, C, n_tb, plan.n_slots, accumulate); break;
        case 32: hipLaunchKernelGGL(k_xcorr_blk<32>, grid, dim3(512), 0, s, tiles, acc, sb, plan.n_tasks, C, n_tb, plan.n_slots, accumulate); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    }
    return hipErrorInvalidValue;
}

hipError_t launch_integrate(hipStream_t s, const int32_t* acc, float* v_store, float* v_add, int N,
                            int C, int n_slots) {
    hipLaunchKernelGGL(k_integrate, dim3(N / 32, C), dim3(256), 0, s, acc, v_store, v_add, N, C,
                       n_slots);
    return hipGetLastError();
}

size_t image_lds_bytes(int ng) { return (size_t)ng * ng * 32; }

hipError_t launch_image(hipStream_t s, const int32_t* acc, const float2* tw, float* img,
                        float* vmax, float* vmin, int N, int C, int n_slots, int grid_x, int ng) {
    hipLaunchKernelGGL(k_image, dim3(C), dim3(256), image_lds_bytes(ng), s, acc, tw, img, vmax,
                       vmin, N, C, n_slots, grid_x, ng);
    return hipGetLastError();
}

hipError_t set_image_lds_limit(int ng) {
    return hipFuncSetAttribute((const void*)k_image, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)image_lds_bytes(ng));
}

hipError_t launch_xcorr_plain(hipStream_t s, const int8_t* tiles, int32_t* vre, int32_t* vim, int N,
                              int n_tb, int c) {
    const int nb = N / 32;
    hipLaunchKernelGGL(k_xcorr_plain, dim3((nb * nb + 3) / 4), dim3(256), 0, s, tiles, vre, vim, N,
                       n_tb, c);
    return hipGetLastError();
}

}  // namespace mixer
