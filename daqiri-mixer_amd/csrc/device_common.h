// device_common.h -- device helpers shared by the gfx950 kernels (int8 SWAR, MFMA wrapper,
// LDS slab staging, accumulator tile store).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v4f __attribute__((ext_vector_type(4)));

namespace mixer {


// Per-byte two's-complement negation of 4 packed int8 (wraps -128 to -128 like the
// reference's (int8)(-im), src/rx_host_corr.cu:71 and src/cgemm.cuh:52).
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
// The same test folded into an accumulator: a byte of w equal to 0x80 leaves its bit 7 set in
// ((w ^ 0x80808080) - 0x01010101) & w (the zero-byte test on w ^ 0x80808080, using
// ~(w ^ 0x80808080) & 0x80808080 = w & 0x80808080). Three operations (xor, add, and-or); the
// other bits are don't-cares, so the caller masks the accumulator with 0x80808080 once.
__device__ __forceinline__ unsigned acc80(unsigned s, int w) {
    return s | ((((unsigned)w ^ 0x80808080u) + 0xFEFEFEFFu) & (unsigned)w);
}

// MFMA lane layout (checked on gfx950 hardware with a single-instruction probe):
//   v_mfma_i32_32x32x32_i8: A lane l holds row l%32, k = 16*(l/32)..+15 (16 bytes);
//   B lane l holds column l%32 with the same k; D lane l register r holds
//   row 8*(r/4) + 4*(l/32) + r%4, column l%32.
// Because A and B share the k mapping, loading both operands from the same element x time
// tile gives sum_k x_i[k] y_j[k] whatever the hardware's k order inside a lane.
__device__ __forceinline__ v16i mfma(v4i a, v4i b, v16i c) {
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------------------------------------
// Correlator building blocks. A (channel, 32-sample block) slab of NE elements is staged into
// LDS as [plane][element][32 samples] in 16-byte units,
// unit(p, e, g) = p*2*NE + 2e + (g ^ bit2(e) ^ bit3(e)): in a fragment read (lane l: element
// l%32, k-half l/32) every 8 consecutive lanes hit 8 distinct 16-byte bank groups of a 128-byte
// row and every 16 consecutive lanes 16 distinct groups of a 256-byte row, and the staging
// writes (64 consecutive units per wave) stay distinct.
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
    return p * 2 * NE + 2 * e + (g ^ (((e >> 2) ^ (e >> 3)) & 1));
}

// Fragment of element block B (32 elements x 16 samples per lane) of plane p from LDS.
template <int NE>
__device__ __forceinline__ v4i frag(const uint8_t* lds, int p, int B, int lrow, int kg) {
    return *(const v4i*)(lds + (lds_unit<NE>(p, B * 32 + lrow, kg) << 4));
}

// Loads and stores of device (global) memory through a generic pointer. A pointer read from
// device memory (a job table) is generic to the compiler, and a plain access through it becomes a
// FLAT instruction, which counts on the LDS counter too: every wait for an LDS fragment read
// would then also wait for the in-flight global prefetch. Accessing through a global-address-
// space pointer gives global loads and stores (every pointer passed here is device memory).
template <typename T>
__device__ __forceinline__ T gld(const T* p) {
    return *(const __attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ __forceinline__ void gst(T* p, const T& v) {
    *(__attribute__((address_space(1))) T*)p = v;
}
// Buffer loads of one staged block: the block's base in a scalar descriptor and one 32-bit offset
// register per load (the same offsets for every block), where a global load keeps a 64-bit
// address per load. Every offset is below 2^31 (one time block of one array). Used for the QUAD
// staging of whole 128-byte lines (0.215 -> 0.205 ms at 256 x 128 x 8192); the gathers of
// 32-byte runs (fold, blocked) measured slower with them (1024 x 256 x 4096: 2.96 vs 2.85 ms)
// and stay on global loads.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t block_rsrc(const void* base) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, 0x7fffffff, 0x00020000);
}
__device__ __forceinline__ v4i bload(__amdgpu_buffer_rsrc_t rs, uint32_t off) {
    return __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off, 0, 0));
}
__device__ __forceinline__ void gadd(unsigned long long* p, unsigned long long v) {
    __hip_atomic_fetch_add((__attribute__((address_space(1))) unsigned long long*)p, v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// The image grid G holds one orientation of every baseline: cell q = (dy, dx) sums V_ab over the
// pairs a > b (row above column in V) with that cell, plus the autocorrelations V_aa at q = 0. The
// image kernel completes it to F(q) = G(q) + conj(G(-q)) for q != 0 and F(0) = G(0), since V_ba =
// conj(V_ab) lies at -q (exact in int64, and half the atomics of adding both orientations). A sum of
// pairs a > b (off_diag) that lands on q = 0 (only if the grid aliases) counts both orientations
// there at once: 2 Re.
__device__ __forceinline__ void grid_add_lower(unsigned long long* Gc, int ng, int dy, int dx,
                                               unsigned long long re, unsigned long long im,
                                               bool off_diag) {
    const int y = ((dy % ng) + ng) % ng, x = ((dx % ng) + ng) % ng;
    if (off_diag && y == 0 && x == 0) {
        gadd(&Gc[0], re + re);
        return;
    }
    gadd(&Gc[2 * (y * ng + x)], re);
    gadd(&Gc[2 * (y * ng + x) + 1], im);
}

// Copies a list of element blocks of one slab global -> registers -> LDS. NBLK blocks, each
// 2 planes x 32 elements x 32 samples = 128 chunks of 16 B; NT threads with local ids ltid.
template <int NE, int NBLK, int NT>
struct Stager {
    static constexpr int CHUNKS = NBLK * 128;
    static constexpr int PER = (CHUNKS + NT - 1) / NT;
    v4i r[PER];
    // blk(i) gives the element block of list entry i
    template <typename BF>
    __device__ __forceinline__ void load(const int8_t* slab, BF blk, int ltid) {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int u = ltid + i * NT;
            if (CHUNKS % NT == 0 || u < CHUNKS) {
                const int b = u >> 7, w = u & 127, p = w >> 6, eg = w & 63;
                r[i] = gld((const v4i*)(slab + (size_t)p * NE * 32 + blk(b) * 1024 + eg * 16));
            }
        }
    }
    template <typename BF>
    __device__ __forceinline__ void store(uint8_t* lds, BF blk, int ltid) const {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int u = ltid + i * NT;
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
        if (accumulate) v += gld(p);
        gst(p, v);
    }
}

// The four MFMAs of one tile and one 32-sample step, in the stacked real form of the
// reference's four real products (cgemm.cuh:59-62):
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
// The 16x16x64 tile shape. v_mfma_i32_16x16x64_i8 (checked on gfx950 with a single-instruction
// probe): A lane l holds row l%16, k = 16*(l/16)..+15; B lane l holds column l%16 with the same k;
// D lane l register r holds row 4*(l/16) + r, column l%16. One instruction contracts 64 samples,
// two 32-sample blocks, so a stage holds two blocks [kb][plane][element][32] (each block's slab
// swizzled as lds_unit<NE>) and lane l reads block (l/16)/2, 16-byte half (l/16)%2. A 32x32 tile
// of the accumulator is four 16x16 sub-tiles s = 2*hr + hc (rows 16*hr.., columns 16*hc..).
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ v4i mfma16(v4i a, v4i b, v4i c) {
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// Rows 16*h .. 16*h+15 of element block B for the 64 samples of a two-block stage at L.
template <int NE>
__device__ __forceinline__ Frag load_frag16(const uint8_t* L, int B, int h, int lane) {
    const int e = B * 32 + 16 * h + (lane & 15), g = lane >> 4;
    const uint8_t* blk = L + (g >> 1) * (NE * 64);
    Frag f;
    f.re = *(const v4i*)(blk + (lds_unit<NE>(0, e, g & 1) << 4));
    f.im = *(const v4i*)(blk + (lds_unit<NE>(1, e, g & 1) << 4));
    f.ng = neg8x16(f.im);
    return f;
}

// One 32x32 tile over a 64-sample stage: rows a[hr], columns b[hc], the same four real products
// as tile_step on each sub-tile.
__device__ __forceinline__ void tile_step16(const Frag (&a)[2], const Frag (&b)[2], v4i (&vr)[4],
                                            v4i (&vi)[4]) {
#pragma unroll
    for (int hr = 0; hr < 2; ++hr)
#pragma unroll
        for (int hc = 0; hc < 2; ++hc) {
            const int s = 2 * hr + hc;
            vr[s] = mfma16(a[hr].re, b[hc].re, vr[s]);
            vi[s] = mfma16(a[hr].re, b[hc].ng, vi[s]);
            vr[s] = mfma16(a[hr].ng, b[hc].ng, vr[s]);
            vi[s] = mfma16(a[hr].im, b[hc].re, vi[s]);
        }
}

// Half HC (columns 16*HC..) of one 32x32 tile over a 64-sample stage: rows a[hr], columns b.
template <int HC>
__device__ __forceinline__ void tile_half16(const Frag (&a)[2], const Frag& b, v4i (&vr)[4], v4i (&vi)[4]) {
#pragma unroll
    for (int hr = 0; hr < 2; ++hr) {
        const int s = 2 * hr + HC;
        vr[s] = mfma16(a[hr].re, b.re, vr[s]);
        vi[s] = mfma16(a[hr].re, b.ng, vi[s]);
        vr[s] = mfma16(a[hr].ng, b.ng, vr[s]);
        vi[s] = mfma16(a[hr].im, b.re, vi[s]);
    }
}

// The same on a tile held as one v16i (sub-tile s in registers 4s..4s+3).
template <int S>
__device__ __forceinline__ void mfma16_sub(v16i& acc, v4i a, v4i b) {
    v4i c = {acc[4 * S], acc[4 * S + 1], acc[4 * S + 2], acc[4 * S + 3]};
    c = mfma16(a, b, c);
    acc[4 * S] = c.x;
    acc[4 * S + 1] = c.y;
    acc[4 * S + 2] = c.z;
    acc[4 * S + 3] = c.w;
}
template <int HC>
__device__ __forceinline__ void tile_half16v(const Frag (&a)[2], const Frag& b, v16i& vr, v16i& vi) {
    mfma16_sub<HC>(vr, a[0].re, b.re);
    mfma16_sub<HC>(vi, a[0].re, b.ng);
    mfma16_sub<HC>(vr, a[0].ng, b.ng);
    mfma16_sub<HC>(vi, a[0].im, b.re);
    mfma16_sub<2 + HC>(vr, a[1].re, b.re);
    mfma16_sub<2 + HC>(vi, a[1].re, b.ng);
    mfma16_sub<2 + HC>(vr, a[1].ng, b.ng);
    mfma16_sub<2 + HC>(vi, a[1].im, b.re);
}
__device__ __forceinline__ void store_tile16v(const v16i& x, int32_t* __restrict__ dst, int lane,
                                              int accumulate) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int s = r >> 2, q = r & 3;
        const int row = 16 * (s >> 1) + 4 * (lane >> 4) + q, col = 16 * (s & 1) + (lane & 15);
        int32_t* p = dst + row * 32 + col;
        int v = x[r];
        if (accumulate) v += gld(p);
        gst(p, v);
    }
}

// Store one 32x32 tile held as four 16x16 sub-tiles to [32][32] row-major memory.
__device__ __forceinline__ void store_tile16(const v4i (&x)[4], int32_t* __restrict__ dst, int lane,
                                             int accumulate) {
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = 16 * (s >> 1) + 4 * (lane >> 4) + q, col = 16 * (s & 1) + (lane & 15);
            int32_t* p = dst + row * 32 + col;
            int v = x[s][q];
            if (accumulate) v += gld(p);
            gst(p, v);
        }
}

}  // namespace mixer
