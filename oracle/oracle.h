/* oracle.h -- CPU restatement of the reference X-engine path (corner-turn, complex int8 GEMM,
 * int32->f32 flush, diff, gridding, forward 2-D DFT, image publish).
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (libmixer, the CLI, the serving loop) links,
 * loads or calls this code. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use it, as the checker or as the timed CPU baseline.
 *
 * Every function cites the reference file:line it restates (paths relative to the reference repo).
 * All entry points are extern "C" with plain pointers so ctypes can call them.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- a1: wire contract (src/config.h:10-38) ---------------------------------------------- */
/* Zero the 64-byte header and store seq big-endian at byte 48 (src/tx_host.cu:140-141). */
void orc_write_header(uint8_t* pkt, int hdr_bytes, int seq_byte, uint32_t seq);
/* Read seq big-endian at byte seq_byte (src/rx_host_corr.cu:134). */
uint32_t orc_read_seq(const uint8_t* pkt, int seq_byte);

/* ---- a3: corner-turn k_packAB (src/rx_host_corr.cu:56-72) --------------------------------- */
/* bb: header-stripped heaps [n_tb][N][2*C*T] (planar re[C*T] then im[C*T], plane index c*T+t).
 * Writes A=X and B=conj(X), all four [C][N][K] with K = n_tb*T and k = tb*T + t;
 * bim = (int8)(-im), which wraps -128 to -128 exactly as the reference's narrowing does.
 * Any of the four output pointers may be NULL (skipped). */
void orc_unpack(const uint8_t* bb, int n_tb, int N, int C, int T,
                int8_t* are, int8_t* aim, int8_t* bre, int8_t* bim);

/* ---- a4: cgemm_blk_t<ACC> (src/cgemm.cuh:28-69) -------------------------------------------- */
/* Per batch b: Cre += Are*Bre + wrap8(-Aim)*Bim ; Cim += Are*Bim + Aim*Bre, with A [M][K]
 * row-major and B [N][K] (the reference reads its [N][K] B column-major, cgemm.cuh:55).
 * acc=0 overwrites C, acc=1 adds into it (cgemm.cuh:42-45). Sums are formed in int64; the
 * return value is the number of outputs that do not fit int32 (0 = exact). Full square. */
long orc_cgemm(int M, int N, int K, int batch,
               const int8_t* are, const int8_t* aim, const int8_t* bre, const int8_t* bim,
               int32_t* cre, int32_t* cim, int acc);

/* The CPU correlator used as the timed CPU baseline: heaps -> V (full square, reference
 * semantics, int32) for channels [c0, c1) only, OpenMP over (channel, row). Output
 * vre/vim are [c1-c0][N][N]. nthreads<=0 uses the OpenMP default. Returns overflow count. */
long orc_xengine(const uint8_t* bb, int n_tb, int N, int C, int T, int c0, int c1,
                 int32_t* vre, int32_t* vim, int nthreads);

/* ---- a5: k_flush_cube (src/rx_host_corr.cu:86-95) ----------------------------------------- */
/* vcube [C][2][N][N] f32 += (float)vi*; vi* zeroed. vire/viim are [C][N][N]. */
void orc_flush(float* vcube, int32_t* vire, int32_t* viim, int C, int N);

/* ---- a6: k_diff (src/rx_host_corr.cu:73-77) ------------------------------------------------ */
void orc_diff(const float* vs, float* vp, float* dv, size_t n);

/* ---- a7: k_grid_cube (src/rx_host_corr.cu:78-85), generalised to a grid_x-wide element grid
 * and an ng x ng image grid: element a sits at (a % grid_x, a / grid_x); cell
 * (dy,dx) = ((qa-qb) mod ng, (pa-pb) mod ng). The reference hard-codes grid_x=16, ng=32.
 * dv is the planar f32 cube [C][2][N][N]; g is [C][ng][ng][2] (re, im) in FP64. */
void orc_grid(const float* dv, int C, int N, int grid_x, int ng, double* g);

/* ---- a8: cufftExecC2C(..., CUFFT_FORWARD) on ng x ng, batched over C (src/rx_host_corr.cu:205-206,222)
 * restated as the FP64 direct 2-D DFT: F[ky][kx] = sum G[y][x] exp(-2 pi i (ky y + kx x)/ng),
 * unnormalised. g and f are [C][ng][ng][2]. */
void orc_dft2(const double* g, int C, int ng, double* f);

/* ---- a9: image publish (src/rx_host_corr.cu:229-236) -------------------------------------- */
/* img[c][y][x] = Re F[c][(y+ng/2)%ng][(x+ng/2)%ng]; vmax/vmin per channel. */
void orc_publish(const double* f, int C, int ng, float* img, float* vmax, float* vmin);

/* ---- signal model of k_sky (src/tx_host.cu:41-60), in double with round-half-even --------- */
/* payload for one antenna a at direction (l, m): re/im planes [C*T], values clipped to +-127;
 * the same value for all T samples. nch_gen is the generator's channel sweep (256). */
void orc_sky_payload(int a, int grid_x, int C, int T, int nch_gen, double l, double m, uint8_t* payload);

#ifdef __cplusplus
}
#endif
