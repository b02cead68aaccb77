/* oracle.cpp -- CPU restatement of the reference X-engine path. TEST INFRASTRUCTURE ONLY:
 * the product never links or calls this file; tests/, smoke() and bench.py's cpu_baseline
 * leg use it as the checker and as the timed CPU baseline.
 *
 * Pinning (what makes each function trustworthy without running the reference, which needs
 * CUDA/cuFFT/DAQIRI and cannot be built here):
 *   orc_write_header/orc_read_seq  byte-level round trips + the config.h offsets (tests/test_oracle.py)
 *   orc_unpack                      bijection/permutation check + numpy reshape formulation
 *   orc_cgemm/orc_xengine           numpy complex128 X @ X^H (exact for |x|<=127), rank-1 closed
 *                                   forms of the k_sky signal, Hermitian/diagonal invariants, a
 *                                   hand-computed -128 wrap case, golden fixtures in tests/golden/
 *   orc_flush/orc_diff              FP64 sums; single int32->f32 rounding bound
 *   orc_grid                        all-ones closed form (gx-|dx|)(gy-|dy|), Hermitian G(-d)=conj G(d)
 *   orc_dft2                        numpy.fft.fft2, delta -> flat, plane wave -> peak
 *   orc_publish                     the reference peek32 viewer (built from source into oracle/_ref)
 *                                   reads a Corr32 written from this output; fftshift closed form
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

/* src/config.h:10-14 (header map), src/tx_host.cu:140-141 (seq baked big-endian at byte 48). */
void orc_write_header(uint8_t* pkt, int hdr_bytes, int seq_byte, uint32_t seq) {
    memset(pkt, 0, (size_t)hdr_bytes);
    pkt[seq_byte + 0] = (uint8_t)(seq >> 24);
    pkt[seq_byte + 1] = (uint8_t)(seq >> 16);
    pkt[seq_byte + 2] = (uint8_t)(seq >> 8);
    pkt[seq_byte + 3] = (uint8_t)(seq);
}

/* src/rx_host_corr.cu:134: seq0 = bswap32(*(u32*)(p0+48)). */
uint32_t orc_read_seq(const uint8_t* pkt, int seq_byte) {
    return ((uint32_t)pkt[seq_byte] << 24) | ((uint32_t)pkt[seq_byte + 1] << 16) |
           ((uint32_t)pkt[seq_byte + 2] << 8) | (uint32_t)pkt[seq_byte + 3];
}

/* src/rx_host_corr.cu:61-72. The reference walks i = ant*TOTAL_SAMP + s, heap = s/SAMP_PER_HEAP,
 * sl = s%SAMP_PER_HEAP, c = sl/T, t = sl%T, klocal = heap*T + t, reading bb[heap][ant][sl] and
 * bb[heap][ant][SAMP_PER_HEAP+sl], writing index c*N*K + ant*K + klocal of all four arrays. */
void orc_unpack(const uint8_t* bb, int n_tb, int N, int C, int T,
                int8_t* are, int8_t* aim, int8_t* bre, int8_t* bim) {
    const size_t samp = (size_t)C * T;         /* SAMP_PER_HEAP */
    const size_t nb = 2 * samp;                /* payload bytes per heap */
    const size_t K = (size_t)n_tb * T;
    for (int ant = 0; ant < N; ++ant)
        for (int heap = 0; heap < n_tb; ++heap) {
            const uint8_t* h = bb + ((size_t)heap * N + ant) * nb;
            for (size_t sl = 0; sl < samp; ++sl) {
                size_t c = sl / T, t = sl % T, k = (size_t)heap * T + t;
                int8_t re = (int8_t)h[sl];
                int8_t im = (int8_t)h[samp + sl];
                size_t a = c * (size_t)N * K + (size_t)ant * K + k;
                if (are) are[a] = re;
                if (aim) aim[a] = im;
                if (bre) bre[a] = re;
                if (bim) bim[a] = (int8_t)(-(int)im);   /* (MixerI8)(-im): wraps -128 -> -128 */
            }
        }
}

static inline int32_t sat_count(int64_t v, long* of) {
    if (v > INT32_MAX || v < INT32_MIN) ++*of;
    return (int32_t)(uint32_t)(uint64_t)v;
}

/* src/cgemm.cuh:28-69: per k, aneg = -aim narrowed to int8 (cgemm.cuh:52), then
 * cre += are*bre + aneg*bim, cim += are*bim + aim*bre (cgemm.cuh:59-62). B is [N][K] read
 * column-major, so B[k][n] = b[n*K + k] (cgemm.cuh:55). */
long orc_cgemm(int M, int N, int K, int batch,
               const int8_t* are, const int8_t* aim, const int8_t* bre, const int8_t* bim,
               int32_t* cre, int32_t* cim, int acc) {
    long of = 0;
    for (int b = 0; b < batch; ++b) {
        const int8_t* ar = are + (size_t)b * M * K;
        const int8_t* ai = aim + (size_t)b * M * K;
        const int8_t* br = bre + (size_t)b * N * K;
        const int8_t* bi = bim + (size_t)b * N * K;
        int32_t* cr = cre + (size_t)b * M * N;
        int32_t* ci = cim + (size_t)b * M * N;
        for (int m = 0; m < M; ++m)
            for (int n = 0; n < N; ++n) {
                int64_t sr = acc ? cr[(size_t)m * N + n] : 0;
                int64_t si = acc ? ci[(size_t)m * N + n] : 0;
                for (int k = 0; k < K; ++k) {
                    int8_t a_r = ar[(size_t)m * K + k], a_i = ai[(size_t)m * K + k];
                    int8_t aneg = (int8_t)(-(int)a_i);
                    int8_t b_r = br[(size_t)n * K + k], b_i = bi[(size_t)n * K + k];
                    sr += (int64_t)a_r * b_r + (int64_t)aneg * b_i;
                    si += (int64_t)a_r * b_i + (int64_t)a_i * b_r;
                }
                cr[(size_t)m * N + n] = sat_count(sr, &of);
                ci[(size_t)m * N + n] = sat_count(si, &of);
            }
    }
    return of;
}

/* The CPU correlator (baseline): the corner-turn of rx_host_corr.cu:61-72 for one channel
 * followed by the cgemm.cuh math over the full square. re / neg=wrap8(-im) / im rows are
 * formed once per channel; dot products run over contiguous k in int32 chunks (each chunk of
 * 4096 samples is < 2^31 in magnitude) summed in int64. */
long orc_xengine(const uint8_t* bb, int n_tb, int N, int C, int T, int c0, int c1,
                 int32_t* vre, int32_t* vim, int nthreads) {
    const size_t samp = (size_t)C * T, nb = 2 * samp, K = (size_t)n_tb * T;
    const int nc = c1 - c0;
    std::vector<int8_t> re((size_t)nc * N * K), im((size_t)nc * N * K), ng((size_t)nc * N * K);
    long of = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    /* corner-turn into [c][ant][k] */
#pragma omp parallel for collapse(2) schedule(static)
    for (int ant = 0; ant < N; ++ant)
        for (int heap = 0; heap < n_tb; ++heap) {
            const uint8_t* h = bb + ((size_t)heap * N + ant) * nb;
            for (int c = c0; c < c1; ++c)
                for (int t = 0; t < T; ++t) {
                    size_t sl = (size_t)c * T + t;
                    size_t a = ((size_t)(c - c0) * N + ant) * K + (size_t)heap * T + t;
                    int8_t r = (int8_t)h[sl], i = (int8_t)h[samp + sl];
                    re[a] = r; im[a] = i; ng[a] = (int8_t)(-(int)i);
                }
        }
    const size_t CH = 4096;
#pragma omp parallel for collapse(2) schedule(dynamic, 4) reduction(+ : of)
    for (int c = 0; c < nc; ++c)
        for (int m = 0; m < N; ++m) {
            const int8_t* ar = &re[((size_t)c * N + m) * K];
            const int8_t* ai = &im[((size_t)c * N + m) * K];
            const int8_t* an = &ng[((size_t)c * N + m) * K];
            for (int n = 0; n < N; ++n) {
                const int8_t* br = &re[((size_t)c * N + n) * K];
                const int8_t* bn = &ng[((size_t)c * N + n) * K];
                int64_t sr = 0, si = 0;
                for (size_t k0 = 0; k0 < K; k0 += CH) {
                    size_t k1 = k0 + CH < K ? k0 + CH : K;
                    int32_t pr = 0, pi = 0;
                    for (size_t k = k0; k < k1; ++k) {
                        pr += (int32_t)ar[k] * br[k] + (int32_t)an[k] * bn[k];
                        pi += (int32_t)ar[k] * bn[k] + (int32_t)ai[k] * br[k];
                    }
                    sr += pr; si += pi;
                }
                long o = 0;
                vre[((size_t)c * N + m) * N + n] = sat_count(sr, &o);
                vim[((size_t)c * N + m) * N + n] = sat_count(si, &o);
                of += o;
            }
        }
    return of;
}

/* src/rx_host_corr.cu:90-95. */
void orc_flush(float* vcube, int32_t* vire, int32_t* viim, int C, int N) {
    const size_t mn = (size_t)N * N;
    for (int ch = 0; ch < C; ++ch)
        for (size_t j = 0; j < mn; ++j) {
            size_t i = (size_t)ch * mn + j;
            vcube[(size_t)ch * 2 * mn + j] += (float)vire[i]; vire[i] = 0;
            vcube[(size_t)ch * 2 * mn + mn + j] += (float)viim[i]; viim[i] = 0;
        }
}

/* src/rx_host_corr.cu:74-77. */
void orc_diff(const float* vs, float* vp, float* dv, size_t n) {
    for (size_t i = 0; i < n; ++i) { dv[i] = vs[i] - vp[i]; vp[i] = vs[i]; }
}

/* src/rx_host_corr.cu:79-85 with a>>4 -> a / grid_x, a&15 -> a % grid_x, NG -> ng. */
void orc_grid(const float* dv, int C, int N, int grid_x, int ng, double* g) {
    const size_t mn = (size_t)N * N;
    memset(g, 0, sizeof(double) * 2 * (size_t)C * ng * ng);
    for (int ch = 0; ch < C; ++ch) {
        const float* d = dv + (size_t)ch * 2 * mn;
        double* gc = g + (size_t)ch * 2 * ng * ng;
        for (int a = 0; a < N; ++a)
            for (int b = 0; b < N; ++b) {
                int dy = (((a / grid_x) - (b / grid_x)) % ng + ng) % ng;
                int dx = (((a % grid_x) - (b % grid_x)) % ng + ng) % ng;
                size_t i = (size_t)a * N + b;
                gc[2 * (dy * ng + dx)] += (double)d[i];
                gc[2 * (dy * ng + dx) + 1] += (double)d[mn + i];
            }
    }
}

/* FP64 direct DFT, separable: rows then columns, each a direct O(ng^2) 1-D DFT with
 * twiddle exp(-2 pi i (j*k mod ng)/ng) computed in double. */
void orc_dft2(const double* g, int C, int ng, double* f) {
    std::vector<double> cw((size_t)ng), sw((size_t)ng), tmp(2 * (size_t)ng * ng);
    for (int j = 0; j < ng; ++j) {
        cw[j] = cos(2.0 * M_PI * j / ng);
        sw[j] = -sin(2.0 * M_PI * j / ng);
    }
    for (int ch = 0; ch < C; ++ch) {
        const double* gc = g + (size_t)ch * 2 * ng * ng;
        double* fc = f + (size_t)ch * 2 * ng * ng;
        for (int y = 0; y < ng; ++y)                       /* along x */
            for (int kx = 0; kx < ng; ++kx) {
                double sr = 0, si = 0;
                for (int x = 0; x < ng; ++x) {
                    int w = (int)(((long)kx * x) % ng);
                    double ar = gc[2 * (y * ng + x)], ai = gc[2 * (y * ng + x) + 1];
                    sr += ar * cw[w] - ai * sw[w];
                    si += ar * sw[w] + ai * cw[w];
                }
                tmp[2 * (y * ng + kx)] = sr; tmp[2 * (y * ng + kx) + 1] = si;
            }
        for (int kx = 0; kx < ng; ++kx)                    /* along y */
            for (int ky = 0; ky < ng; ++ky) {
                double sr = 0, si = 0;
                for (int y = 0; y < ng; ++y) {
                    int w = (int)(((long)ky * y) % ng);
                    double ar = tmp[2 * (y * ng + kx)], ai = tmp[2 * (y * ng + kx) + 1];
                    sr += ar * cw[w] - ai * sw[w];
                    si += ar * sw[w] + ai * cw[w];
                }
                fc[2 * (ky * ng + kx)] = sr; fc[2 * (ky * ng + kx) + 1] = si;
            }
    }
}

/* src/rx_host_corr.cu:229-235: v = Re h[((y+NG/2)%NG)*NG + (x+NG/2)%NG]; running max/min
 * start at -1e30/1e30 exactly as the reference does. */
void orc_publish(const double* f, int C, int ng, float* img, float* vmax, float* vmin) {
    for (int ch = 0; ch < C; ++ch) {
        const double* fc = f + (size_t)ch * 2 * ng * ng;
        float mx = -1e30f, mn = 1e30f;
        for (int y = 0; y < ng; ++y)
            for (int x = 0; x < ng; ++x) {
                float v = (float)fc[2 * (((y + ng / 2) % ng) * ng + (x + ng / 2) % ng)];
                img[(size_t)ch * ng * ng + y * ng + x] = v;
                mx = v > mx ? v : mx; mn = v < mn ? v : mn;
            }
        vmax[ch] = mx; vmin[ch] = mn;
    }
}

/* src/tx_host.cu:41-60: Astar=2, px=a&15, qy=a>>4 (here a%grid_x, a/grid_x),
 * Apl = 0.71*Astar*sqrt(2c/(NCH-1)), phi = pi*(px*l + qy*m),
 * qr = rint((Astar + Apl cos phi)*25), qi = rint((Apl sin phi)*25), clipped to +-127, written
 * for all T samples of channel c in each plane. Double precision, round-half-even. */
void orc_sky_payload(int a, int grid_x, int C, int T, int nch_gen, double l, double m, uint8_t* payload) {
    const double astar = 2.0;
    int px = a % grid_x, qy = a / grid_x;
    for (int c = 0; c < C; ++c) {
        double apl = 0.71 * astar * sqrt(2.0 * c / (double)(nch_gen - 1));
        double phi = M_PI * (px * l + qy * m);
        long qr = lrint((astar + apl * cos(phi)) * 25.0);
        long qi = lrint((apl * sin(phi)) * 25.0);
        qr = qr < -127 ? -127 : (qr > 127 ? 127 : qr);
        qi = qi < -127 ? -127 : (qi > 127 ? 127 : qi);
        for (int t = 0; t < T; ++t) {
            payload[c * T + t] = (uint8_t)(int8_t)qr;
            payload[(size_t)C * T + c * T + t] = (uint8_t)(int8_t)qi;
        }
    }
}

}  /* extern "C" */
