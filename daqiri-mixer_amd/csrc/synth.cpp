// synth.cpp -- seeded synthetic F-engine packets in the config.h wire format. Replaces the
// DAQIRI capture (out of scope) and restates the transmitter's signal model, k_sky
// (src/tx_host.cu:41-60), plus Gaussian noise and extra point sources. Header and seq baking
// follow src/tx_host.cu:136-143 (64-byte header, seq big-endian at byte 48, element = seq % N).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "mixer.h"

namespace {

// splitmix64 finaliser: a counter-based generator keyed by (seed, element, channel, sample).
inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Two independent N(0,1) by Box-Muller from one 64-bit key.
inline void gauss2(uint64_t key, double* a, double* b) {
    const uint64_t r1 = mix64(key), r2 = mix64(key ^ 0xD1B54A32D192ED03ull);
    const double u1 = ((double)(r1 >> 11) + 1.0) * (1.0 / 9007199254740993.0);  // (0, 1]
    const double u2 = (double)(r2 >> 11) * (1.0 / 9007199254740992.0);          // [0, 1)
    const double rad = std::sqrt(-2.0 * std::log(u1));
    *a = rad * std::cos(2.0 * M_PI * u2);
    *b = rad * std::sin(2.0 * M_PI * u2);
}

inline int8_t q8(double v) {
    long q = std::lrint(v);
    return (int8_t)(q < -127 ? -127 : (q > 127 ? 127 : q));
}

void put16(uint8_t* p, uint32_t v) {
    p[0] = (uint8_t)(v >> 8);
    p[1] = (uint8_t)v;
}
void put32(uint8_t* p, uint32_t v) {
    p[0] = (uint8_t)(v >> 24);
    p[1] = (uint8_t)(v >> 16);
    p[2] = (uint8_t)(v >> 8);
    p[3] = (uint8_t)v;
}

std::mutex g_wire_mtx;
mixer_wire g_wire;
bool g_wire_set = false;

}  // namespace

extern "C" {

void mixer_default_wire(mixer_wire* w) {
    if (!w) return;
    memset(w, 0, sizeof(*w));
    const uint8_t dst[6] = {0xa0, 0x88, 0xc2, 0x0d, 0x5e, 0x28};  // MIXER_DEF_ETH_DST, src/config.h:56
    memcpy(w->eth_dst, dst, 6);
    w->ip_src = (10u << 24) | 1u;  // 10.0.0.1, src/config.h:57
    w->ip_dst = (10u << 24) | 2u;  // 10.0.0.2, src/config.h:58
    w->udp_port = 4096;            // MIXER_UDP_PORT, src/config.h:52
}

// The standard IPv4 header checksum: one's-complement sum of the big-endian 16-bit words of
// the 20-byte header (checksum field zero), folded and complemented.
int mixer_build_header(const mixer_array_cfg* c, const mixer_wire* w, uint32_t seq, uint8_t* h) {
    if (!c || !w || !h) return MIXER_ERR_ARG;
    if (c->hdr_bytes < 42 || c->seq_byte < 42 || c->seq_byte + 4 > c->hdr_bytes) return MIXER_ERR_ARG;
    memset(h, 0, (size_t)c->hdr_bytes);
    const size_t wire = mixer_wire_bytes(c);
    memcpy(h, w->eth_dst, 6);
    memcpy(h + 6, w->eth_src, 6);
    put16(h + 12, 0x0800);  // ETH_P_IP
    uint8_t* ip = h + 14;
    ip[0] = 0x45;           // version 4, IHL 5
    put16(ip + 2, (uint32_t)std::min<size_t>(wire - 14, 65535));
    ip[8] = 64;             // TTL
    ip[9] = 17;             // UDP
    put32(ip + 12, w->ip_src);
    put32(ip + 16, w->ip_dst);
    uint32_t sum = 0;
    for (int i = 0; i < 20; i += 2) sum += ((uint32_t)ip[i] << 8) | ip[i + 1];
    while (sum >> 16) sum = (sum & 0xffffu) + (sum >> 16);
    put16(ip + 10, ~sum & 0xffffu);
    uint8_t* udp = h + 34;
    put16(udp, w->udp_port);
    put16(udp + 2, w->udp_port);
    put16(udp + 4, (uint32_t)std::min<size_t>(wire - 34, 65535));
    put32(h + c->seq_byte, seq);
    return MIXER_OK;
}

int mixer_synth_set_wire(const mixer_wire* w) {
    if (!w) return MIXER_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_wire_mtx);
    g_wire = *w;
    g_wire_set = true;
    return MIXER_OK;
}

void mixer_default_sky(mixer_sky* s) {
    if (!s) return;
    memset(s, 0, sizeof(*s));
    s->planet_l = 0.525;   // rho = 0.525 at orbit phase 0 (src/tx_host.cu:86,106)
    s->planet_m = 0.0;
    s->sigma = 0.5;
    s->star_amp = 2.0;     // Astar (src/tx_host.cu:42)
    s->planet_scale = 0.71;  // src/tx_host.cu:48
    s->nch_gen = 256;      // MIXER_NCH (src/config.h:28)
    s->n_src = 0;
}

int mixer_synth_packets(const mixer_array_cfg* cfg, const mixer_sky* sky, uint64_t seed,
                        uint64_t seq0, int npkts, uint8_t* out, size_t pitch, int nthreads) {
    if (!cfg || !sky || !out || npkts < 0) return MIXER_ERR_ARG;
    int st = mixer_validate_cfg(cfg);
    if (st) return st;
    if (pitch < mixer_wire_bytes(cfg) || sky->n_src < 0 || sky->n_src > 4 || sky->nch_gen < 2)
        return MIXER_ERR_ARG;
    const mixer_array_cfg c = *cfg;
    const int N = c.n_elem, C = c.n_chan, T = c.t_pkt;
    // per (element) phase terms and per (channel) planet amplitude, in double
    std::vector<double> pre(N), pim(N), apl(C);
    std::vector<double> sre((size_t)N, 0.0), sim((size_t)N, 0.0);
    for (int a = 0; a < N; ++a) {
        const int px = a % c.grid_x, qy = a / c.grid_x;
        const double phi = M_PI * (px * sky->planet_l + qy * sky->planet_m);
        pre[a] = std::cos(phi);
        pim[a] = std::sin(phi);
        for (int k = 0; k < sky->n_src; ++k) {
            const double ph = M_PI * (px * sky->src_l[k] + qy * sky->src_m[k]);
            sre[a] += sky->src_amp[k] * std::cos(ph);
            sim[a] += sky->src_amp[k] * std::sin(ph);
        }
    }
    for (int ch = 0; ch < C; ++ch)
        apl[ch] = sky->planet_scale * sky->star_amp * std::sqrt(2.0 * ch / (double)(sky->nch_gen - 1));
    const double sig = sky->sigma;
    mixer_wire wire;
    {
        std::lock_guard<std::mutex> lk(g_wire_mtx);
        if (g_wire_set) wire = g_wire;
        else mixer_default_wire(&wire);
    }
    std::vector<uint8_t> tmpl((size_t)c.hdr_bytes);
    st = mixer_build_header(&c, &wire, 0, tmpl.data());
    if (st) return st;
    auto work = [&](int i0, int i1) {
        for (int i = i0; i < i1; ++i) {
            const uint64_t seq = seq0 + (uint64_t)i;
            const int a = (int)(seq % (uint64_t)N);
            const uint64_t tb = seq / (uint64_t)N;
            uint8_t* p = out + (size_t)i * pitch;
            memcpy(p, tmpl.data(), (size_t)c.hdr_bytes);
            put32(p + c.seq_byte, (uint32_t)seq);
            int8_t* re = (int8_t*)(p + c.hdr_bytes);
            int8_t* im = re + (size_t)C * T;
            for (int ch = 0; ch < C; ++ch) {
                const double mr = sky->star_amp + apl[ch] * pre[a] + sre[a];
                const double mi = apl[ch] * pim[a] + sim[a];
                const uint64_t kac = mix64(mix64(seed ^ (uint64_t)a) ^ ((uint64_t)ch << 32));
                for (int t = 0; t < T; ++t) {
                    double nr = 0, ni = 0;
                    if (sig > 0) gauss2(mix64(kac ^ (tb * (uint64_t)T + (uint64_t)t)), &nr, &ni);
                    re[ch * T + t] = q8(25.0 * (mr + sig * nr));
                    im[ch * T + t] = q8(25.0 * (mi + sig * ni));
                }
            }
        }
    };
    int nt = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt > 64) nt = 64;
    if (npkts < 64 * nt) nt = 1 + npkts / 64;
    std::vector<std::thread> th;
    for (int k = 0; k < nt; ++k) {
        const int i0 = (int)((long)npkts * k / nt), i1 = (int)((long)npkts * (k + 1) / nt);
        th.emplace_back(work, i0, i1);
    }
    for (auto& t : th) t.join();
    return MIXER_OK;
}

// Orbiting planet of the reference transmitter's sky-update thread (src/tx_host.cu:98-100,
// 116-120): (l, m) = rho (cos w t, sin w t), w = 2 pi / torbit, t seconds since motion started.
int mixer_orbit_sky(const mixer_sky* in, double rho, double torbit_s, double t_s, mixer_sky* out) {
    if (!in || !out || !(torbit_s > 0)) return MIXER_ERR_ARG;
    const double w = 2.0 * M_PI / torbit_s;
    *out = *in;
    out->planet_l = rho * std::cos(w * t_s);
    out->planet_m = rho * std::sin(w * t_s);
    return MIXER_OK;
}

}  // extern "C"
