// cli.cpp -- mixer_corr: the reference's rx_host_corr command line (src/rx_host_corr.cu:177-246)
// on libmixer, fed by the seeded synthetic source instead of the DAQIRI capture.
//
//   mixer_corr [cfg|-] [--seconds N] [--fps F] [--device D] [--noflop] [--flush N] [--dump PATH]
//              [--dump-mode cumulative|integ] [--elements N] [--grid-x G] [--channels C]
//              [--k-update K] [--ng NG] [--rate GBps] [--sigma S] [--seed X] [--shm NAME]
//              [--rho R] [--torbit S] [--refresh-hz F] [--eth-dst MAC] [--frame dv|latest]
//              [--slots N] [--no-wrap-check]
// --rho/--torbit/--refresh-hz move the synthetic planet on the orbit (l, m) = rho (cos, sin)
// (2 pi t / torbit), re-generating one replay slot per refresh (src/tx_host.cu:84-86,116-125);
// --eth-dst sets the wire header's destination MAC (src/tx_host.cu:87,96-98).
// --frame dv (default) publishes the image of everything integrated since the previous frame,
// as the reference's monitor does (dV, src/rx_host_corr.cu:218-222); --frame latest publishes the
// latest integration's image. --slots N bounds the ingest slots (default 96, the reference's
// NSLOT, src/rx_host_corr.cu:46): a time block that finds none free is dropped (imissed).
// --no-wrap-check skips the imaginary -128 test (the synthetic source clips to [-127, 127], as
// the reference's transmitter does, src/tx_host.cu:54-55).
//
// The first positional argument (the reference's DAQIRI YAML) is accepted and ignored.
// --flush N sets the samples per integration to N V-updates (the reference flushes its int32
// cube to f32 every N V-updates). --dump appends f32 [C][2][N][N] records about once per
// second, the cumulative-since-start cube by default as the reference does, and writes a JSON
// sidecar PATH.json describing the records (the reference's records are headerless). The image cube is
// published to /dev/shm/corr32 (Corr32 layout) at --fps. Any library error exits with 1.
#include <fcntl.h>
#include <sys/mman.h>
#include <unistd.h>

#include <atomic>
#include <chrono>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "mixer.h"

static volatile std::sig_atomic_t g_stop = 0;
static void on_sig(int) { g_stop = 1; }

#define CK(x)                                                                              \
    do {                                                                                   \
        int s_ = (x);                                                                      \
        if (s_) {                                                                          \
            fprintf(stderr, "[mixer_corr] %s:%d: %s (%d)\n", __FILE__, __LINE__, mixer_strerror(s_), s_); \
            std::exit(1);                                                                  \
        }                                                                                  \
    } while (0)

static uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char** argv) {
    mixer_array_cfg cfg;
    mixer_default_cfg(&cfg);
    cfg.k_update = 1024;  // KPCH (src/rx_host_corr.cu:40)
    int seconds = 0, device = 0, flush = 8;
    float fps = 20.f;
    bool noflop = false, dump_cum = true, frame_dv = true, no_wrap = false;
    int slots = 96;
    const char* dump_path = nullptr;
    const char* shm_name = "/corr32";
    double rate = 0, sigma = 0.5, rho = 0.525, torbit = 3.0, refresh_hz = 0.0;
    uint64_t seed = 0x4D495832ull;
    mixer_wire hdr_wire;
    mixer_default_wire(&hdr_wire);
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto nxt = [&]() -> const char* {
            if (i + 1 >= argc) { fprintf(stderr, "missing value for %s\n", a.c_str()); std::exit(1); }
            return argv[++i];
        };
        if (a == "--seconds") seconds = atoi(nxt());
        else if (a == "--fps") fps = (float)atof(nxt());
        else if (a == "--device") device = atoi(nxt());
        else if (a == "--noflop") noflop = true;
        else if (a == "--flush") flush = atoi(nxt());
        else if (a == "--dump") dump_path = nxt();
        else if (a == "--dump-mode") dump_cum = std::string(nxt()) != "integ";
        else if (a == "--elements") cfg.n_elem = atoi(nxt());
        else if (a == "--grid-x") cfg.grid_x = atoi(nxt());
        else if (a == "--channels") cfg.n_chan = atoi(nxt());
        else if (a == "--k-update") cfg.k_update = atoi(nxt());
        else if (a == "--ng") cfg.ng = atoi(nxt());
        else if (a == "--rate") rate = atof(nxt());
        else if (a == "--sigma") sigma = atof(nxt());
        else if (a == "--seed") seed = strtoull(nxt(), nullptr, 0);
        else if (a == "--shm") shm_name = nxt();
        else if (a == "--rho") rho = atof(nxt());
        else if (a == "--torbit") torbit = atof(nxt());
        else if (a == "--refresh-hz") refresh_hz = atof(nxt());
        else if (a == "--frame") frame_dv = std::string(nxt()) != "latest";
        else if (a == "--slots") slots = atoi(nxt());
        else if (a == "--no-wrap-check") no_wrap = true;
        else if (a == "--eth-dst") {
            const char* s = nxt();
            unsigned m[6];
            if (sscanf(s, "%x:%x:%x:%x:%x:%x", &m[0], &m[1], &m[2], &m[3], &m[4], &m[5]) != 6) {
                fprintf(stderr, "bad MAC %s\n", s);
                return 1;
            }
            for (int k = 0; k < 6; ++k) hdr_wire.eth_dst[k] = (uint8_t)m[k];
        }
        else if (a == "-h" || a == "--help") {
            fprintf(stderr, "usage: %s [cfg] [--seconds N][--fps F][--device D][--noflop][--flush N][--dump PATH]\n"
                            "  [--dump-mode cumulative|integ][--elements N][--grid-x G][--channels C][--k-update K]\n"
                            "  [--ng NG][--rate GBps][--sigma S][--seed X][--shm NAME]\n"
                            "  [--rho R][--torbit S][--refresh-hz F][--eth-dst MAC][--frame dv|latest][--slots N]\n"
                            "  [--no-wrap-check]\n",
                    argv[0]);
            return 1;
        } else if (a[0] == '-' && a != "-") { fprintf(stderr, "unknown flag %s\n", a.c_str()); return 1; }
    }
    if (fps < 1.f) { fprintf(stderr, "--fps must be >= 1\n"); return 1; }
    if (flush < 1) { fprintf(stderr, "--flush must be >= 1\n"); return 1; }
    cfg.k_integ = flush * cfg.k_update;
    if (dump_path && dump_cum) cfg.flags |= MIXER_CFG_CUMULATIVE;
    if (noflop) cfg.flags |= MIXER_CFG_NOFLOP;
    if (frame_dv) cfg.flags |= MIXER_CFG_FRAME_DV;
    if (no_wrap) cfg.flags |= MIXER_CFG_NO_WRAP_CHECK;
    if (slots < 0) { fprintf(stderr, "--slots must be >= 0\n"); return 1; }
    CK(mixer_validate_cfg(&cfg));
    std::signal(SIGINT, on_sig);
    std::signal(SIGTERM, on_sig);
    CK(mixer_set_device(device));
    mixer_ctx* ctx = nullptr;
    CK(mixer_create(&cfg, device, &ctx));
    CK(mixer_set_ingest_slots(ctx, slots));

    // Corr32 in shared memory (src/rx_host_corr.cu:100-105)
    const size_t cbytes = mixer_corr32_bytes(&cfg);
    shm_unlink(shm_name);
    int fd = shm_open(shm_name, O_CREAT | O_RDWR, 0666);
    if (fd < 0) { perror("shm_open"); return 1; }
    if (ftruncate(fd, (off_t)cbytes)) { perror("ftruncate"); return 1; }
    void* shm = mmap(nullptr, cbytes, PROT_READ | PROT_WRITE, MAP_SHARED, fd, 0);
    close(fd);
    if (shm == MAP_FAILED) { perror("mmap"); return 1; }
    CK(mixer_corr32_init(shm, &cfg));

    // synthetic source: a pinned ring of RB time blocks, generated once and replayed with fresh seq
    const int N = cfg.n_elem;
    const size_t pitch = (size_t)cfg.pkt_pitch, wire = mixer_wire_bytes(&cfg);
    const int RB = 64, NSLOT = 8, per_slot = RB / NSLOT;
    uint8_t* ring = nullptr;
    CK(mixer_host_alloc((size_t)RB * N * pitch, (void**)&ring));
    mixer_sky sky;
    mixer_default_sky(&sky);
    sky.sigma = sigma;
    sky.planet_l = rho;  // orbit phase 0 (src/tx_host.cu:106)
    sky.planet_m = 0.0;
    CK(mixer_synth_set_wire(&hdr_wire));
    CK(mixer_synth_packets(&cfg, &sky, seed, 0, RB * N, ring, pitch, 0));
    std::vector<uint64_t> slot_tok(NSLOT, 0);
    std::atomic<uint64_t> blocks{0};
    const size_t payload = mixer_payload_bytes(&cfg);
    fprintf(stderr, "[mixer_corr] %d elements x %d channels, %d samples/integration (%d per V-update), ng %d, "
                    "device %d @ %.0f Hz, shm %s%s\n", N, cfg.n_chan, cfg.k_integ, cfg.k_update, cfg.ng, device,
            fps, shm_name, noflop ? " [noflop]" : "");

    std::thread src([&] {
        uint64_t seq = 0;
        const uint64_t t0 = now_ns();
        uint64_t next_refresh = t0;
        int slot = 0;
        while (!g_stop) {
            if (slot_tok[slot]) CK(mixer_token_wait(ctx, slot_tok[slot]));  // slot free again
            uint8_t* base = ring + (size_t)slot * per_slot * N * pitch;
            if (refresh_hz > 0 && torbit > 0 && now_ns() >= next_refresh) {
                // the planet moved: re-generate this slot at its new position (fresh noise too)
                mixer_sky s2;
                CK(mixer_orbit_sky(&sky, rho, torbit, (now_ns() - t0) * 1e-9, &s2));
                CK(mixer_synth_packets(&cfg, &s2, seed, seq, per_slot * N, base, pitch, 0));
                next_refresh += (uint64_t)(1e9 / refresh_hz);
            }
            for (int k = 0; k < per_slot * N; ++k) {  // bake the current seq (src/tx_host.cu:141)
                const uint32_t s = (uint32_t)(seq + k);
                uint8_t* p = base + (size_t)k * pitch + cfg.seq_byte;
                p[0] = (uint8_t)(s >> 24); p[1] = (uint8_t)(s >> 16); p[2] = (uint8_t)(s >> 8); p[3] = (uint8_t)s;
            }
            if (rate > 0) {  // pace to rate GB/s of payload
                const double due = (double)(blocks.load() * N) * payload / (rate * 1e9);
                while (!g_stop && (now_ns() - t0) * 1e-9 < due) std::this_thread::sleep_for(std::chrono::microseconds(50));
            }
            CK(mixer_submit_heaps(ctx, base, pitch, per_slot * N, now_ns(), &slot_tok[slot]));
            seq += (uint64_t)per_slot * N;
            blocks += per_slot;
            slot = (slot + 1) % NSLOT;
        }
    });

    FILE* dumpf = nullptr;
    std::vector<float> hv;
    if (dump_path) {
        dumpf = fopen(dump_path, "wb");
        if (!dumpf) { perror("dump fopen"); g_stop = 1; src.join(); return 1; }
        hv.resize((size_t)cfg.n_chan * 2 * N * N);
        const std::string side = std::string(dump_path) + ".json";
        FILE* jf = fopen(side.c_str(), "w");
        if (!jf) { perror("dump sidecar fopen"); g_stop = 1; src.join(); return 1; }
        fprintf(jf,
                "{\"format\": \"mixer-dump\", \"dtype\": \"float32-le\", "
                "\"layout\": \"[n_chan][2][n_elem][n_elem] (plane 0 = Re, 1 = Im; row a, column b)\", "
                "\"n_chan\": %d, \"n_elem\": %d, \"grid_x\": %d, \"record_bytes\": %zu, "
                "\"mode\": \"%s\", \"samples_per_integration\": %d, \"records_per_second\": 1}\n",
                cfg.n_chan, N, cfg.grid_x, hv.size() * sizeof(float),
                dump_cum ? "cumulative" : "integration", cfg.k_integ);
        fclose(jf);
    }
    const auto t0 = std::chrono::steady_clock::now();
    auto next = t0;
    uint32_t frame = 0;
    uint64_t last_blocks = 0;
    while (!g_stop) {
        if (seconds > 0 && std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >= seconds) break;
        next += std::chrono::duration_cast<std::chrono::steady_clock::duration>(std::chrono::duration<double>(1.0 / fps));
        std::this_thread::sleep_until(next);
        if (!noflop) CK(mixer_publish_image(ctx, shm, cbytes));
        const bool dump_now = dumpf && frame % (uint32_t)fps == 0;
        if (dump_now && !noflop) {
            uint64_t idx = 0;
            CK(mixer_dump(ctx, hv.data(), dump_cum ? 1 : 0, &idx));
            if (idx) {
                fwrite(hv.data(), sizeof(float), hv.size(), dumpf);
                fflush(dumpf);
            }
        }
        if (frame % (uint32_t)fps == 0) {
            mixer_stats st;
            CK(mixer_get_stats(ctx, &st));
            const uint64_t b = blocks.load();
            fprintf(stderr, "[mixer_corr] frame %u blocks/s=%.0f (%.2f GB/s) integrations=%lu latency=%.2f ms "
                            "imissed=%lu misalign=%lu gaps=%lu wrap128=%lu\n",
                    frame, (double)(b - last_blocks) * fps / (frame ? fps : 1),
                    (double)(b - last_blocks) * N * payload / 1e9, (unsigned long)st.integrations,
                    st.last_dump_latency_ms, (unsigned long)st.imissed, (unsigned long)st.misalign,
                    (unsigned long)st.seq_gaps, (unsigned long)st.wrap128);
            last_blocks = b;
        }
        frame++;
    }
    g_stop = 1;
    src.join();
    CK(mixer_sync(ctx));
    mixer_stats st;
    CK(mixer_get_stats(ctx, &st));
    fprintf(stderr, "[mixer_corr] stop after %u frames; V-updates=%lu integrations=%lu heaps=%lu wire_bytes=%zu\n",
            frame, (unsigned long)st.vupdates, (unsigned long)st.integrations, (unsigned long)st.heaps_in, wire);
    if (dumpf) fclose(dumpf);
    CK(mixer_destroy(ctx));
    CK(mixer_host_free(ring));
    munmap(shm, cbytes);
    shm_unlink(shm_name);
    return 0;
}
