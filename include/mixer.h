/* mixer.h -- C ABI of libmixer, the MI355X (gfx950) X-engine.
 *
 * The reference (DAQIRI-MIXER) is one executable with no library API; its operator surface is
 * four seams, each replaced here by extern "C" entry points with plain pointers and sizes:
 *   input seam   Slot{mr_ptr} of 256 packets at pitch MR_STRIDE   src/rx_host_corr.cu:120-121,136-141,159
 *   kernel seam  k_packAB(bb, Are, Aim, Bre, Bim)                  src/rx_host_corr.cu:61-72
 *                cgemm_blk_t<ACC>(M,N,K,Are,Aim,Bre,Bim,Cre,Cim)  src/cgemm.cuh:28-69
 *                k_flush_cube(Vcube, Vire, Viim)                   src/rx_host_corr.cu:90-95
 *                k_diff / k_grid_cube / cufftExecC2C               src/rx_host_corr.cu:74-85,205-206,222
 *   output seam  --dump record [C][2][N][N] f32                    src/rx_host_corr.cu:187,210-227
 *                Corr32 in /dev/shm/corr32                          src/rx_host_corr.cu:97-105,229-236
 *   CLI          rx_host_corr <cfg> [--seconds][--fps][--device][--noflop][--flush][--dump]
 *                                                                   src/rx_host_corr.cu:177-183
 *
 * Error convention: every int-returning call returns MIXER_OK (0) or a negative mixer status;
 * the reference instead exits the process on any CUDA error (CK -> exit(1),
 * src/rx_host_corr.cu:35). The CLI maps a nonzero status to exit code 1.
 *
 * Threading: one submitting thread per context (the reference's gpu_thread role); mixer_dump,
 * mixer_read_image and mixer_get_stats may be called from one other thread (its main/pub role).
 *
 * Value domain: inputs are complex int8 with components in [-127, 127], which the reference's
 * own transmitter guarantees (clip at src/tx_host.cu:54-55). Inside that domain V is exactly
 * Hermitian and every output is bit-identical to the reference's full-square product. An
 * imaginary sample of -128 is counted (mixer_stats.wrap128): the lower triangle and the
 * diagonal tiles stay bit-identical to the reference even then, but the mirrored upper
 * triangle is conj(lower), which the reference's wrapped negation does not reproduce. The image
 * grids each baseline's lower orientation and its conjugate, so it is the image of the Hermitian
 * V that conj(lower) defines (inside the value domain: of V itself).
 */
#ifndef MIXER_H
#define MIXER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------- */
enum {
    MIXER_OK = 0,
    MIXER_ERR_ARG = -1,      /* bad argument or unsupported geometry */
    MIXER_ERR_HIP = -2,      /* a HIP runtime call failed */
    MIXER_ERR_NOMEM = -3,    /* device or pinned host allocation failed */
    MIXER_ERR_STATE = -4,    /* call not valid in the context's current state */
    MIXER_ERR_BUDGET = -5,   /* HBM budget would be exceeded (serving admission) */
    MIXER_ERR_NODEV = -6,    /* no usable gfx950 device */
    MIXER_ERR_OVERFLOW = -7, /* samples per integration exceed the exact int32 range */
};

/* ---- array geometry and wire format -------------------------------------------------------
 * Defaults (mixer_default_cfg) are the reference's compile-time constants:
 * MIXER_NANT=256, MIXER_GRID=16, MIXER_ACTIVE_NCH=128, MIXER_TPKT=32, MIXER_HDR_BYTES=64,
 * MIXER_SEQ_BYTE=48 (src/config.h:26-38); KPCH=1024 per V-update, MR_STRIDE=16384, NG=32
 * (src/rx_host_corr.cu:40-51); 8192 samples per 20 Hz integration.
 * Packet: hdr_bytes header (seq big-endian u32 at seq_byte), then a payload of
 * 2*n_chan*t_pkt bytes, PLANAR: re[c*t_pkt + t] for all c, then im[...] (src/config.h:10-23).
 * Element of a packet = seq % n_elem. n_elem consecutive packets (seq % n_elem == 0 first)
 * form one time block of t_pkt samples. */
typedef struct mixer_array_cfg {
    int32_t n_elem;    /* elements N; a multiple of 32, 32..1024 */
    int32_t grid_x;    /* elements per grid row, a divisor of 32; element a sits at (a % grid_x, a / grid_x) */
    int32_t n_chan;    /* channels per packet C, 1..1024 */
    int32_t t_pkt;     /* samples per packet per channel; must be 32 (one MFMA K step) */
    int32_t hdr_bytes; /* header bytes before the payload (64) */
    int32_t seq_byte;  /* byte offset of the big-endian u32 seq (48) */
    int32_t k_integ;   /* samples per integration (dump); multiple of k_update, <= 65536 */
    int32_t k_update;  /* samples per V-update (one corner-turn + GEMM launch); multiple of t_pkt */
    int32_t ng;        /* image grid side (32 or 64); >= 2*grid_x and >= 2*(n_elem/grid_x) */
    int32_t flags;     /* MIXER_CFG_* bits */
    int64_t pkt_pitch; /* host ring row pitch in bytes (16384) */
} mixer_array_cfg;

#define MIXER_CFG_CUMULATIVE 1 /* also keep the reference's cumulative-since-start f32 cube */
#define MIXER_CFG_NO_IMAGE 2   /* skip gridding + FFT at each integration */
#define MIXER_CFG_NOFLOP 4     /* ingest only: copy heaps, skip all compute (--noflop,
                                  src/rx_host_corr.cu:161) */
#define MIXER_CFG_AUTOTUNE 8   /* mixer_create runs mixer_autotune when the size class has no
                                  tuning yet */
#define MIXER_CFG_FRAME_DV 16  /* keep an exact frame grid: mixer_read_frame returns the image of
                                  everything integrated since the previous frame (the reference's
                                  dV monitor, src/rx_host_corr.cu:218-222), and
                                  mixer_publish_image publishes it */
#define MIXER_CFG_NO_WRAP_CHECK 32 /* skip the imaginary -128 test (mixer_stats.wrap128 stays 0):
                                  for sources that clip to [-127, 127] as the reference's
                                  transmitter does (src/tx_host.cu:54-55); about 5 % of the
                                  256 x 128 correlator. No reference counterpart. */

void mixer_default_cfg(mixer_array_cfg* cfg);
int mixer_validate_cfg(const mixer_array_cfg* cfg);
size_t mixer_payload_bytes(const mixer_array_cfg* cfg); /* 2*n_chan*t_pkt */
size_t mixer_wire_bytes(const mixer_array_cfg* cfg);    /* hdr_bytes + payload */

const char* mixer_version(void);
const char* mixer_strerror(int status);
int mixer_device_count(int* n); /* visible gfx950 devices */

/* ---- operator entry points on device pointers ---------------------------------------------
 * stream is a hipStream_t (NULL = the default stream). Shapes come from cfg; n_tb is the
 * number of time blocks (each t_pkt samples) in the call.
 *
 * Corner-turn: replaces k_packAB (src/rx_host_corr.cu:56-72).
 *   d_heaps  header-stripped payloads [n_tb][n_elem][payload] (the reference's g_bb layout)
 *   d_tiles  int8 element x time tiles [n_chan][n_tb][2][n_elem][32]: for channel c and time
 *            block tb, a re tile [n_elem][32] followed by an im tile [n_elem][32]
 *            (mixer_tiles_bytes). The reference's B = conj(X) is derived inside the GEMM.
 *   d_wrap   optional u32 counter (may be NULL): incremented once per workgroup that saw an
 *            imaginary sample of -128. */
size_t mixer_tiles_bytes(const mixer_array_cfg* cfg, int n_tb);
int mixer_corner_turn(void* stream, const mixer_array_cfg* cfg, const uint8_t* d_heaps, int n_tb,
                      int8_t* d_tiles, uint32_t* d_wrap);

/* Correlate: replaces cgemm_blk_t<ACC> (src/cgemm.cuh:28-69) on the lower triangle.
 *   d_acc    int32 accumulator, mixer_acc_bytes(cfg) bytes: [n_slots][n_chan][n_tiles][2][32][32]
 *            with tile t = I*(I+1)/2 + J for 32x32 element blocks I >= J, plane 0 = Re, 1 = Im,
 *            row = element of block I, column = element of block J, and n_slots partial sums
 *            (split over time) that mixer_integrate adds. accumulate=0 overwrites (ACC=false),
 *            1 adds into it (ACC=true). Diagonal tiles hold the full 32x32 block. */
size_t mixer_acc_bytes(const mixer_array_cfg* cfg);
int mixer_acc_slots(const mixer_array_cfg* cfg);
int mixer_xcorr(void* stream, const mixer_array_cfg* cfg, const int8_t* d_tiles, int n_tb,
                int32_t* d_acc, int accumulate);

/* Integrate: replaces k_flush_cube (src/rx_host_corr.cu:86-95). Sums the accumulator slots in
 * int32 (exact), converts once to f32, and writes the full square
 *   d_v [n_chan][2][n_elem][n_elem] f32 (the --dump record layout), upper triangle = conj(lower).
 * accumulate=0 stores V; 1 adds V into d_v (the reference's cumulative cube). */
int mixer_integrate(void* stream, const mixer_array_cfg* cfg, const int32_t* d_acc, float* d_v,
                    int accumulate);

/* Fused corner-turn + correlate: k_packAB + cgemm_blk_t<ACC> (src/rx_host_corr.cu:61-72,
 * src/cgemm.cuh:28-69) in one launch. The correlator stages each 32-sample element row
 * straight from the header-free heaps d_heaps [n_tb][n_elem][re C*32 | im C*32] (every row is
 * one 32-byte run of a heap plane), so no tiles are written or read. Same accumulator as
 * mixer_xcorr. d_wrap (optional) counts workgroups that staged an imaginary -128. */
int mixer_xcorr_heaps(void* stream, const mixer_array_cfg* cfg, const uint8_t* d_heaps, int n_tb,
                      int32_t* d_acc, int accumulate, uint32_t* d_wrap);

/* Image: replaces k_diff + k_grid_cube + cufftExecC2C + the host fftshift/Re/min-max
 * (src/rx_host_corr.cu:74-85,205-206,219-235). Grids the integration's V by baseline
 * (dy, dx) = ((qa-qb) mod ng, (pa-pb) mod ng) in exact integer arithmetic, runs a forward,
 * unnormalised ng x ng FP32 FFT per channel in LDS, and writes
 *   d_img [n_chan][ng][ng] f32 = Re F[(y+ng/2)%ng][(x+ng/2)%ng], d_vmax/d_vmin [n_chan]. */
int mixer_image(void* stream, const mixer_array_cfg* cfg, const int32_t* d_acc, float* d_img,
                float* d_vmax, float* d_vmin);

/* Image without any device allocation: no counterpart in the reference. For callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same inputs, outputs and results as mixer_image, but the int64 grid workspace, the integrate
 * segment list and the FFT twiddles live in one static device array of the library (16 MiB of
 * grid plus 4.6 KiB of tables, held on every GPU the library runs on until the process ends):
 * n_chan * ng * ng * 16 bytes must not exceed 16 MiB (256 channels at ng = 64, 1024 at ng = 32),
 * or MIXER_ERR_ARG. Every call on a GPU uses that GPU's array: calls that overlap on one GPU
 * (other streams, other threads) race. Each call clears the workspace it uses first. */
int mixer_image_noalloc(void* stream, const mixer_array_cfg* cfg, const int32_t* d_acc, float* d_img,
                        float* d_vmax, float* d_vmin);

/* Launch tuning (the autotuner's search space), per array size class n_elem:
 *   slots     time split of one correlator call into int32 partial sums (1..8)
 *   prefetch  correlator time-loop pipeline: 1 = register staging one 32-sample block ahead,
 *             3 = register staging two blocks ahead, 5 = register staging three blocks ahead
 *             (fold and blocked layouts), 2 or 4 = direct global->LDS loads, that many blocks
 *             per barrier
 *   ct_chan   corner-turn channels per workgroup (8, 16 or 32)
 *   staging   how contexts feed the correlator: 1 = straight from the heaps (the corner-turn
 *             is fused into the correlator's loads, mixer_xcorr_heaps), 2 = a separate
 *             corner-turn pass into tiles (mixer_corner_turn + mixer_xcorr)
 *   mfma      MFMA tile shape: 32 = v_mfma_i32_32x32x32_i8 (32x32 tiles, one 32-sample block per
 *             instruction), 16 = v_mfma_i32_16x16x64_i8 (each 32x32 tile as four 16x16 tiles,
 *             two 32-sample blocks per instruction and per barrier; arrays of <= 256 elements,
 *             on the FOLD layout, prefetch ignored)
 *   integrate contexts on the BLOCKED layout (384..1024 elements) whose integration is one
 *             correlator launch of one time slot (k_update = k_integ, prefetch 3, staging 1):
 *             2 = write V and grid it from the correlator's epilogue (no int32 accumulator round
 *             trip, no integrate launch), 1 = the separate integrate kernel. mixer_xcorr and
 *             mixer_xcorr_heaps always fill the accumulator.
 * 0 in a field means automatic: staging 1 with prefetch 3, the 32x32x32 shape (the measured
 * fastest choice on MI355X for both the fold and the blocked correlator) and integrate 2. The table is
 * process-wide. slots changes the accumulator
 * layout (mixer_acc_slots, mixer_acc_bytes): set it before allocating accumulators or creating
 * contexts. No reference counterpart (the reference fixes its launch shape at compile time,
 * src/cgemm.cuh:18-24). */
typedef struct mixer_tuning {
    int32_t slots;
    int32_t prefetch;
    int32_t ct_chan;
    int32_t staging;
    int32_t mfma;
    int32_t integrate;
} mixer_tuning;
int mixer_set_tuning(int n_elem, const mixer_tuning* t);
int mixer_get_tuning(const mixer_array_cfg* cfg, mixer_tuning* t); /* the plan in effect */
/* Autotuner: times (corner-turn +) correlate + integrate + image of one integration of cfg on
 * synthetic device data for every (staging, slots, prefetch) candidate in the 32x32x32 shape and,
 * for arrays of <= 256 elements, every (staging, slots) candidate in the 16x16x64 shape, then
 * every ct_chan if the separate corner-turn won, on device `device`; stores the fastest in the
 * tuning table and returns it in *best. */
int mixer_autotune(const mixer_array_cfg* cfg, int device, mixer_tuning* best);

/* Plain per-channel-launch form (baseline): one corner-turn launch and one full-square GEMM
 * launch per channel, exactly the reference's 4-real-product math (cgemm.cuh:58-63) on the
 * whole square, no lower-triangle saving, no persistence. Output d_vre/d_vim [n_chan][N][N]
 * int32 (overwrite). d_tiles is scratch of mixer_tiles_bytes(cfg, n_tb). */
int mixer_xcorr_plain(void* stream, const mixer_array_cfg* cfg, const uint8_t* d_heaps, int n_tb,
                      int8_t* d_tiles, int32_t* d_vre, int32_t* d_vim);

/* ---- engine: the reference's gpu_thread + publish loop as a context -------------------------
 * (src/rx_host_corr.cu:143-175 ingest/correlate, :213-239 monitor/dump). */
typedef struct mixer_ctx mixer_ctx;

typedef struct mixer_stats {
    uint64_t heaps_in;      /* packets accepted */
    uint64_t time_blocks;   /* complete time blocks correlated */
    uint64_t vupdates;      /* V-updates launched (corner-turn + GEMM) */
    uint64_t integrations;  /* integrations completed (V + image ready) */
    uint64_t imissed;       /* time blocks dropped: no free slot (src/rx_host_corr.cu:140) */
    uint64_t misalign;      /* submissions whose first seq % n_elem != 0 (src/rx_host_corr.cu:135) */
    uint64_t seq_gaps;      /* time blocks dropped because a seq inside them was missing */
    uint64_t wrap128;       /* corner-turn or correlator workgroups that saw an imaginary -128 */
    double last_dump_latency_ms; /* arrival of the integration's last heap -> V/image ready, for
                                    the latest probed integration (one probe in flight) */
} mixer_stats;

int mixer_create(const mixer_array_cfg* cfg, int device, mixer_ctx** out);
int mixer_destroy(mixer_ctx* ctx);
int mixer_get_cfg(mixer_ctx* ctx, mixer_array_cfg* cfg);

/* Input seam (src/rx_host_corr.cu:136-141,159). pkts points to npkts packets at a pitch of
 * `pitch` bytes in host memory (pinned memory gives an asynchronous copy). npkts must be a
 * multiple of n_elem and the first seq a multiple of n_elem; otherwise the leading partial
 * group is skipped and misalign is counted, as the reference's drain thread does. Packets
 * are checked for consecutive seq; a time block with a gap is dropped (seq_gaps). The header
 * strip is one hipMemcpy2DAsync per run of consecutive valid time blocks (at most 16384
 * packets, within one ring buffer) on the context's copy stream. *token receives a completion
 * token: the caller's buffer is free for reuse once mixer_token_done/wait reports it (the
 * reference returns its packets to the NIC before the copy has run; this API does not).
 * At most 256 tokens are outstanding per context: a call that asks for a token while the
 * token 256 before it is still pending first waits for that one, so a token reported done
 * is always complete. arrival_ns (steady clock) of the last packet feeds the dump-latency
 * statistics. */
int mixer_submit_heaps(mixer_ctx* ctx, const uint8_t* pkts, size_t pitch, int npkts,
                       uint64_t arrival_ns, uint64_t* token);
/* Bounded ingest, the reference's fixed slot pool (NSLOT = 96, src/rx_host_corr.cu:46,120-121):
 * with n_slots > 0 every time block that mixer_submit_heaps accepts holds one slot until its
 * header-strip copy has completed, and a time block that finds no free slot is dropped and
 * counted in mixer_stats.imissed (src/rx_host_corr.cu:139-140). Slots are returned at the start
 * of each mixer_submit_heaps call whose copies have completed by then (the reference's
 * gpu_thread polls its in-flight slots the same way, src/rx_host_corr.cu:154). 0 (the default)
 * is unbounded: nothing is dropped and the copy stream queues without limit. */
int mixer_set_ingest_slots(mixer_ctx* ctx, int n_slots);
/* Device-resident input: header-stripped payloads [n_tb][n_elem][payload] already in HBM
 * (the replayed-ring benchmark path). The buffer must stay valid until the token completes.
 * For both submit calls token may be NULL: then no completion event is recorded (an event
 * record idles the compute stream for a few microseconds), for callers that never reuse the
 * buffer or order its reuse themselves (mixer_sync). */
int mixer_submit_device(mixer_ctx* ctx, const uint8_t* d_payloads, int n_tb, uint64_t arrival_ns,
                        uint64_t* token);
int mixer_token_done(mixer_ctx* ctx, uint64_t token); /* 1 done, 0 pending, <0 error */
int mixer_token_wait(mixer_ctx* ctx, uint64_t token);
int mixer_sync(mixer_ctx* ctx);                       /* wait for all submitted work */

/* Output seams. mixer_dump copies the most recently completed integration's V cube
 * ([n_chan][2][n_elem][n_elem] f32, the reference's --dump record) to host memory; with
 * cumulative=1 it copies the cube summed over all integrations since creation, which is the
 * reference's record (requires MIXER_CFG_CUMULATIVE). *integ_index receives the number of
 * completed integrations the snapshot covers (0 = none yet, nothing copied). The snapshot is
 * consistent: it is taken between integrations, never while one is being written. */
int mixer_dump(mixer_ctx* ctx, float* host_out, int cumulative, uint64_t* integ_index);
/* The latest image cube: img [n_chan][ng][ng], vmax/vmin [n_chan] (any may be NULL). */
int mixer_read_image(mixer_ctx* ctx, float* img, float* vmax, float* vmin, uint64_t* integ_index);
/* Monitor frame (needs MIXER_CFG_FRAME_DV): the image of dV, the sum of V over all integrations
 * completed since the previous mixer_read_frame (the reference images dV = cube - previous
 * frame's cube, src/rx_host_corr.cu:218-222), gridded in exact int64 and transformed once.
 * *n_in_frame receives the number of integrations the frame covers (0: nothing copied). */
int mixer_read_frame(mixer_ctx* ctx, float* img, float* vmax, float* vmin, uint64_t* n_in_frame);

/* Corr32 shared-memory image cube (src/rx_host_corr.cu:97-105): {u32 magic=0x43523343, w, h,
 * nch; u64 write_seq; f32 img[nch*h*w]; f32 vmax[nch]; f32 vmin[nch]} = 525336 bytes at the
 * default geometry. mixer_publish_image publishes the latest integration's image, or with
 * MIXER_CFG_FRAME_DV the frame since the previous publish (nothing when no integration
 * completed since). write_seq has bit 63 set while a frame is being written and holds
 * frame+1 (release order) once it is complete, so readers that check the bit never see a torn
 * frame and the reference's peek32 viewer reads the region unchanged. */
size_t mixer_corr32_bytes(const mixer_array_cfg* cfg);
int mixer_corr32_init(void* region, const mixer_array_cfg* cfg);
int mixer_publish_image(mixer_ctx* ctx, void* region, size_t region_bytes);

int mixer_get_stats(mixer_ctx* ctx, mixer_stats* st);

/* Kernel timing on the context's compute stream: HIP events bracket every launch of the
 * stages in `enable`, a bit mask (bit s = stage s; 0 = off, 0xF = all). Stage ids: 0
 * corner-turn, 1 GEMM, 2 integrate, 3 image. */
int mixer_set_timing(mixer_ctx* ctx, int enable);
int mixer_get_timing(mixer_ctx* ctx, int stage, double* total_ms, uint64_t* launches);
void* mixer_compute_stream(mixer_ctx* ctx);

/* ---- batched correlation of many independent arrays --------------------------------------
 * (no reference counterpart: the reference correlates one fixed array per process). A batch
 * holds n_jobs arrays of any supported geometry; each run correlates one integration
 * (cfg.k_integ samples) of every job with (optionally) one corner-turn launch, two persistent
 * correlator launches (arrays above / up to 256 elements) whose workgroups pull (job, channel
 * group, super-block, time slot) work items from
 * per-XCD queues (whole channels per queue, largest first; an idle workgroup takes from the
 * other queues), one integrate + gridding launch and one image launch. The
 * launch sequence is captured once as a hipGraph and replayed. Tuning:
 *   pack       small arrays (<= 256 elements): channels packed per correlator workgroup:
 *              0 = as many as fit (8 waves), 1 = one channel per workgroup
 *   tb_target  32-sample blocks per work item when a job is split in time (time slots are
 *              ceil(n_tb / tb_target), at most 8, only for jobs with fewer than 512 items)
 *   use_graph  1 = replay a captured hipGraph, 0 = plain launches on the batch stream
 *   staging    1 = the correlator stages straight from the heaps (no corner-turn launch),
 *              2 = a separate corner-turn launch into tiles; 0 = automatic: 1 (the 64-array
 *              batch runs 108 ms fused, 121 ms with corner-turned tiles above 256 elements;
 *              fixed at creation)
 *   mfma       MFMA tile shape of the fold correlator (arrays up to 256 elements): 32 =
 *              v_mfma_i32_32x32x32_i8, 16 = v_mfma_i32_16x16x64_i8 (two-block stages); 0 = 32 */
typedef struct mixer_batch mixer_batch;
typedef struct mixer_batch_tuning {
    int32_t pack;
    int32_t tb_target;
    int32_t use_graph;
    int32_t staging;
    int32_t mfma;
} mixer_batch_tuning;
typedef struct mixer_batch_info {
    int32_t n_jobs;
    int32_t n_items;        /* correlator work items */
    int32_t n_wgs;          /* persistent correlator workgroups */
    int32_t graph;          /* 1 if runs replay a hipGraph */
    uint64_t payload_bytes; /* input bytes of one run (all jobs) */
    uint64_t exec_ops;      /* int8 MFMA operations of one run */
    uint64_t device_bytes;  /* device memory held by the batch */
} mixer_batch_info;

void mixer_batch_default_tuning(mixer_batch_tuning* t);
/* cfgs[j]: geometry of job j (k_integ samples per run; k_update is ignored). tuning may be NULL. */
int mixer_batch_create(const mixer_array_cfg* cfgs, int n_jobs, int device,
                       const mixer_batch_tuning* tuning, mixer_batch** out);
int mixer_batch_destroy(mixer_batch* b);
/* d_heaps[j]: job j's header-stripped heaps [k_integ/32][n_elem][payload] in device memory;
 * they must stay valid until the run completes (mixer_batch_sync). */
int mixer_batch_run(mixer_batch* b, const uint8_t* const* d_heaps);
int mixer_batch_sync(mixer_batch* b);
void* mixer_batch_stream(mixer_batch* b);
int mixer_batch_get_info(mixer_batch* b, mixer_batch_info* info);
/* Results of the last completed run of job j: V [n_chan][2][n_elem][n_elem] f32, the image
 * cube, and (device pointer) the job's int32 accumulator with its number of time slots. */
int mixer_batch_dump(mixer_batch* b, int job, float* host_v);
int mixer_batch_read_image(mixer_batch* b, int job, float* img, float* vmax, float* vmin);
int mixer_batch_acc(mixer_batch* b, int job, const int32_t** d_acc, int* n_slots);
/* Autotuner for a batch: times runs of b on its current inputs (the last d_heaps given to
 * mixer_batch_run, which must have been called once) for each mfma x pack x tb_target candidate, each
 * as the median over 3 blocks of reps back-to-back runs after a warm-up run, and leaves b on the
 * fastest, returned in *best. It moves off the tuning b had only for a candidate at least 2 %
 * faster, confirmed by timing the two alternately once the search is over (the GPU's clock
 * drifts during a search of seconds), so noise cannot leave b slower than it started. */
int mixer_batch_autotune(mixer_batch* b, int reps, mixer_batch_tuning* best);
/* Re-plan b under tuning t (its staging field is ignored: fixed at creation; the work list and
 * graph are rebuilt), or read the tuning in effect. */
int mixer_batch_set_tuning(mixer_batch* b, const mixer_batch_tuning* t);
int mixer_batch_get_tuning(mixer_batch* b, mixer_batch_tuning* t);
/* Fills n bytes of device memory with seeded uniform int8 values in [-127, 127]. */
int mixer_fill_random(void* stream, void* d_ptr, size_t bytes, uint64_t seed);

/* ---- serving loop: many arrays, integrations admitted as their data arrives ----------------
 * No reference counterpart (the reference serves one array per process, its gpu_thread,
 * src/rx_host_corr.cu:143-175). A server holds the arrays admitted under an HBM budget; for
 * each it keeps the int32 accumulator, the f32 V cube (double-buffered), the image grid and an
 * input ring resident on the device. mixer_server_submit hands over one chunk of an array's
 * header-free heaps (a whole number of V-updates, device memory, e.g. written there by a NIC)
 * with the arrival time of its last heap; the server correlates it on the array's stream and,
 * when the chunk completes an integration, integrates + images it and copies the image cube to
 * pinned host memory. A completion thread timestamps each such dump: latency = dump complete on
 * the host - arrival of the integration's last chunk (steady clock).
 * Arrays are spread over n_streams compute streams (0 = 1: correlator kernels of different
 * arrays running side by side spread each other's workgroups over the XCDs and lose the
 * per-channel L2 reuse; 8 arrays of 256..1024 elements at 500 GB/s offered are served at
 * 506 GB/s with one stream, 443 GB/s with four). A chunk that arrives while its
 * array's previous correlator launch is still running is held, and the chunks contiguous with
 * it join it in one longer launch (never across an integration); held chunks are launched when
 * a later chunk finds the array idle, when its integration completes, or by drain/token_wait.
 * Submission is single-threaded per array. */
typedef struct mixer_server mixer_server;
typedef struct mixer_server_stats {
    uint64_t hbm_budget;    /* bytes */
    uint64_t hbm_used;      /* bytes held by admitted arrays */
    uint64_t admitted;      /* arrays admitted so far */
    uint64_t rejected;      /* admissions refused for budget */
    uint64_t chunks;        /* chunks submitted */
    uint64_t bytes_in;      /* payload bytes submitted */
    uint64_t integrations;  /* dumps completed on the host */
    uint64_t lat_n;         /* latency samples */
    double lat_p50_ms, lat_p99_ms, lat_max_ms, lat_mean_ms;
    uint64_t launches;      /* correlator launches (fewer than updates when chunks coalesce) */
} mixer_server_stats;

int mixer_server_create(int device, uint64_t hbm_budget, int n_streams, mixer_server** out);
int mixer_server_destroy(mixer_server* srv);
/* Device bytes an admitted array of this geometry holds (its admission cost). */
uint64_t mixer_server_footprint(const mixer_array_cfg* cfg);
/* Admit an array (cfg.k_update = samples per chunk unit, cfg.k_integ = samples per
 * integration); MIXER_ERR_BUDGET if its footprint does not fit the remaining budget. */
int mixer_server_admit(mixer_server* srv, const mixer_array_cfg* cfg, int* array_id);
int mixer_server_release(mixer_server* srv, int array_id);
/* The array's device input ring: n_chunks chunks of chunk_bytes (= k_update/32 time blocks of
 * header-free heaps each) that a producer may fill and then submit. */
int mixer_server_input(mixer_server* srv, int array_id, uint8_t** d_ring, int* n_chunks,
                       size_t* chunk_bytes);
/* Submit n_tb time blocks (a multiple of k_update/32) at device address d_heaps. The memory is
 * read asynchronously: *token (optional) completes when it may be reused
 * (mixer_server_token_wait). */
int mixer_server_submit(mixer_server* srv, int array_id, const uint8_t* d_heaps, int n_tb,
                        uint64_t arrival_ns, uint64_t* token);
int mixer_server_token_wait(mixer_server* srv, int array_id, uint64_t token);
/* Chunk coalescing on (1, the default) or off (0: every chunk is launched when submitted). */
int mixer_server_set_coalescing(mixer_server* srv, int enable);
/* Wait until everything submitted is dumped. */
int mixer_server_drain(mixer_server* srv);
int mixer_server_get_stats(mixer_server* srv, mixer_server_stats* st);
void mixer_server_reset_latency(mixer_server* srv);
/* Last dumped integration of an array: the host image copy (img [C][ng][ng], vmax/vmin [C]),
 * V from device (host_v [C][2][N][N]) and its integration count. */
int mixer_server_read_image(mixer_server* srv, int array_id, float* img, float* vmax, float* vmin,
                            uint64_t* integ_index);
int mixer_server_dump(mixer_server* srv, int array_id, float* host_v, uint64_t* integ_index);

/* Load generator: Poisson arrivals of chunks (k_update samples) from every admitted array for
 * duration_s seconds at aggregate_gbps of payload, shared sample clock (each array's rate is
 * proportional to its bytes per sample), data from the arrays' input rings (filled with seeded
 * int8 noise at admission). Submits each chunk at its arrival time from the calling thread and
 * reports throughput and latency of the dumps that completed. */
typedef struct mixer_serve_spec {
    double duration_s;
    double aggregate_gbps;
    uint64_t seed;
} mixer_serve_spec;
typedef struct mixer_serve_report {
    uint64_t chunks, integrations, late_submits; /* late: submitted > 1 ms after arrival */
    double offered_gbps, achieved_gbps, wall_s;
    double lat_p50_ms, lat_p99_ms, lat_max_ms;
} mixer_serve_report;
int mixer_serve_run(mixer_server* srv, const mixer_serve_spec* spec, mixer_serve_report* rep);

/* ---- HIP plumbing for hosts without a HIP binding of their own -------------------------------
 * (no reference counterpart: the reference calls the CUDA runtime directly). The Python tests
 * and bench.py reach device memory, streams and events only through these, so one HIP
 * runtime serves the whole process. kind: 0 host->device, 1 device->host, 2 device->device. */
int mixer_set_device(int device);
int mixer_dev_alloc(size_t bytes, void** ptr);
int mixer_dev_free(void* ptr);
int mixer_host_alloc(size_t bytes, void** ptr); /* pinned host memory */
int mixer_host_free(void* ptr);
int mixer_memcpy(void* dst, const void* src, size_t bytes, int kind, void* stream);
int mixer_memset(void* ptr, int value, size_t bytes, void* stream);
int mixer_stream_create(void** stream);
int mixer_stream_destroy(void* stream);
int mixer_stream_sync(void* stream);
int mixer_event_create(void** ev);
int mixer_event_destroy(void* ev);
int mixer_event_record(void* ev, void* stream);
int mixer_event_elapsed_ms(void* start, void* stop, float* ms); /* synchronises on stop */
int mixer_device_sync(void);

/* ---- synthetic F-engine source (replaces the DAQIRI capture and tx_host) ------------------
 * Signal model of k_sky (src/tx_host.cu:41-60) plus seeded noise: element a at
 * (px, qy) = (a % grid_x, a / grid_x); x_a(c,t) = clip127(rint(25*[2 + Apl(c) e^{i pi (px l + qy m)}
 * + sigma (n_r + i n_i)])), Apl(c) = 0.71*2*sqrt(2c/(nch_gen-1)), n ~ N(0,1) from a counter-based
 * generator keyed by (seed, element, channel, absolute sample). n_src extra point sources
 * (l_k, m_k, amp_k) may be added. Packets are written for seq in [seq0, seq0+npkts) at
 * `pitch` bytes apart: header (seq big-endian at seq_byte) + planar payload. */
typedef struct mixer_sky {
    double planet_l, planet_m; /* planet direction cosines */
    double sigma;              /* noise std in units of the 25-LSB scale (0 = noise-free) */
    double star_amp;           /* 2.0 in the reference */
    double planet_scale;       /* 0.71 in the reference (planet amplitude factor) */
    int32_t nch_gen;           /* 256: the generator's channel sweep (src/config.h:28) */
    int32_t n_src;             /* extra point sources, 0..4 */
    double src_l[4], src_m[4], src_amp[4];
} mixer_sky;

void mixer_default_sky(mixer_sky* sky);
/* Synthetic packets seq0 .. seq0+npkts-1 at `pitch`: the wire header (mixer_build_header with
 * the template set by mixer_synth_set_wire, default mixer_default_wire) and the k_sky payload
 * (src/tx_host.cu:41-60,136-143) with seeded Gaussian noise. Host only. */
int mixer_synth_packets(const mixer_array_cfg* cfg, const mixer_sky* sky, uint64_t seed,
                        uint64_t seq0, int npkts, uint8_t* out, size_t pitch, int nthreads);

/* Wire header of the reference transmitter (src/tx_host.cu:62-79, src/config.h:10-14,53-58):
 * eth (dst MAC, src MAC, ethertype IPv4), IPv4 (version 4, IHL 5, TTL 64, protocol UDP, total
 * length = wire bytes - 14, header checksum), UDP (source = destination port, length = wire
 * bytes - 34, checksum 0), zeros to hdr_bytes, seq big-endian at seq_byte. Lengths above
 * 65535 (payloads beyond a jumbo frame) are clamped to 65535. */
typedef struct mixer_wire {
    uint8_t eth_dst[6];
    uint8_t eth_src[6];   /* zero: the NIC fills it on transmit (src/config.h:11) */
    uint32_t ip_src;      /* host byte order */
    uint32_t ip_dst;
    uint16_t udp_port;
    uint16_t reserved;
} mixer_wire;
void mixer_default_wire(mixer_wire* w); /* a0:88:c2:0d:5e:28, 10.0.0.1 -> 10.0.0.2, port 4096 */
int mixer_build_header(const mixer_array_cfg* cfg, const mixer_wire* w, uint32_t seq, uint8_t* hdr);
/* Header template used by mixer_synth_packets from now on (process-wide). */
int mixer_synth_set_wire(const mixer_wire* w);
/* The orbiting planet of the reference transmitter (src/tx_host.cu:98-100,116-120): copies *in to
 * *out with the planet at (l, m) = rho (cos wt, sin wt), w = 2 pi / torbit_s, t_s seconds after
 * motion started (t_s = 0: l = rho, m = 0, the transmitter's first frame, src/tx_host.cu:106).
 * mixer_corr --rho/--torbit/--refresh-hz re-generates its replay slots with it. */
int mixer_orbit_sky(const mixer_sky* in, double rho, double torbit_s, double t_s, mixer_sky* out);

#ifdef __cplusplus
}
#endif
#endif /* MIXER_H */
