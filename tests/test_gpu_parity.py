"""GPU parity of the HIP path against the oracle, through the C ABI.

Bit-exact for the corner-turn bytes and the int32 V; the f32 integration within one int32->f32
rounding (<= 2^-24 relative, tolerance 1e-6 relative to FP64); images against the FP64 grid +
FP64 DFT oracle within 1e-5 of the channel's peak |image|.
"""
import os
import sys

import numpy as np
import pytest

import mixer_amd as M
import pyoracle as O

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(0x4D495839)


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rand_heaps(n_tb, N, C, lo=-127, hi=127, rng=RNG):
    return rng.integers(lo, hi + 1, size=(n_tb, N, 64 * C), dtype=np.int16).astype(np.int8).view(np.uint8)


def cfg_for(N, C, k_update, gx=None, ng=None, **kw):
    gx = gx or (16 if N % 16 == 0 and N // 16 <= 16 else 32)
    gy = N // gx
    ng = ng or max(32, 1 << int(np.ceil(np.log2(2 * max(gx, gy)))))
    return M.default_cfg(n_elem=N, grid_x=gx, n_chan=C, k_update=k_update, k_integ=k_update, ng=ng,
                         pkt_pitch=max(16384, 64 + 64 * C), **kw)


@pytest.fixture(scope="module", autouse=True)
def device():
    assert M.device_count() >= 1, "no gfx950 device visible to libmixer"
    M.check(M.lib().mixer_set_device(0))
    yield


# ---------------------------------------------------------------- corner-turn
@pytest.mark.parametrize("N,C,n_tb", [(256, 128, 4), (64, 40, 3), (32, 1, 2), (1024, 8, 2)])
def test_corner_turn_bytes(N, C, n_tb):
    log(f"corner_turn N={N} C={C} n_tb={n_tb}")
    cfg = cfg_for(N, C, 32 * n_tb)
    bb = rand_heaps(n_tb, N, C, -128, 127)
    d_bb = M.DeviceBuffer.from_array(bb)
    d_t = M.DeviceBuffer(M.tiles_bytes(cfg, n_tb))
    d_w = M.DeviceBuffer(4)
    d_w.zero()
    M.corner_turn(cfg, d_bb, n_tb, d_t, d_w)
    t = d_t.download(np.int8, (C, n_tb, 2, N, 32))
    are, aim, _, _ = O.unpack(bb, n_tb, N, C, 32)
    # oracle [C][N][K] -> tiles [C][tb][N][32]
    exp_re = are.reshape(C, N, n_tb, 32).transpose(0, 2, 1, 3)
    exp_im = aim.reshape(C, N, n_tb, 32).transpose(0, 2, 1, 3)
    assert np.array_equal(t[:, :, 0], exp_re) and np.array_equal(t[:, :, 1], exp_im)
    w = d_w.download(np.uint32, (1,))[0]
    assert w > 0  # random bytes include imaginary -128
    # clean input: no wrap flagged
    bb2 = rand_heaps(n_tb, N, C)
    d_bb.upload(bb2)
    d_w.zero()
    M.corner_turn(cfg, d_bb, n_tb, d_t, d_w)
    assert d_w.download(np.uint32, (1,))[0] == 0


# ---------------------------------------------------------------- GEMM
def gpu_v(cfg, bb, n_tb, accumulate_parts=1):
    """corner-turn + xcorr over n_tb blocks, split into accumulate_parts V-updates."""
    N, C = cfg.n_elem, cfg.n_chan
    per = n_tb // accumulate_parts
    c2 = cfg_for(N, C, 32 * per, gx=cfg.grid_x, ng=cfg.ng)
    d_acc = M.DeviceBuffer(M.acc_bytes(c2))
    d_t = M.DeviceBuffer(M.tiles_bytes(c2, per))
    for p in range(accumulate_parts):
        d_bb = M.DeviceBuffer.from_array(np.ascontiguousarray(bb[p * per:(p + 1) * per]))
        M.corner_turn(c2, d_bb, per, d_t)
        M.xcorr(c2, d_t, per, d_acc, accumulate=p > 0)
    acc = d_acc.download(np.int32, (M.acc_bytes(c2) // 4,))
    return c2, d_acc, acc


@pytest.mark.parametrize("N,C,n_tb,parts", [
    (256, 4, 8, 1), (256, 3, 16, 2), (64, 5, 4, 1), (128, 3, 6, 3), (160, 2, 4, 1), (32, 2, 3, 1),
    (96, 2, 2, 1), (384, 2, 2, 1), (512, 2, 3, 1), (1024, 1, 2, 1),
])
def test_xcorr_bit_exact(N, C, n_tb, parts):
    log(f"xcorr N={N} C={C} n_tb={n_tb} parts={parts}")
    cfg = cfg_for(N, C, 32 * n_tb)
    bb = rand_heaps(n_tb, N, C)
    c2, _, acc = gpu_v(cfg, bb, n_tb, parts)
    vre, vim = M.unpack_acc(acc, c2)
    ore, oim = O.v_full_from_heaps(bb, n_tb, N, C, 32)
    assert np.array_equal(vre, ore) and np.array_equal(vim, oim)


def test_xcorr_extremes_and_single_element():
    """All +-127 (int32 headroom) and one non-zero element."""
    N, C, n_tb = 256, 2, 64
    bb = np.where(RNG.random((n_tb, N, 64 * C)) < 0.5, 127, -127).astype(np.int8).view(np.uint8)
    cfg = cfg_for(N, C, 32 * n_tb)
    c2, _, acc = gpu_v(cfg, bb, n_tb)
    vre, vim = M.unpack_acc(acc, c2)
    ore, oim = O.v_full_from_heaps(bb, n_tb, N, C, 32)
    assert np.array_equal(vre, ore) and np.array_equal(vim, oim)
    assert vre[0, 0, 0] == 2 * 127 * 127 * 32 * n_tb
    bb = np.zeros((2, N, 64 * C), np.uint8)
    bb[:, 77, :] = rand_heaps(2, 1, C)[:, 0, :]
    c2, _, acc = gpu_v(cfg_for(N, C, 64), bb, 2)
    vre, vim = M.unpack_acc(acc, c2)
    ore, oim = O.v_full_from_heaps(bb, 2, N, C, 32)
    assert np.array_equal(vre, ore) and np.array_equal(vim, oim)
    assert np.count_nonzero(vre) == C and np.count_nonzero(vim) == 0


K_MAX = 65504  # the longest integration a context accepts (longer ones: MIXER_ERR_OVERFLOW)
LIMIT_CASES = [(64, 3), (256, 4), (1024, 2)]  # fold, QUAD (two time slots), blocked (fused integrate)


def max_magnitude_block(N, C, rng):
    """One 32-sample heap block whose samples are constant in time and as large as int8 allows:
    real -128 or 127, imaginary -127 or 127 (an imaginary -128 wraps when negated)."""
    r = np.where(rng.random((N, C)) < 0.5, -128, 127).astype(np.int8)
    i = np.where(rng.random((N, C)) < 0.5, -127, 127).astype(np.int8)
    r[0] = -128  # at least one autocorrelation reaches the bound
    blk = np.empty((N, 64 * C), np.int8)
    blk[:, :32 * C] = np.repeat(r, 32, axis=1)
    blk[:, 32 * C:] = np.repeat(i, 32, axis=1)
    return blk.view(np.uint8)


def limit_case(N, C):
    """Heaps of K_MAX samples (2047 blocks, an odd count) repeating one maximal block, and the
    exact V: 2047 times the one-block V of the oracle. The largest |V| is the autocorrelation of
    an element with real -128: 65504 (128^2 + 127^2) = 2 129 732 352, 99.2 % of 2^31 - 1."""
    n_tb = K_MAX // 32
    blk = max_magnitude_block(N, C, RNG)
    r1, i1 = O.v_full_from_heaps(np.ascontiguousarray(blk[None]), 1, N, C, 32)
    ore, oim = r1 * n_tb, i1 * n_tb
    assert np.abs(ore).max() == K_MAX * (128 * 128 + 127 * 127) < 2 ** 31
    bb = np.ascontiguousarray(np.broadcast_to(blk, (n_tb, N, 64 * C)))
    return n_tb, bb, np.stack([ore, oim], axis=1).astype(np.float32)


@pytest.mark.parametrize("N,C", LIMIT_CASES)
def test_longest_integration_at_the_int32_limit(N, C):
    """The int32 sums (src/cgemm.cuh:42-45,58-63) at the longest integration, through a context
    (default tuning) and through the operator entry points (mixer_xcorr_heaps +
    mixer_integrate): V is one f32 rounding of the exact sum, bit for bit."""
    n_tb, bb, exact = limit_case(N, C)
    d_bb = M.DeviceBuffer.from_array(bb)
    cfg = cfg_for(N, C, K_MAX)
    log(f"longest integration N={N} C={C} n_tb={n_tb}: context")
    with M.Engine(cfg) as e:
        e.submit_device(d_bb, n_tb, token=False)
        e.sync()
        v, idx = e.dump()
    assert idx == 1 and np.array_equal(v, exact), f"context V differs (N={N})"
    log(f"longest integration N={N} C={C}: mixer_xcorr_heaps + mixer_integrate")
    d_acc = M.DeviceBuffer(M.acc_bytes(cfg))
    d_v = M.DeviceBuffer(C * 2 * N * N * 4)
    M.xcorr_heaps(cfg, d_bb, n_tb, d_acc)
    M.integrate(cfg, d_acc, d_v)
    assert np.array_equal(d_v.download(np.float32, (C, 2, N, N)), exact), f"operator V differs (N={N})"


def test_longest_integration_at_the_int32_limit_batched():
    """The same three arrays as one batch: the planner splits their few channels into time slots
    whose int32 partial sums the batch integrate adds exactly."""
    cases = [limit_case(N, C) for N, C in LIMIT_CASES]
    cfgs = [cfg_for(N, C, K_MAX) for N, C in LIMIT_CASES]
    bufs = [M.DeviceBuffer.from_array(bb) for _, bb, _ in cases]
    log("longest integration: batch of 3 arrays")
    with M.Batch(cfgs) as b:
        b.run(bufs)
        b.sync()
        for j, (_, _, exact) in enumerate(cases):
            assert np.array_equal(b.dump(j), exact), f"batch job {j} {LIMIT_CASES[j]} V differs"


def channel_v(acc, cfg, c):
    """Full-square (Vre, Vim) of channel c from a tile-packed accumulator [slots][C][tiles][2][32][32]."""
    nb = cfg.n_elem // 32
    one = acc.reshape(-1, cfg.n_chan, nb * (nb + 1) // 2 * 2048)[:, c:c + 1]
    c1 = cfg_for(cfg.n_elem, 1, cfg.k_update, gx=cfg.grid_x, ng=cfg.ng)
    vre, vim = M.unpack_acc(np.ascontiguousarray(one).reshape(-1), c1)
    return vre[0], vim[0]


@pytest.mark.parametrize("N,C", [(32, 1024), (256, 1021), (512, 1024)])
def test_most_channels(N, C):
    """The largest channel count a context accepts (1024: 64 KiB per heap element, up to 32 MiB
    per time block) and a ragged 1021 (the last quad of a heap line partly empty) on the fold,
    QUAD and blocked layouts straight from the heaps: channels 0, 1, C/2, C-2 and C-1
    bit-exact against the oracle's correlator (src/cgemm.cuh:28-69)."""
    n_tb = 3
    bb = rand_heaps(n_tb, N, C)
    cfg = cfg_for(N, C, 32 * n_tb)
    log(f"most channels N={N} C={C} n_tb={n_tb}")
    d_bb = M.DeviceBuffer.from_array(bb)
    d_acc = M.DeviceBuffer(M.acc_bytes(cfg))
    M.xcorr_heaps(cfg, d_bb, n_tb, d_acc)
    acc = d_acc.download(np.int32, (M.acc_bytes(cfg) // 4,))
    for c in (0, 1, C // 2, C - 2, C - 1):
        ore, oim, of = O.xengine(bb, n_tb, N, C, 32, c, c + 1, nthreads=min(16, os.cpu_count() or 1))
        assert of == 0
        vre, vim = channel_v(acc, cfg, c)
        assert np.array_equal(vre, ore[0]) and np.array_equal(vim, oim[0]), f"channel {c} differs"


def test_xcorr_minus128_lower_triangle_matches_reference():
    """With imaginary -128 present the reference's V is not Hermitian; the lower triangle and
    the diagonal tiles stay bit-identical to it."""
    N, C, n_tb = 128, 2, 4
    bb = rand_heaps(n_tb, N, C, -128, 127)
    c2, _, acc = gpu_v(cfg_for(N, C, 32 * n_tb), bb, n_tb)
    nb = N // 32
    a = acc.reshape(-1, C, nb * (nb + 1) // 2, 2, 32, 32).astype(np.int64).sum(0)
    ore, oim = O.v_full_from_heaps(bb, n_tb, N, C, 32)
    t = 0
    for I in range(nb):
        for J in range(I + 1):
            assert np.array_equal(a[:, t, 0], ore[:, I * 32:(I + 1) * 32, J * 32:(J + 1) * 32])
            assert np.array_equal(a[:, t, 1], oim[:, I * 32:(I + 1) * 32, J * 32:(J + 1) * 32])
            t += 1


def test_plain_form_full_square():
    N, C, n_tb = 256, 3, 4
    cfg = cfg_for(N, C, 32 * n_tb)
    bb = rand_heaps(n_tb, N, C, -128, 127)  # the full square reproduces the reference even with -128
    d_bb = M.DeviceBuffer.from_array(bb)
    d_t = M.DeviceBuffer(M.tiles_bytes(cfg, n_tb))
    d_r = M.DeviceBuffer(C * N * N * 4)
    d_i = M.DeviceBuffer(C * N * N * 4)
    M.xcorr_plain(cfg, d_bb, n_tb, d_t, d_r, d_i)
    vre = d_r.download(np.int32, (C, N, N))
    vim = d_i.download(np.int32, (C, N, N))
    ore, oim = O.v_full_from_heaps(bb, n_tb, N, C, 32)
    assert np.array_equal(vre, ore) and np.array_equal(vim, oim)


# ---------------------------------------------------------------- integrate
def test_integrate_f32_and_cumulative():
    N, C, n_tb = 256, 3, 8
    bb = rand_heaps(n_tb, N, C)
    c2, d_acc, acc = gpu_v(cfg_for(N, C, 32 * n_tb), bb, n_tb)
    d_v = M.DeviceBuffer(C * 2 * N * N * 4)
    M.integrate(c2, d_acc, d_v)
    v = d_v.download(np.float32, (C, 2, N, N))
    ore, oim = O.v_full_from_heaps(bb, n_tb, N, C, 32)
    exact = np.stack([ore, oim], axis=1)
    assert np.array_equal(v, exact.astype(np.float32))  # one rounding of the exact int32 sum
    rel = np.abs(v.astype(np.float64) - exact) / np.maximum(np.abs(exact), 1)
    assert rel.max() <= 1e-6
    # cumulative: the reference's Vcube += (float)Vi (rx_host_corr.cu:93-94) over two integrations
    M.integrate(c2, d_acc, d_v, accumulate=True)
    v2 = d_v.download(np.float32, (C, 2, N, N))
    assert np.array_equal(v2, exact.astype(np.float32) + exact.astype(np.float32))


# ---------------------------------------------------------------- image
@pytest.mark.parametrize("N,gx,C,ng,n_tb", [(256, 16, 4, 32, 4), (1024, 32, 2, 64, 2), (64, 8, 3, 32, 2),
                                             (32, 4, 3, 16, 2), (64, 8, 2, 16, 2), (256, 16, 2, 64, 2),
                                             (128, 8, 2, 32, 2), (96, 32, 2, 64, 2)])
def test_image_vs_fp64(N, gx, C, ng, n_tb):
    """Every grid size a geometry can ask for (16, 32 in registers, 64) and rectangular arrays
    (8 x 16, 32 x 3 elements) against the FP64 grid + DFT."""
    log(f"image N={N} gx={gx} C={C} ng={ng}")
    cfg = cfg_for(N, C, 32 * n_tb, gx=gx, ng=ng)
    sky = M.default_sky(sigma=0.5, planet_l=0.5, planet_m=0.25)
    pk = M.synth_packets(cfg, sky, seed=3, seq0=0, npkts=n_tb * N)
    bb = np.ascontiguousarray(pk[:, 64:]).reshape(n_tb, N, 64 * C)
    c2, d_acc, acc = gpu_v(cfg, bb, n_tb)
    d_img = M.DeviceBuffer(C * ng * ng * 4)
    d_mx = M.DeviceBuffer(C * 4)
    d_mn = M.DeviceBuffer(C * 4)
    M.image(c2, d_acc, d_img, d_mx, d_mn)
    img = d_img.download(np.float32, (C, ng, ng))
    vmax = d_mx.download(np.float32, (C,))
    vmin = d_mn.download(np.float32, (C,))
    ore, oim = O.v_full_from_heaps(bb, n_tb, N, C, 32)
    oimg, omax, omin, _, _ = O.image_from_v(ore, oim, gx, ng)
    scale = np.abs(oimg).reshape(C, -1).max(axis=1)
    err = (np.abs(img.astype(np.float64) - oimg).reshape(C, -1).max(axis=1) / scale)
    log(f"  image max error / peak = {err.max():.3e}")
    assert err.max() <= 1e-5
    assert np.array_equal(vmax, img.reshape(C, -1).max(1)) and np.array_equal(vmin, img.reshape(C, -1).min(1))
    # star at the centre
    for c in range(C):
        assert np.unravel_index(np.argmax(img[c]), (ng, ng)) == (ng // 2, ng // 2)


# ---------------------------------------------------------------- engine
def heaps_of(pk, N, C):
    return np.ascontiguousarray(pk[:, 64:64 + 64 * C]).reshape(-1, N, 64 * C)


def test_engine_host_packets_dump_and_image():
    N, C, k_upd, upd = 256, 8, 256, 4
    cfg = cfg_for(N, C, k_upd, flags=M.CFG_CUMULATIVE)
    cfg.k_integ = k_upd * upd
    n_tb = cfg.k_integ // 32
    sky = M.default_sky(sigma=0.5)
    pitch = 1024  # > wire bytes (576): exercises the header strip at a pitch
    pk = M.synth_packets(cfg, sky, seed=11, seq0=0, npkts=2 * n_tb * N, pitch=pitch)
    with M.Engine(cfg) as e:
        # feed in uneven pieces: a group boundary inside a call and one spanning calls
        cuts = [0, 3 * N, 3 * N + 5 * N, 20 * N, 2 * n_tb * N]
        for a, b in zip(cuts[:-1], cuts[1:]):
            tok = e.submit_heaps(np.ascontiguousarray(pk[a:b]), pitch, b - a)
            e.token_wait(tok)
        e.sync()
        v, idx = e.dump()
        vc, _ = e.dump(cumulative=True)
        img, vmax, vmin, _ = e.read_image()
        st = e.stats()
    assert idx == 2 and st.integrations == 2 and st.vupdates == 2 * upd and st.misalign == 0
    bb = heaps_of(pk, N, C)
    r1, i1 = O.v_full_from_heaps(np.ascontiguousarray(bb[:n_tb]), n_tb, N, C, 32)
    r2, i2 = O.v_full_from_heaps(np.ascontiguousarray(bb[n_tb:]), n_tb, N, C, 32)
    assert np.array_equal(v, np.stack([r2, i2], 1).astype(np.float32))
    assert np.array_equal(vc, np.stack([r1, i1], 1).astype(np.float32) + np.stack([r2, i2], 1).astype(np.float32))
    oimg, _, _, _, _ = O.image_from_v(r2, i2, 16, 32)
    assert np.abs(img - oimg).max() / np.abs(oimg).max() <= 1e-5


def test_engine_misalign_and_gap():
    N, C = 64, 4
    cfg = cfg_for(N, C, 64)
    pitch = 640
    sky = M.default_sky()
    pk = M.synth_packets(cfg, sky, seed=5, seq0=N - 7, npkts=7 + 4 * N, pitch=pitch)
    # drop one packet inside the third block
    keep = np.ones(len(pk), bool)
    keep[7 + 2 * N + 10] = False
    pk2 = np.ascontiguousarray(pk[keep])
    with M.Engine(cfg) as e:
        e.token_wait(e.submit_heaps(pk2, pitch, len(pk2)))
        e.sync()
        st = e.stats()
        v, idx = e.dump()
    assert st.misalign == 1 and st.seq_gaps == 1 and st.time_blocks == 2 and idx == 1
    # blocks 0,1 and 3 were complete; the first integration is blocks 0,1
    bb = heaps_of(pk[7:], N, C)
    r, i = O.v_full_from_heaps(np.ascontiguousarray(bb[:2]), 2, N, C, 32)
    assert np.array_equal(v, np.stack([r, i], 1).astype(np.float32))


def test_engine_gap_splits_a_copy_run():
    """One call with a packet missing in block 2 of 9: blocks 0, 1 | 3, 4 share a ring buffer
    (the run before the gap is copied, the run after it starts at the next slot), and the run
    3.. crosses into the next ring buffer. Both integrations are exact."""
    N, C = 64, 4
    cfg = cfg_for(N, C, 128, flags=M.CFG_CUMULATIVE)
    pitch = 640
    pk = M.synth_packets(cfg, M.default_sky(sigma=0.5), seed=9, seq0=0, npkts=9 * N, pitch=pitch)
    keep = np.ones(len(pk), bool)
    keep[2 * N + 17] = False
    with M.Engine(cfg) as e:
        e.token_wait(e.submit_heaps(np.ascontiguousarray(pk[keep]), pitch, int(keep.sum())))
        e.sync()
        st = e.stats()
        v, idx = e.dump()
        vc, _ = e.dump(cumulative=True)
    assert st.seq_gaps == 1 and st.time_blocks == 8 and idx == 2
    bb = heaps_of(pk, N, C)
    r1, i1 = O.v_full_from_heaps(np.ascontiguousarray(bb[[0, 1, 3, 4]]), 4, N, C, 32)
    r2, i2 = O.v_full_from_heaps(np.ascontiguousarray(bb[5:9]), 4, N, C, 32)
    assert np.array_equal(v, np.stack([r2, i2], 1).astype(np.float32))
    assert np.array_equal(vc, np.stack([r1, i1], 1).astype(np.float32) + np.stack([r2, i2], 1).astype(np.float32))


def test_engine_device_submit_matches_host_submit():
    N, C = 256, 16
    cfg = cfg_for(N, C, 512)
    cfg.k_integ = 1024
    n_tb = 32
    pk = M.synth_packets(cfg, M.default_sky(), seed=9, seq0=0, npkts=n_tb * N)
    bb = heaps_of(pk, N, C)
    d = M.DeviceBuffer.from_array(bb)
    with M.Engine(cfg) as e:
        tb_bytes = N * 64 * C
        # uneven device pieces: 5 + 20 + 7 blocks; the first two without completion tokens
        assert e.submit_device(d, 5, offset=0, token=False) == 0
        e.submit_device(d, 20, offset=5 * tb_bytes, token=False)
        e.token_wait(e.submit_device(d, 7, offset=25 * tb_bytes))
        e.sync()
        v, idx = e.dump()
    r, i = O.v_full_from_heaps(bb, n_tb, N, C, 32)
    assert idx == 1 and np.array_equal(v, np.stack([r, i], 1).astype(np.float32))


def test_engine_concurrent_reader_gets_whole_integrations():
    """A reader thread's dumps and images while another thread submits back to back: each is
    exactly the integration whose index it reports (inputs alternate between two buffers)."""
    import threading
    N, C, n_tb = 256, 8, 8
    cfg = cfg_for(N, C, 32 * n_tb)
    bufs = [rand_heaps(n_tb, N, C) for _ in range(2)]
    exp_v, exp_img = [], []
    for bb in bufs:
        r, i = O.v_full_from_heaps(bb, n_tb, N, C, 32)
        exp_v.append(np.stack([r, i], 1).astype(np.float32))
        exp_img.append(O.image_from_v(r, i, cfg.grid_x, cfg.ng)[0])
    d = [M.DeviceBuffer.from_array(bb) for bb in bufs]
    n_integ = 300
    with M.Engine(cfg) as e:
        def submitter():
            for k in range(n_integ):
                e.submit_device(d[k % 2], n_tb)
        th = threading.Thread(target=submitter)
        th.start()
        seen = set()
        while True:
            v, idx = e.dump()
            img, _, _, idx2 = e.read_image()
            if idx:
                assert np.array_equal(v, exp_v[(idx - 1) % 2]), idx
                seen.add(idx)
            if idx2:
                o = exp_img[(idx2 - 1) % 2]
                assert np.abs(img - o).max() / np.abs(o).max() <= 1e-5, idx2
            if not th.is_alive() and idx == n_integ:
                break
        th.join()
        e.sync()
        v, idx = e.dump()
        assert idx == n_integ and np.array_equal(v, exp_v[(n_integ - 1) % 2])
    log(f"concurrent reader: {len(seen)} distinct integrations read")
    assert len(seen) >= 1


# ---------------------------------------------------------------- tuning variants
@pytest.mark.parametrize("N,C", [(256, 3), (128, 2), (1024, 1), (512, 2)])
def test_every_launch_variant_bit_exact(N, C):
    """Every (slots, prefetch, corner-turn width) variant is its own kernel instance: all must
    give the same bytes and the same exact V."""
    n_tb = 8
    bb = rand_heaps(n_tb, N, C)
    ore, oim = O.v_full_from_heaps(bb, n_tb, N, C, 32)
    are, aim, _, _ = O.unpack(bb, n_tb, N, C, 32)
    exp_t = np.stack([are.reshape(C, N, n_tb, 32).transpose(0, 2, 1, 3),
                      aim.reshape(C, N, n_tb, 32).transpose(0, 2, 1, 3)], axis=2)
    d_bb = M.DeviceBuffer.from_array(bb)
    try:
        for slots in (1, 2, 8):
            for pf in (1, 2, 3, 4, 5):
                for ct in (8, 16, 32):
                    if ct != 32 and (slots, pf) != (2, 2):
                        continue
                    M.set_tuning(N, slots, pf, ct)
                    cfg = cfg_for(N, C, 32 * n_tb)
                    assert M.acc_slots(cfg) == slots
                    d_t = M.DeviceBuffer(M.tiles_bytes(cfg, n_tb))
                    d_acc = M.DeviceBuffer(M.acc_bytes(cfg))
                    M.corner_turn(cfg, d_bb, n_tb, d_t)
                    assert np.array_equal(d_t.download(np.int8, exp_t.shape), exp_t), (slots, pf, ct)
                    M.xcorr(cfg, d_t, n_tb, d_acc)
                    vre, vim = M.unpack_acc(d_acc.download(np.int32, (M.acc_bytes(cfg) // 4,)), cfg)
                    assert np.array_equal(vre, ore) and np.array_equal(vim, oim), (slots, pf, ct)
    finally:
        M.lib().mixer_set_tuning(N, M.C.byref(M.Tuning(0, 0, 0, 0, 0)))


def test_autotune_sets_a_valid_plan():
    cfg = cfg_for(128, 16, 1024)
    try:
        best = M.autotune(cfg)
        log(f"autotune 128 x 16 x 1024 -> {best.as_dict()}")
        assert best.slots in (1, 2, 4, 8) and best.ct_chan in (8, 16, 32) and best.staging in (1, 2)
        assert (best.mfma, best.prefetch) in [(32, 1), (32, 2), (32, 3), (32, 4), (32, 5), (16, 0)]
        t = M.get_tuning(cfg)
        assert (t.slots, t.ct_chan, t.mfma) == (best.slots, best.ct_chan, best.mfma)
        assert best.mfma == 16 or t.prefetch == best.prefetch
        bb = rand_heaps(32, 128, 16)
        c2, _, acc = gpu_v(cfg, bb, 32)
        vre, vim = M.unpack_acc(acc, c2)
        ore, oim = O.v_full_from_heaps(bb, 32, 128, 16, 32)
        assert np.array_equal(vre, ore) and np.array_equal(vim, oim)
    finally:
        M.lib().mixer_set_tuning(128, M.C.byref(M.Tuning(0, 0, 0, 0, 0)))


def test_autotune_blocked_then_context_is_exact():
    """mixer_autotune on a blocked size (512 elements x 64 channels, one launch per integration,
    enough workgroups for one time slot) also decides between the integrate in the correlator's
    epilogue and the separate kernel when a one-slot plan wins; a context created afterwards runs
    the chosen plan and its V and images match the oracle (channels 0, 31 and 63)."""
    N, C, n_tb = 512, 64, 32
    cfg = cfg_for(N, C, 32 * n_tb)
    try:
        best = M.autotune(cfg)
        log(f"autotune {N} x {C} x {32 * n_tb} -> {best.as_dict()}")
        assert best.mfma == 32
        # the choice exists only for one-slot plans staged from the heaps two blocks ahead
        one_launch = best.slots == 1 and best.staging == 1 and best.prefetch == 3
        assert best.integrate in ((1, 2) if one_launch else (0,))
        bb = rand_heaps(n_tb, N, C)
        d = M.DeviceBuffer.from_array(bb)
        with M.Engine(cfg) as e:
            e.submit_device(d, n_tb, token=False)
            e.sync()
            v, idx = e.dump()
            img, _, _, _ = e.read_image()
        assert idx == 1
        for c in (0, 31, 63):
            ore, oim, of = O.xengine(bb, n_tb, N, C, 32, c, c + 1, nthreads=min(16, os.cpu_count() or 1))
            assert of == 0
            assert np.array_equal(v[c], np.stack([ore[0], oim[0]]).astype(np.float32)), f"channel {c}"
            oimg, _, _, _, _ = O.image_from_v(ore.astype(np.int64), oim.astype(np.int64), cfg.grid_x, cfg.ng)
            err = np.abs(img[c].astype(np.float64) - oimg[0]).max() / np.abs(oimg[0]).max()
            assert err <= 1e-5, f"channel {c} image error {err:.2e}"
    finally:
        M.set_tuning(N)


# ---------------------------------------------------------------- fused corner-turn + correlate
@pytest.mark.parametrize("N,C,n_tb,parts", [
    (256, 5, 8, 1), (256, 3, 12, 3), (128, 6, 4, 1), (128, 9, 6, 2), (32, 3, 3, 1), (64, 9, 4, 1),
    (96, 2, 2, 1), (160, 3, 5, 1),
    (224, 2, 3, 1), (384, 2, 2, 1), (512, 3, 3, 1), (1024, 1, 2, 1), (640, 1, 4, 2),
    # odd block counts per time slot (7; 3 and 4) through the two-ahead pipelines' last pass
    (256, 6, 7, 1), (512, 2, 7, 1),
    # the three-ahead pipeline's six-block rounds: whole rounds, and rounds plus 1..5 blocks
    (1024, 1, 13, 1), (512, 2, 6, 1), (256, 2, 11, 1), (384, 1, 12, 1), (640, 1, 9, 1),
])
def test_xcorr_heaps_fused_bit_exact(N, C, n_tb, parts):
    """The correlator staging straight from the heaps gives the CPU correlator's V, for every
    pipeline variant and time split."""
    log(f"xcorr_heaps N={N} C={C} n_tb={n_tb} parts={parts}")
    bb = rand_heaps(n_tb, N, C)
    ore, oim = O.v_full_from_heaps(bb, n_tb, N, C, 32)
    per = n_tb // parts
    try:
        for slots in (1, 2, 8):
            for pf in (1, 2, 3, 4, 5):
                if slots > per or (slots == 8 and pf not in (1, 3, 5)):
                    continue
                M.set_tuning(N, slots, pf)
                cfg = cfg_for(N, C, 32 * per)
                d_acc = M.DeviceBuffer(M.acc_bytes(cfg))
                for p in range(parts):
                    d_bb = M.DeviceBuffer.from_array(np.ascontiguousarray(bb[p * per:(p + 1) * per]))
                    M.xcorr_heaps(cfg, d_bb, per, d_acc, accumulate=p > 0)
                vre, vim = M.unpack_acc(d_acc.download(np.int32, (M.acc_bytes(cfg) // 4,)), cfg)
                assert np.array_equal(vre, ore) and np.array_equal(vim, oim), (slots, pf)
    finally:
        M.set_tuning(N)


# ---------------------------------------------------------------- the 16x16x64 tile shape
@pytest.mark.parametrize("N,C,n_tb,parts", [
    (256, 5, 8, 1), (256, 3, 7, 1), (256, 6, 12, 3), (224, 2, 5, 1), (192, 3, 6, 2), (160, 3, 5, 1),
    (128, 9, 6, 2), (96, 2, 3, 1), (64, 9, 4, 1), (32, 3, 3, 1), (32, 1, 1, 1),
])
def test_mfma16_shape_bit_exact(N, C, n_tb, parts):
    """The FOLD correlator in the v_mfma_i32_16x16x64_i8 shape (two 32-sample blocks per
    instruction, an odd last block paired with zeros) gives the CPU correlator's V straight from
    the heaps and from corner-turned tiles, for every time split and accumulated V-updates."""
    log(f"mfma16 N={N} C={C} n_tb={n_tb} parts={parts}")
    bb = rand_heaps(n_tb, N, C)
    ore, oim = O.v_full_from_heaps(bb, n_tb, N, C, 32)
    per = n_tb // parts
    try:
        for slots in (1, 2, 3):
            if slots > per:
                continue
            M.set_tuning(N, slots=slots, mfma=16)
            cfg = cfg_for(N, C, 32 * per)
            assert M.get_tuning(cfg).mfma == 16 and M.acc_slots(cfg) == slots
            d_acc = M.DeviceBuffer(M.acc_bytes(cfg))
            d_acc2 = M.DeviceBuffer(M.acc_bytes(cfg))
            d_t = M.DeviceBuffer(M.tiles_bytes(cfg, per))
            for p in range(parts):
                d_bb = M.DeviceBuffer.from_array(np.ascontiguousarray(bb[p * per:(p + 1) * per]))
                M.xcorr_heaps(cfg, d_bb, per, d_acc, accumulate=p > 0)
                M.corner_turn(cfg, d_bb, per, d_t)
                M.xcorr(cfg, d_t, per, d_acc2, accumulate=p > 0)
            for acc in (d_acc, d_acc2):
                vre, vim = M.unpack_acc(acc.download(np.int32, (M.acc_bytes(cfg) // 4,)), cfg)
                assert np.array_equal(vre, ore) and np.array_equal(vim, oim), slots
    finally:
        M.set_tuning(N)


def test_mfma16_shape_counts_minus128_and_runs_the_engine():
    """The 16x16x64 shape counts imaginary -128 like the other variants, and a context on it
    dumps the exact V and its image."""
    try:
        M.set_tuning(256, mfma=16)
        _check_minus128(256, 4)
        cfg = cfg_for(256, 8, 32 * 6)
        bb = rand_heaps(6, 256, 8)
        d = M.DeviceBuffer.from_array(bb)
        with M.Engine(cfg) as e:
            e.token_wait(e.submit_device(d, 6))
            e.sync()
            v, idx = e.dump()
            img, _, _, _ = e.read_image()
        r, i = O.v_full_from_heaps(bb, 6, 256, 8, 32)
        assert idx == 1 and np.array_equal(v, np.stack([r, i], 1).astype(np.float32))
        o = O.image_from_v(r, i, cfg.grid_x, cfg.ng)[0]
        assert np.abs(img - o).max() / np.abs(o).max() <= 1e-5
    finally:
        M.set_tuning(256)


@pytest.mark.parametrize("N,C,pf", [(256, 4, 1), (256, 4, 3), (256, 4, 4), (96, 3, 2), (160, 2, 3), (512, 2, 1),
                                    (512, 2, 2), (512, 2, 3), (256, 4, 5), (512, 2, 5)])
def test_xcorr_heaps_counts_imaginary_minus128(N, C, pf):
    try:
        M.set_tuning(N, prefetch=pf)
        _check_minus128(N, C)
    finally:
        M.set_tuning(N)


def _check_minus128(N, C):
    n_tb = 4
    cfg = cfg_for(N, C, 32 * n_tb)
    bb = rand_heaps(n_tb, N, C, -127, 127)
    d_acc = M.DeviceBuffer(M.acc_bytes(cfg))
    d_w = M.DeviceBuffer(4)
    d_w.zero()
    d_bb = M.DeviceBuffer.from_array(bb)
    M.xcorr_heaps(cfg, d_bb, n_tb, d_acc, d_wrap=d_w)
    assert d_w.download(np.uint32, (1,))[0] == 0
    # a real -128 is harmless; an imaginary one (element 37, channel 1, sample 5) is counted
    bb2 = bb.copy()
    bb2[1, 37, 1 * 32 + 5] = 0x80
    d_bb.upload(bb2)
    M.xcorr_heaps(cfg, d_bb, n_tb, d_acc, d_wrap=d_w)
    assert d_w.download(np.uint32, (1,))[0] == 0
    bb2[1, 37, C * 32 + 1 * 32 + 5] = 0x80
    d_bb.upload(bb2)
    M.xcorr_heaps(cfg, d_bb, n_tb, d_acc, d_wrap=d_w)
    assert d_w.download(np.uint32, (1,))[0] >= 1
    # the lower triangle stays exact with the -128 present
    nb = N // 32
    a = d_acc.download(np.int32, (M.acc_bytes(cfg) // 4,)).reshape(-1, C, nb * (nb + 1) // 2, 2, 32, 32)
    a = a.astype(np.int64).sum(0)
    ore, oim = O.v_full_from_heaps(bb2, n_tb, N, C, 32)
    t = 0
    for I in range(nb):
        for J in range(I + 1):
            assert np.array_equal(a[:, t, 0], ore[:, I * 32:(I + 1) * 32, J * 32:(J + 1) * 32])
            assert np.array_equal(a[:, t, 1], oim[:, I * 32:(I + 1) * 32, J * 32:(J + 1) * 32])
            t += 1


@pytest.mark.parametrize("N,C,k_update,staging,check", [
    (256, 8, 256, 1, True), (256, 8, 256, 2, True), (256, 8, 256, 1, False), (256, 8, 256, 2, False),
    (512, 2, 512, 1, True), (512, 2, 512, 1, False)])
def test_engine_staging_modes_agree(N, C, k_update, staging, check):
    """Contexts correlate from the heaps (staging 1; at 512 elements with one launch per
    integration, with the integrate in the correlator's epilogue) or through a separate
    corner-turn pass (staging 2); all give the exact lower triangle with an imaginary -128 present
    and count it, unless the context was created with MIXER_CFG_NO_WRAP_CHECK (wrap128 stays 0)."""
    cfg = cfg_for(N, C, k_update)
    cfg.k_integ = 512
    cfg.flags = 0 if check else M.CFG_NO_WRAP_CHECK
    n_tb = 16
    bb = rand_heaps(n_tb, N, C)
    bb[3, 100, C * 32 + 7] = 0x80
    ore, oim = O.v_full_from_heaps(bb, n_tb, N, C, 32)
    d = M.DeviceBuffer.from_array(bb)
    try:
        M.set_tuning(N, staging=staging)
        assert M.get_tuning(cfg).staging == staging
        with M.Engine(cfg) as e:
            e.token_wait(e.submit_device(d, n_tb))
            e.sync()
            v, idx = e.dump()
            st = e.stats()
        assert idx == 1 and (st.wrap128 >= 1 if check else st.wrap128 == 0), st.wrap128
        nb = N // 32
        for I in range(nb):  # lower triangle (exact even with the -128)
            for J in range(I + 1):
                sl = (slice(None), slice(I * 32, (I + 1) * 32), slice(J * 32, (J + 1) * 32))
                assert np.array_equal(v[:, 0][sl], ore[sl].astype(np.float32))
                assert np.array_equal(v[:, 1][sl], oim[sl].astype(np.float32))
    finally:
        M.set_tuning(N)
