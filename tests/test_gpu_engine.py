"""GPU tests of the engine's ingest and monitor paths through the C ABI, against the oracle:

- bounded ingest: time blocks that find no free slot are dropped and counted in imissed, as the
  reference's drain thread does (src/rx_host_corr.cu:139-140), and the integration built from
  the blocks that were accepted is bit-exact;
- completion tokens: a token reported done is complete even after its event slot was reused
  (more than 256 tokens outstanding), so a caller that refills its buffer as soon as the token
  says so never corrupts V;
- monitor frames: the image of everything integrated since the previous frame (the reference's
  dV, src/rx_host_corr.cu:218-222) against the FP64 image of the summed V.
"""
import sys

import numpy as np
import pytest

import mixer_amd as M
import pyoracle as O

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(0x4D49583A)


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cfg_for(N, C, k_update, k_integ=None, gx=8, ng=32, **kw):
    return M.default_cfg(n_elem=N, grid_x=gx, n_chan=C, k_update=k_update,
                         k_integ=k_integ or k_update, ng=ng, pkt_pitch=max(1024, 64 + 64 * C), **kw)


def heaps_of(pk, N, C):
    return np.ascontiguousarray(pk[:, 64:64 + 64 * C]).reshape(-1, N, 64 * C)


def pinned(nbytes):
    p = M.C.c_void_p()
    M.check(M.lib().mixer_host_alloc(nbytes, M.C.byref(p)))
    return p, np.ctypeslib.as_array((M.C.c_uint8 * nbytes).from_address(p.value))


@pytest.fixture(scope="module", autouse=True)
def device():
    assert M.device_count() >= 1, "no gfx950 device visible to libmixer"
    M.check(M.lib().mixer_set_device(0))
    yield


def exact(v, bb, n_tb, N, C):
    r, i = O.v_full_from_heaps(np.ascontiguousarray(bb), n_tb, N, C, 32)
    return np.array_equal(v, np.stack([r, i], 1).astype(np.float32))


def test_ingest_slot_exhaustion_drops_and_counts():
    """5 slots: one call of 12 time blocks keeps the first 5 and drops 7 (imissed = 7); after the
    copies complete the next call gets its slots back. The integration of the 8 accepted blocks
    (0-4 and 12-14) is bit-exact; a 0-slot (unbounded) context drops nothing."""
    N, C, n_tb = 64, 8, 8
    cfg = cfg_for(N, C, 32 * n_tb)
    pitch = int(cfg.pkt_pitch)
    pk = M.synth_packets(cfg, M.default_sky(sigma=0.5), seed=21, seq0=0, npkts=15 * N, pitch=pitch)
    hp, ring = pinned(15 * N * pitch)
    try:
        ring[:] = pk.reshape(-1)
        host = ring.reshape(15 * N, pitch)
        with M.Engine(cfg) as e:
            e.set_ingest_slots(5)
            log("slots=5: submit 12 blocks in one call")
            e.token_wait(e.submit_heaps(host[:12 * N], pitch, 12 * N))
            st = e.stats()
            assert st.imissed == 7 and st.heaps_in == 5 * N, (st.imissed, st.heaps_in)
            e.sync()
            e.token_wait(e.submit_heaps(host[12 * N:], pitch, 3 * N))
            e.sync()
            st = e.stats()
            v, idx = e.dump()
        assert st.imissed == 7 and st.heaps_in == 8 * N and st.integrations == 1 and idx == 1
        bb = heaps_of(pk, N, C)
        assert exact(v, bb[[0, 1, 2, 3, 4, 12, 13, 14]], n_tb, N, C)
        with M.Engine(cfg) as e:  # unbounded: nothing dropped
            e.token_wait(e.submit_heaps(host[:8 * N], pitch, 8 * N))
            e.sync()
            st = e.stats()
            v, idx = e.dump()
        assert st.imissed == 0 and idx == 1 and exact(v, bb[:8], n_tb, N, C)
    finally:
        M.check(M.lib().mixer_host_free(hp))


def test_tokens_past_the_ring_are_complete_before_reuse():
    """300 single-block calls from 300 pinned buffers, one token each (more than the 256 token
    slots). After every call, each buffer whose token reports done is overwritten at once with
    noise, as a NIC would refill it. Had a token been reported done before its copy ran, V would
    hold the noise: it equals the oracle's V of the original data."""
    N, C, n_tb = 64, 16, 300
    cfg = cfg_for(N, C, 32 * 60, k_integ=32 * n_tb)
    pitch = int(cfg.pkt_pitch)
    pk = M.synth_packets(cfg, M.default_sky(sigma=0.5), seed=22, seq0=0, npkts=n_tb * N, pitch=pitch)
    hp, ring = pinned(n_tb * N * pitch)
    try:
        host = ring.reshape(n_tb * N, pitch)
        host[:] = pk
        toks, done = [], np.zeros(n_tb, bool)
        refilled = 0
        with M.Engine(cfg) as e:
            for k in range(n_tb):
                toks.append(e.submit_heaps(host[k * N:(k + 1) * N], pitch, N))
                for j in np.flatnonzero(~done[:k + 1]):
                    if e.token_done(toks[j]):
                        done[j] = True
                        host[j * N:(j + 1) * N, 64:] = RNG.integers(0, 256, size=(N, pitch - 64), dtype=np.uint8)
                        refilled += 1
            for j in range(n_tb):
                e.token_wait(toks[j])
            e.sync()
            v, idx = e.dump()
        log(f"{refilled} buffers refilled while later calls were still being submitted")
        assert idx == 1 and toks[-1] - toks[0] == n_tb - 1
        assert exact(v, heaps_of(pk, N, C), n_tb, N, C)
    finally:
        M.check(M.lib().mixer_host_free(hp))


def test_frame_is_the_image_of_everything_since_the_previous_frame():
    """MIXER_CFG_FRAME_DV: after 3 integrations the frame is the image of V1 + V2 + V3; after 2
    more, of V4 + V5; with none in between the frame is empty. Against the FP64 grid + DFT of the
    summed exact V, within 1e-5 of the peak; the latest-integration image is unchanged."""
    N, C, n_tb, ng = 64, 4, 4, 32
    cfg = cfg_for(N, C, 32 * n_tb, flags=M.CFG_FRAME_DV)
    bbs = [RNG.integers(-127, 128, size=(n_tb, N, 64 * C), dtype=np.int16).astype(np.int8).view(np.uint8)
           for _ in range(5)]
    vs = [O.v_full_from_heaps(bb, n_tb, N, C, 32) for bb in bbs]
    d = [M.DeviceBuffer.from_array(bb) for bb in bbs]
    with M.Engine(cfg) as e:
        _, _, _, n0 = e.read_frame()
        assert n0 == 0
        for k in range(3):
            e.submit_device(d[k], n_tb, token=False)
        img, vmax, vmin, n = e.read_frame()
        assert n == 3
        r = sum(v[0].astype(np.int64) for v in vs[:3])
        i = sum(v[1].astype(np.int64) for v in vs[:3])
        o, ovmax, ovmin, _, _ = O.image_from_v(r, i, cfg.grid_x, ng)
        err = np.abs(img - o).max() / np.abs(o).max()
        log(f"frame of 3 integrations: max err / peak {err:.2e}")
        assert err <= 1e-5
        assert np.abs(vmax - ovmax).max() <= 1e-5 * np.abs(o).max()
        assert np.abs(vmin - ovmin).max() <= 1e-5 * np.abs(o).max()
        for k in (3, 4):
            e.submit_device(d[k], n_tb, token=False)
        img, _, _, n = e.read_frame()
        assert n == 2
        r = vs[3][0].astype(np.int64) + vs[4][0]
        i = vs[3][1].astype(np.int64) + vs[4][1]
        o = O.image_from_v(r, i, cfg.grid_x, ng)[0]
        err = np.abs(img - o).max() / np.abs(o).max()
        log(f"frame of 2 integrations: max err / peak {err:.2e}")
        assert err <= 1e-5
        assert e.read_frame()[3] == 0
        lat, _, _, idx = e.read_image()  # the latest integration alone
        o = O.image_from_v(vs[4][0], vs[4][1], cfg.grid_x, ng)[0]
        assert idx == 5 and np.abs(lat - o).max() / np.abs(o).max() <= 1e-5


@pytest.mark.parametrize("N,C,gx,n_tb,flags", [(1024, 16, 32, 8, 0), (512, 64, 32, 6, M.CFG_CUMULATIVE),
                                               (384, 96, 16, 5, 0), (768, 24, 32, 3, M.CFG_CUMULATIVE)])
def test_fused_integrate_matches_oracle_and_separate_kernel(N, C, gx, n_tb, flags):
    """Blocked arrays whose integration is one correlator launch of one time slot write V and the
    image grid from the correlator's epilogue (tuning integrate 2). Over two integrations the
    dumps (per integration and, with CFG_CUMULATIVE, cumulative) and the images equal those of the
    separate integrate kernel (integrate 1) bit for bit, and V equals the oracle's exact sums
    rounded once to f32 (src/rx_host_corr.cu:86-95)."""
    ng = 64
    cfg = cfg_for(N, C, 32 * n_tb, gx=gx, ng=ng, flags=flags)
    pitch = int(cfg.pkt_pitch)
    sky = M.default_sky(sigma=0.5, planet_l=0.3, planet_m=-0.2)
    pk = M.synth_packets(cfg, sky, seed=31 + N, seq0=0, npkts=2 * n_tb * N, pitch=pitch)
    out = {}
    try:
        for integ in (2, 1):
            M.set_tuning(N, slots=1, prefetch=3, integrate=integ)
            t = M.get_tuning(cfg)
            assert t.integrate == integ and t.slots == 1, t.as_dict()
            log(f"fused integrate N={N} C={C} gx={gx} n_tb={n_tb} flags={flags}: integrate={integ}")
            res = []
            with M.Engine(cfg) as e:
                for k in range(2):
                    e.token_wait(e.submit_heaps(pk[k * n_tb * N:(k + 1) * n_tb * N], pitch, n_tb * N))
                    e.sync()
                    v, idx = e.dump()
                    img, vmax, vmin, iidx = e.read_image()
                    cum = e.dump(cumulative=True)[0] if flags & M.CFG_CUMULATIVE else None
                    assert idx == k + 1 and iidx == k + 1
                    res.append((v, img, vmax, vmin, cum))
            out[integ] = res
    finally:
        M.set_tuning(N)
    for k in range(2):
        for a, b in zip(out[2][k], out[1][k]):
            if a is not None:
                assert np.array_equal(a, b)
    bb = heaps_of(pk, N, C)
    for k in range(2):
        assert exact(out[2][k][0], bb[k * n_tb:(k + 1) * n_tb], n_tb, N, C)
    if flags & M.CFG_CUMULATIVE:
        assert np.array_equal(out[2][1][4], out[2][0][0] + out[2][1][0])
