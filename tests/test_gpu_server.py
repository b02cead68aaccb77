"""GPU tests of the serving loop through the C ABI: admission by HBM budget, chunks correlated
as they arrive with device-resident accumulators, V bit-identical to the oracle's exact sums,
images within 1e-5 of the FP64 oracle, one latency sample per dumped integration."""
import sys

import numpy as np
import pytest

import mixer_amd as M
import pyoracle as O

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(0x4D495836)


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cfg_for(N, C, k_update, k_integ):
    gx = 16 if N % 16 == 0 and N // 16 <= 16 else 32
    gy = N // gx
    ng = max(32, 1 << int(np.ceil(np.log2(2 * max(gx, gy)))))
    return M.default_cfg(n_elem=N, grid_x=gx, n_chan=C, k_update=k_update, k_integ=k_integ, ng=ng,
                         pkt_pitch=max(16384, 64 + 64 * C))


def heaps(N, C, n_tb):
    return RNG.integers(-127, 128, size=(n_tb, N, 64 * C), dtype=np.int16).astype(np.int8).view(np.uint8)


@pytest.fixture(scope="module", autouse=True)
def device():
    assert M.device_count() >= 1, "no gfx950 device visible to libmixer"
    M.check(M.lib().mixer_set_device(0))
    yield


def test_server_admission_exact_dumps_and_latency():
    ca = cfg_for(256, 16, 256, 1024)   # 8 time blocks per chunk, 4 chunks per integration
    cb = cfg_for(512, 4, 128, 256)     # BLOCKED geometry, 2 chunks per integration
    cc = cfg_for(64, 8, 64, 192)
    fa, fb, fc = (M.server_footprint(c) for c in (ca, cb, cc))
    assert min(fa, fb, fc) > 0
    with M.Server(fa + fb + fc // 2) as S:
        a = S.admit(ca)
        b = S.admit(cb)
        assert S.try_admit(cc) is None  # does not fit the remaining budget
        st = S.stats()
        assert st.rejected == 1 and st.admitted == 2 and st.hbm_used == fa + fb
        # array a: two integrations in uneven submissions (1 + 3 + 2 + 2 chunks)
        na = 2 * 1024 // 32
        ha = heaps(256, 16, na)
        da = M.DeviceBuffer.from_array(ha)
        tb_a = 256 * 64 * 16
        off = 0
        for n in (8, 24, 16, 16):
            S.submit(a, da.ptr.value + off * tb_a, n, arrival_ns=0)
            off += n
        # array b: three integrations, one chunk per submission
        nb_ = 3 * 256 // 32
        hb = heaps(512, 4, nb_)
        db = M.DeviceBuffer.from_array(hb)
        tb_b = 512 * 64 * 4
        for k in range(0, nb_, 4):
            S.submit(b, db.ptr.value + k * tb_b, 4, arrival_ns=0)
        S.drain()
        st = S.stats()
        assert st.integrations == 2 + 3 and st.lat_n == 5 and st.chunks == 4 + 6
        va, ia = S.dump(a)
        vb, ib = S.dump(b)
        assert ia == 2 and ib == 3
        r, i = O.v_full_from_heaps(np.ascontiguousarray(ha[32:]), 32, 256, 16, 32)
        assert np.array_equal(va, np.stack([r, i], 1).astype(np.float32))
        img, vmax, vmin, idx = S.read_image(a)
        oimg, _, _, _, _ = O.image_from_v(r, i, ca.grid_x, ca.ng)
        scale = np.abs(oimg).reshape(16, -1).max(1)
        assert idx == 2 and (np.abs(img - oimg).reshape(16, -1).max(1) / scale).max() <= 1e-5
        assert np.array_equal(vmax, img.reshape(16, -1).max(1))
        r, i = O.v_full_from_heaps(np.ascontiguousarray(hb[16:]), 8, 512, 4, 32)
        assert np.array_equal(vb, np.stack([r, i], 1).astype(np.float32))
        # releasing b frees its budget for c
        S.release(b)
        c = S.admit(cc)
        assert S.stats().hbm_used == fa + fc
        hc = heaps(64, 8, 6)
        dc = M.DeviceBuffer.from_array(hc)
        S.token_wait(c, S.submit(c, dc, 6, arrival_ns=0))
        S.drain()
        vc, icx = S.dump(c)
        r, i = O.v_full_from_heaps(hc, 6, 64, 8, 32)
        assert icx == 1 and np.array_equal(vc, np.stack([r, i], 1).astype(np.float32))


def test_serve_run_poisson_arrivals():
    """Short run of the load generator: every integration that completes is dumped and timed,
    the offered rate is delivered, and the last dump is the exact product of the replayed ring."""
    # array 0 integrates one chunk at a time, so its dump count names the last ring slot used
    cfgs = [cfg_for(256, 32, 1024, 1024), cfg_for(512, 16, 512, 1024), cfg_for(128, 64, 256, 1024)]
    with M.Server(64 << 30) as S:
        ids = [S.admit(c) for c in cfgs]
        rep = S.run(duration_s=0.6, aggregate_gbps=20.0)
        log(f"serve: {rep.as_dict()}")
        assert rep.chunks > 0 and rep.integrations > 0
        assert rep.lat_p50_ms > 0 and rep.lat_p99_ms >= rep.lat_p50_ms and rep.lat_p99_ms < 50.0
        assert 0.5 * 20.0 < rep.achieved_gbps < 1.5 * 20.0
        st = S.stats()
        assert st.integrations == st.lat_n
        p, n_chunks, cbytes = S.input(ids[0])
        c = cfgs[0]
        n_tb = c.k_integ // 32
        assert n_chunks == 2 and cbytes == n_tb * c.n_elem * c.payload
        v, idx = S.dump(ids[0])
        assert idx >= 1
        slot = np.empty((n_tb, c.n_elem, c.payload), np.uint8)
        M.check(M.lib().mixer_memcpy(M.ptr(slot), M.C.c_void_p(p + ((idx - 1) % 2) * cbytes), slot.nbytes,
                                     M.D2H, None))
        assert np.abs(slot.view(np.int8).astype(np.int16)).max() <= 127
        r, i = O.v_full_from_heaps(slot, n_tb, c.n_elem, c.n_chan, 32)
        assert np.array_equal(v, np.stack([r, i], 1).astype(np.float32))


@pytest.mark.parametrize("coalesce,staging", [(True, 0), (False, 0), (True, 2)])
def test_server_coalesced_chunks_exact(coalesce, staging):
    """Chunks submitted while their array's previous launch still runs (here: queued behind a
    long launch of a large array on the same stream) are held and correlated in fewer, longer
    launches; the dumps stay bit-identical to the oracle. With coalescing off every update is
    its own launch. staging 2 (separate corner-turn) launches a held run one update at a time
    through the tile buffer."""
    big = cfg_for(1024, 32, 8192, 8192)
    small = cfg_for(512, 8, 128, 1024)  # 4 time blocks per chunk, 8 chunks per integration
    M.set_tuning(512, staging=staging)
    try:
        with M.Server(64 << 30, n_streams=1) as S:
            S.set_coalescing(coalesce)
            hb = S.admit(big)
            s = S.admit(small)
            p, _, _ = S.input(hb)
            n_upd = 3 * 8
            hs = heaps(512, 8, 4 * n_upd)
            ds = M.DeviceBuffer.from_array(hs)
            tb = 512 * 64 * 8
            for _ in range(4):
                S.submit(hb, p, 256, arrival_ns=0)
            toks = [S.submit(s, ds.ptr.value + k * 4 * tb, 4, arrival_ns=0) for k in range(n_upd)]
            S.token_wait(s, toks[-1])
            S.drain()
            st = S.stats()
            log(f"coalesce={coalesce}: {st.launches} launches for {4 + n_upd} updates")
            assert st.integrations == 4 + 3
            if coalesce and staging != 2:
                assert st.launches < 4 + n_upd
            else:
                assert st.launches == 4 + n_upd
            v, idx = S.dump(s)
            assert idx == 3
            r, i = O.v_full_from_heaps(np.ascontiguousarray(hs[64:]), 32, 512, 8, 32)
            assert np.array_equal(v, np.stack([r, i], 1).astype(np.float32))
    finally:
        M.set_tuning(512)


def test_server_holds_at_most_half_the_token_ring():
    """An integration of 128 one-block chunks queued behind a long launch: the held run is
    launched every 32 tokens (a token older than the 64-event ring would count as complete), and
    the dump is exact."""
    big = cfg_for(1024, 32, 8192, 8192)
    small = cfg_for(512, 4, 32, 4096)  # 1 time block per chunk, 128 chunks per integration
    with M.Server(64 << 30, n_streams=1) as S:
        hb = S.admit(big)
        s = S.admit(small)
        p, _, _ = S.input(hb)
        hs = heaps(512, 4, 128)
        ds = M.DeviceBuffer.from_array(hs)
        tb = 512 * 64 * 4
        for _ in range(4):
            S.submit(hb, p, 256, arrival_ns=0)
        toks = [S.submit(s, ds.ptr.value + k * tb, 1, arrival_ns=0) for k in range(128)]
        S.token_wait(s, toks[0])
        S.token_wait(s, toks[40])
        S.drain()
        st = S.stats()
        log(f"128 chunks: {st.launches - 4} launches")
        assert 128 // 32 <= st.launches - 4 <= 128
        v, idx = S.dump(s)
        assert idx == 1
        r, i = O.v_full_from_heaps(hs, 128, 512, 4, 32)
        assert np.array_equal(v, np.stack([r, i], 1).astype(np.float32))
