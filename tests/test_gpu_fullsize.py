"""Full-size GPU parity at every benchmark configuration, through the C ABI, against the oracle.

- cfg2  256 elements x 128 channels x 8192 samples (the headline): every channel's V bit-exact
        against the CPU correlator and every channel's 32 x 32 image within 1e-5 of its peak of
        the FP64 grid + FP64 DFT.
- cfg3  1024 elements x 256 channels x 4096 samples, noise + 3 point sources, 64 x 64 images:
        8 channels bit-exact and their images against FP64; Hermitian, real diagonal and
        Cauchy-Schwarz invariants on all 256 channels.
- cfg4  the 64-array batch (seed 0x4D495834, the benchmark's job list and inputs) in one run:
        one job of every element class (the 1024-element job with the most samples among them)
        bit-exact on 2 channels; invariants on 2 channels of every job from its accumulator.
- cfg5  8 arrays served under Poisson arrivals at 500 GB/s offered: the last dump of every array
        bit-exact on 3 channels against the chunks of its input ring that formed it, Hermitian
        invariants on all its channels.

The integer V follows src/cgemm.cuh:28-69 (bit-exact int32, one f32 rounding), the images
src/rx_host_corr.cu:79-85,205-206,222,229-236.
"""
import os
import sys
import time

import numpy as np
import pytest

import bench
import mixer_amd as M
import pyoracle as O

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    assert M.device_count() >= 1, "no gfx950 device visible to libmixer"
    M.check(M.lib().mixer_set_device(0))
    yield


def oracle_channels(bb, n_tb, N, C, chans):
    out = {}
    for c in chans:
        r, i, of = O.xengine(bb, n_tb, N, C, 32, c, c + 1, nthreads=THREADS)
        assert of == 0
        out[c] = (r[0], i[0])
    return out


def check_invariants(v):
    """V [C][2][N][N] f32 of an integration: Hermitian, real positive diagonal, and
    |V_ab|^2 <= V_aa V_bb (Cauchy-Schwarz) for every pair."""
    assert np.array_equal(v[:, 0], v[:, 0].transpose(0, 2, 1))
    assert np.array_equal(v[:, 1], -v[:, 1].transpose(0, 2, 1))
    d = np.diagonal(v[:, 0], axis1=1, axis2=2).astype(np.float64)
    assert (np.diagonal(v[:, 1], axis1=1, axis2=2) == 0).all()
    assert (d > 0).all()
    for c in range(v.shape[0]):
        p = v[c, 0].astype(np.float64) ** 2 + v[c, 1].astype(np.float64) ** 2
        assert (p <= d[c][:, None] * d[c][None, :] * (1 + 1e-6)).all(), c


def image_err(img, vre, vim, gx, ng):
    o = O.image_from_v(vre, vim, gx, ng)[0]
    return float((np.abs(img - o).max(axis=(1, 2)) / np.abs(o).max(axis=(1, 2))).max())


def run_engine(cfg, bb, n_tb):
    d = M.DeviceBuffer.from_array(bb)
    try:
        with M.Engine(cfg) as e:
            e.token_wait(e.submit_device(d, n_tb))
            e.sync()
            v, idx = e.dump()
            img, vmax, vmin, _ = e.read_image()
    finally:
        d.free()
    assert idx == 1
    return v, img, vmax, vmin


def test_cfg2_every_channel_and_image():
    name = "cfg2"
    cfg = bench.make_cfg(name)
    N, C, K = cfg.n_elem, cfg.n_chan, cfg.k_integ
    n_tb = K // 32
    log(f"{name}: {N} el x {C} ch x {K} samples, seed {bench.SEEDS[name]:#x}")
    bb = bench.synth_heaps(cfg, bench.make_sky(name), bench.SEEDS[name], n_tb)
    v, img, vmax, vmin = run_engine(cfg, bb, n_tb)
    t0 = time.perf_counter()
    r, i, of = O.xengine(bb, n_tb, N, C, 32, 0, C, nthreads=THREADS)
    log(f"{name}: oracle over all {C} channels in {time.perf_counter() - t0:.1f} s")
    assert of == 0
    assert np.array_equal(v[:, 0], r.astype(np.float32)) and np.array_equal(v[:, 1], i.astype(np.float32))
    o, ovmax, ovmin, _, _ = O.image_from_v(r, i, cfg.grid_x, cfg.ng)
    per = np.abs(img - o).max(axis=(1, 2)) / np.abs(o).max(axis=(1, 2))
    log(f"{name}: image max err / channel peak over {C} channels: {per.max():.3e} (median {np.median(per):.3e})")
    assert per.max() <= 1e-5
    peak = np.abs(o).max(axis=(1, 2))
    assert (np.abs(vmax - ovmax) <= 1e-5 * peak).all() and (np.abs(vmin - ovmin) <= 1e-5 * peak).all()


def test_cfg3_channel_subset_images_and_invariants():
    name = "cfg3"
    cfg = bench.make_cfg(name)
    N, C, K = cfg.n_elem, cfg.n_chan, cfg.k_integ
    n_tb = K // 32
    log(f"{name}: {N} el x {C} ch x {K} samples, ng {cfg.ng}, seed {bench.SEEDS[name]:#x}")
    bb = bench.synth_heaps(cfg, bench.make_sky(name), bench.SEEDS[name], n_tb)
    v, img, _, _ = run_engine(cfg, bb, n_tb)
    chans = [0, 1, 63, 64, 127, 128, 200, 255]
    t0 = time.perf_counter()
    ref = oracle_channels(bb, n_tb, N, C, chans)
    log(f"{name}: oracle over {len(chans)} channels in {time.perf_counter() - t0:.1f} s")
    del bb
    errs = []
    for c in chans:
        r, i = ref[c]
        assert np.array_equal(v[c, 0], r.astype(np.float32)) and np.array_equal(v[c, 1], i.astype(np.float32)), c
        errs.append(image_err(img[c:c + 1], r[None], i[None], cfg.grid_x, cfg.ng))
    log(f"{name}: image max err / channel peak over channels {chans}: {max(errs):.3e}")
    assert max(errs) <= 1e-5
    t0 = time.perf_counter()
    for c0 in range(0, C, 32):
        check_invariants(v[c0:c0 + 32])
    log(f"{name}: invariants on all {C} channels in {time.perf_counter() - t0:.1f} s")
    # the brightest source (l = m = 0) at the centre of every channel's image
    assert all(np.unravel_index(np.argmax(img[c]), img[c].shape) == (32, 32) for c in range(C))


def acc_channel(addr, n_slots, C, nb, c):
    """Lower-triangle V (int64, rows >= cols by 32-blocks) of channel c from a tile-packed
    accumulator [n_slots][C][NT][2][32][32]: returns (re, im) [N][N] with the upper blocks 0."""
    NT = nb * (nb + 1) // 2
    tiles = np.zeros((NT, 2, 32, 32), np.int64)
    for s in range(n_slots):
        off = ((s * C + c) * NT) * 2048 * 4
        tiles += M.download_ptr(addr + off, np.int32, (NT, 2, 32, 32))
    N = nb * 32
    re = np.zeros((N, N), np.int64)
    im = np.zeros((N, N), np.int64)
    for I in range(nb):
        for J in range(I + 1):
            t = tiles[I * (I + 1) // 2 + J]
            re[I * 32:I * 32 + 32, J * 32:J * 32 + 32] = t[0]
            im[I * 32:I * 32 + 32, J * 32:J * 32 + 32] = t[1]
    return re, im


def test_cfg4_batch_64_jobs():
    cfgs = bench.batch_jobs()
    payload = [c.k_integ // 32 * c.n_elem * c.payload for c in cfgs]
    log(f"cfg4: {len(cfgs)} jobs, {sum(payload) / 2**30:.1f} GiB per run, seed 0x4D495834")
    bufs = [M.DeviceBuffer(p) for p in payload]
    try:
        for j, d in enumerate(bufs):
            M.fill_random(d, seed=0x4D495834 + j)
        M.check(M.lib().mixer_device_sync())
        # one job per element class; for 1024 elements the one with the most samples
        pick = {}
        for j, c in enumerate(cfgs):
            n = c.n_elem
            if n not in pick or (n == 1024 and c.k_integ > cfgs[pick[n]].k_integ):
                pick[n] = j
        with M.Batch(cfgs) as b:
            b.run(bufs)
            b.sync()
            for n in sorted(pick):
                j = pick[n]
                c = cfgs[j]
                n_tb = c.k_integ // 32
                log(f"cfg4 job {j}: {c.n_elem} el x {c.n_chan} ch x {c.k_integ} samples, exact on 2 channels")
                bb = bufs[j].download(np.uint8, (n_tb, c.n_elem, c.payload))
                v = b.dump(j)
                for ch, (r, i) in oracle_channels(bb, n_tb, c.n_elem, c.n_chan, [0, c.n_chan - 1]).items():
                    assert np.array_equal(v[ch, 0], r.astype(np.float32)), (j, ch)
                    assert np.array_equal(v[ch, 1], i.astype(np.float32)), (j, ch)
                del bb, v
            t0 = time.perf_counter()
            for j, c in enumerate(cfgs):
                addr, n_slots = b.acc(j)
                nb = c.n_elem // 32
                for ch in (0, c.n_chan - 1):
                    re, im = acc_channel(addr, n_slots, c.n_chan, nb, ch)
                    d = np.diagonal(re).astype(np.float64)
                    assert (d > 0).all() and (np.diagonal(im) == 0).all(), (j, ch)
                    for I in range(nb):  # diagonal tiles are Hermitian
                        s = slice(I * 32, I * 32 + 32)
                        assert np.array_equal(re[s, s], re[s, s].T) and np.array_equal(im[s, s], -im[s, s].T), (j, ch, I)
                    lo = np.tril(np.ones((c.n_elem, c.n_elem), bool))
                    p = re.astype(np.float64) ** 2 + im.astype(np.float64) ** 2
                    assert (p[lo] <= (d[:, None] * d[None, :])[lo] * (1 + 1e-9)).all(), (j, ch)
            log(f"cfg4: accumulator invariants on 2 channels of all {len(cfgs)} jobs in {time.perf_counter() - t0:.1f} s")
    finally:
        for d in bufs:
            d.free()


def test_cfg5_serving_last_dumps():
    rate = 500
    cfgs, fs = bench.serve_arrays(rate)
    log(f"cfg5: {[(c.n_elem, c.n_chan) for c in cfgs]}, {cfgs[0].k_integ} samples per integration, "
        f"{rate} GB/s offered, seed 0x4D495835")
    with M.Server(200 << 30) as S:
        ids = [S.admit(c) for c in cfgs]
        rep = S.run(duration_s=2.0, aggregate_gbps=float(rate))
        S.drain()
        log(f"cfg5: achieved {rep.achieved_gbps:.1f} GB/s, {rep.integrations} dumps, p99 {rep.lat_p99_ms:.2f} ms")
        assert rep.integrations >= len(cfgs)
        for a, c in zip(ids, cfgs):
            v, n = S.dump(a)
            assert n >= 1, a
            ring, n_chunks, cb = S.input(a)
            upi = c.k_integ // c.k_update
            tb_per_chunk = c.k_update // 32
            # integration n-1 is the array's chunks (n-1)*upi .. n*upi-1 in submission order, and
            # chunk k was read from ring slot k mod n_chunks
            parts = [M.download_ptr(ring + ((k % n_chunks) * cb), np.uint8, (tb_per_chunk, c.n_elem, c.payload))
                     for k in range((n - 1) * upi, n * upi)]
            bb = np.ascontiguousarray(np.concatenate(parts))
            n_tb = c.k_integ // 32
            chans = [0, c.n_chan // 2, c.n_chan - 1]
            for ch, (r, i) in oracle_channels(bb, n_tb, c.n_elem, c.n_chan, chans).items():
                assert np.array_equal(v[ch, 0], r.astype(np.float32)), (a, ch)
                assert np.array_equal(v[ch, 1], i.astype(np.float32)), (a, ch)
            assert np.array_equal(v[:, 0], v[:, 0].transpose(0, 2, 1))
            assert np.array_equal(v[:, 1], -v[:, 1].transpose(0, 2, 1))
            log(f"cfg5 array {a} ({c.n_elem} x {c.n_chan}): dump {n} exact on channels {chans}")
