"""mixer_image_noalloc, the image entry point that allocates nothing (its int64 grid workspace is a
static device array of the library), against mixer_image, which allocates that workspace per
call: on two calls in a row, the second with a smaller input (fewer elements and channels), the
images and their max/min agree within 1e-5 of each channel's peak and both match the FP64 grid +
DFT oracle to that bar; a grid beyond the static workspace is refused with MIXER_ERR_ARG."""
import ctypes as C
import sys

import numpy as np
import pytest

import mixer_amd as M
import pyoracle as O

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(0x4D49583B)


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    assert M.device_count() >= 1, "no gfx950 device visible to libmixer"
    M.check(M.lib().mixer_set_device(0))
    yield


def test_noalloc_image_matches_mixer_image_on_two_calls():
    f = M.lib().mixer_image_noalloc
    for N, gx, Cn, ng, n_tb in [(256, 16, 64, 32, 4), (64, 8, 5, 32, 2)]:
        log(f"noalloc image: {N} elements x {Cn} channels, ng {ng}, {n_tb} time blocks")
        cfg = M.default_cfg(n_elem=N, grid_x=gx, n_chan=Cn, k_update=32 * n_tb, k_integ=32 * n_tb, ng=ng,
                            pkt_pitch=max(16384, 64 + 64 * Cn))
        bb = RNG.integers(-127, 128, size=(n_tb, N, 64 * Cn), dtype=np.int16).astype(np.int8).view(np.uint8)
        d = M.DeviceBuffer.from_array(bb)
        acc = M.DeviceBuffer(M.acc_bytes(cfg))
        M.xcorr_heaps(cfg, d, n_tb, acc)
        bufs = [[M.DeviceBuffer(Cn * ng * ng * 4), M.DeviceBuffer(Cn * 4), M.DeviceBuffer(Cn * 4)] for _ in range(2)]
        M.image(cfg, acc, *bufs[0])
        M.image_noalloc(cfg, acc, *bufs[1])
        M.check(M.lib().mixer_device_sync())
        out = [(b[0].download(np.float32, (Cn, ng, ng)), b[1].download(np.float32, Cn), b[2].download(np.float32, Cn))
               for b in bufs]
        r, i = O.v_full_from_heaps(bb, n_tb, N, Cn, 32)
        o = O.image_from_v(r, i, gx, ng)[0]
        peak = np.abs(o).max(axis=(1, 2))
        (ia, xa, na), (ib, xb, nb_) = out
        same = np.array_equal(ia, ib) and np.array_equal(xa, xb) and np.array_equal(na, nb_)
        e_ab = float((np.abs(ia - ib).max(axis=(1, 2)) / peak).max())
        e_b = float((np.abs(ib - o).max(axis=(1, 2)) / peak).max())
        log(f"  mixer_image vs noalloc: identical={same}, max diff / peak {e_ab:.2e}; noalloc vs FP64 {e_b:.2e}")
        assert e_ab <= 1e-5 and e_b <= 1e-5
        assert (np.abs(xa - xb) <= 1e-5 * peak).all() and (np.abs(na - nb_) <= 1e-5 * peak).all()
        for b in bufs:
            for x in b:
                x.free()
        acc.free()
        d.free()
    big = M.default_cfg(n_elem=256, grid_x=16, n_chan=512, k_update=32, k_integ=32, ng=64)
    assert f(None, C.byref(big), C.c_void_p(1), C.c_void_p(1), C.c_void_p(1), C.c_void_p(1)) == M.ERR_ARG
