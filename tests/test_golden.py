"""Golden fixtures (tests/golden/xengine_small.npz, written by make_golden.py with numpy only)
against the oracle (CPU) and the HIP path (GPU); the reference's own peek32 viewer, built
from its source into oracle/_ref, reads a Corr32 image cube laid out by libmixer (CPU)."""
import os
import re
import subprocess

import numpy as np
import pytest

import mixer_amd as M
import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = np.load(os.path.join(ROOT, "tests", "golden", "xengine_small.npz"))
N, GX, C, NTB, T, NG = (int(v) for v in G["geometry"])
PEEK = os.path.join(ROOT, "oracle", "_ref", "peek32")


def test_oracle_matches_golden():
    heaps = np.ascontiguousarray(G["heaps"])
    vre, vim = O.v_full_from_heaps(heaps, NTB, N, C, T)
    assert np.array_equal(vre, G["vre"]) and np.array_equal(vim, G["vim"])
    g = O.grid(np.ascontiguousarray(G["v_f32"]), GX, NG)
    assert np.allclose(g, G["grid"], rtol=0, atol=1e-9 * np.abs(G["grid"]).max())
    f = O.dft2(g)
    img, vmax, vmin = O.publish(f)
    assert np.allclose(img, G["img"], rtol=0, atol=1e-6 * np.abs(G["img"]).max())
    assert np.allclose(vmax, G["vmax"], rtol=1e-6) and np.allclose(vmin, G["vmin"], rtol=1e-6)


def corr32_region(img, vmax, vmin, write_seq=1):
    cfg = M.default_cfg(n_elem=N, grid_x=GX, n_chan=C, ng=NG, k_integ=64, k_update=64)
    region = np.zeros(M.corr32_bytes(cfg), np.uint8)
    M.check(M.lib().mixer_corr32_init(M.ptr(region), M.C.byref(cfg)))
    region[16:24] = np.frombuffer(np.uint64(write_seq).tobytes(), np.uint8)
    body = np.concatenate([img.ravel(), vmax, vmin]).astype(np.float32)
    region[24:] = body.view(np.uint8)
    return region


@pytest.mark.skipif(not os.path.exists(PEEK), reason="reference peek32 not built (oracle/_ref)")
def test_reference_peek32_reads_corr32_layout():
    img, vmax, vmin = G["img"], G["vmax"], G["vmin"]
    region = corr32_region(img, vmax, vmin, write_seq=7)
    assert region.nbytes == 24 + 4 * (C * NG * NG + 2 * C)
    shm = "/dev/shm/corr32"
    if os.path.exists(shm):
        pytest.skip("/dev/shm/corr32 in use")
    try:
        region.tofile(shm)
        out = subprocess.run([PEEK, "1"], capture_output=True, text=True, timeout=30).stdout
    finally:
        os.unlink(shm)
    # header line: seq, channel, vmax/vmin printed by the reference viewer
    m = re.search(r"seq=(\d+)\s+channel (\d+)/(\d+)\s+vmax=(\S+) vmin=(\S+)", out)
    assert m, out
    assert (int(m.group(1)), int(m.group(2)), int(m.group(3))) == (7, 1, C)
    assert float(m.group(4)) == pytest.approx(float(vmax[1]), rel=1e-2)
    assert float(m.group(5)) == pytest.approx(float(vmin[1]), rel=1e-2)
    y, x = np.unravel_index(np.argmax(img[1]), (NG, NG))
    m = re.search(r"star peak \(x=(\d+),y=(\d+)\) val=(\S+)", out)
    assert m and (int(m.group(1)), int(m.group(2))) == (x, y)
    assert float(m.group(3)) == pytest.approx(float(img[1].max()), rel=1e-2)
    assert len([ln for ln in out.splitlines() if len(ln) == 2 * NG]) == NG  # the 32-row ASCII map


def test_peek_viewer_reads_corr32_and_respects_writer_flag():
    """The package viewer parses the same region: auto channel pick (brightest off-centre pixel,
    src/peek32.cpp:30-39), the same ASCII map as the reference viewer where that is built, and
    it never returns a frame while the writer flag (bit 63 of write_seq) is set."""
    from mixer_amd import peek
    img, vmax, vmin = G["img"], G["vmax"], G["vmin"]
    region = corr32_region(img, vmax, vmin, write_seq=7)
    ws, im, mx, mn = peek.read_frame(region.tobytes())
    assert ws == 7 and np.array_equal(im, img.astype(np.float32)) and np.array_equal(mx, vmax.astype(np.float32))
    yy, xx = np.mgrid[0:NG, 0:NG]
    off = (xx - NG // 2) ** 2 + (yy - NG // 2) ** 2 > 4
    brightest = [float(np.where(off, img[c], -np.inf).max()) for c in range(C)]
    assert peek.pick_channel(im) == int(np.argmax(brightest))
    text = peek.render(ws, im, mx, mn, ch=1)
    assert len([ln for ln in text.splitlines() if len(ln) == 2 * NG]) == NG
    if os.path.exists(PEEK) and not os.path.exists("/dev/shm/corr32"):
        try:
            region.tofile("/dev/shm/corr32")
            ref = subprocess.run([PEEK, "1"], capture_output=True, text=True, timeout=30).stdout
        finally:
            os.unlink("/dev/shm/corr32")
        rows = lambda t: [ln for ln in t.splitlines() if len(ln) == 2 * NG]
        assert rows(ref) == rows(text)
    busy = corr32_region(img, vmax, vmin, write_seq=(1 << 63) | 8)
    with pytest.raises(TimeoutError):
        peek.read_frame(busy.tobytes(), retries=3)
    assert peek.read_frame(corr32_region(img, vmax, vmin, write_seq=0).tobytes()) is None


@pytest.mark.gpu
def test_hip_path_matches_golden():
    M.check(M.lib().mixer_set_device(0))
    cfg = M.default_cfg(n_elem=N, grid_x=GX, n_chan=C, ng=NG, k_integ=NTB * T, k_update=NTB * T,
                        pkt_pitch=1024)
    d = M.DeviceBuffer.from_array(np.ascontiguousarray(G["heaps"]))
    with M.Engine(cfg) as e:
        e.token_wait(e.submit_device(d, NTB))
        e.sync()
        v, idx = e.dump()
        img, vmax, vmin, _ = e.read_image()
        region = np.zeros(M.corr32_bytes(cfg), np.uint8)
        e.publish(region)
    assert idx == 1 and np.array_equal(v, G["v_f32"])
    err = np.abs(img - G["img"]).max() / np.abs(G["img"]).max()
    print(f"golden image max err/peak {err:.2e}")
    assert err <= 1e-5
    assert np.array_equal(region[:24].view(np.uint32)[:4], [0x43523343, NG, NG, C])
    assert region[16:24].view(np.uint64)[0] == 1
    assert np.array_equal(region[24:24 + img.nbytes].view(np.float32).reshape(img.shape), img)
