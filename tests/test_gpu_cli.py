"""GPU test of the command-line X-engine (bin/mixer_corr, the loop of src/rx_host_corr.cu:178-239):
the built-in synthetic transmitter drives the engine for a few seconds; the Corr32 region in
/dev/shm is read while the program runs (star at the image centre in every channel), and the
per-integration --dump records are Hermitian V cubes of the right size."""
import mmap
import os
import subprocess
import sys
import time

import numpy as np
import pytest

from mixer_amd import peek

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "daqiri-mixer_amd", "bin", "mixer_corr")


def read_shm(name):
    try:
        fd = os.open("/dev/shm/" + name.lstrip("/"), os.O_RDONLY)
    except FileNotFoundError:
        return None
    try:
        if os.fstat(fd).st_size == 0:
            return None
        buf = mmap.mmap(fd, 0, mmap.MAP_SHARED, mmap.PROT_READ)
    finally:
        os.close(fd)
    try:
        return peek.read_frame(buf)
    finally:
        buf.close()


def test_cli_publishes_images_and_dumps_hermitian_cubes(tmp_path):
    N, C = 64, 16
    shm = f"/mixer_cli_test_{os.getpid()}"
    dump = tmp_path / "v.bin"
    cmd = [BIN, "--seconds", "3", "--elements", str(N), "--grid-x", "8", "--channels", str(C),
           "--k-update", "1024", "--ng", "32", "--rate", "0.5", "--fps", "10", "--shm", shm,
           "--dump", str(dump), "--dump-mode", "integ"]
    print(" ".join(cmd), flush=True)
    p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    frame = None
    try:
        deadline = time.time() + 60
        while frame is None and p.poll() is None and time.time() < deadline:
            frame = read_shm(shm)
            time.sleep(0.1)
        out, err = p.communicate(timeout=60)
    finally:
        if p.poll() is None:
            p.kill()
            p.communicate()
    sys.stdout.write(err[-1500:])
    assert p.returncode == 0, err[-2000:]
    assert frame is not None, "no image frame published while the program ran"
    ws, img, vmax, vmin = frame
    assert ws >= 1 and img.shape == (C, 32, 32)
    for c in range(C):
        assert np.unravel_index(np.argmax(img[c]), img[c].shape) == (16, 16), c
        assert vmax[c] == img[c].max() and vmin[c] == img[c].min()
    rec = C * 2 * N * N
    data = np.fromfile(dump, np.float32)
    assert data.size >= rec and data.size % rec == 0
    for v in data.reshape(-1, C, 2, N, N):
        assert np.array_equal(v[:, 0], v[:, 0].transpose(0, 2, 1))
        assert np.array_equal(v[:, 1], -v[:, 1].transpose(0, 2, 1))
        assert (np.diagonal(v[:, 1], axis1=1, axis2=2) == 0).all()
        assert (np.diagonal(v[:, 0], axis1=1, axis2=2) > 0).all()
    stats = [l for l in err.splitlines() if "imissed=" in l]
    assert stats and all("imissed=0 misalign=0" in l for l in stats)


def test_cli_orbiting_planet_moves_along_the_orbit():
    """--rho/--torbit/--refresh-hz (src/tx_host.cu:98-100,116-122,144-147): the synthetic
    transmitter regenerates its replay slots with the planet at rho (cos wt, sin wt). The
    published frames (dV since the previous frame) put the planet, the brightest pixel more than
    2 pixels off centre in the highest channel, at radius 16 rho = 8 pixels around the centre, and
    its angle advances counter-clockwise over more than half a turn."""
    N, C, rho, torbit = 256, 64, 0.5, 4.0
    shm = f"/mixer_cli_orbit_{os.getpid()}"
    cmd = [BIN, "--seconds", "6", "--elements", str(N), "--grid-x", "16", "--channels", str(C),
           "--k-update", "1024", "--flush", "2", "--ng", "32", "--rate", "0.5", "--fps", "5",
           "--rho", str(rho), "--torbit", str(torbit), "--refresh-hz", "100", "--shm", shm,
           "--no-wrap-check"]  # the synthetic source clips to [-127, 127]
    print(" ".join(cmd), flush=True)
    p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    pos = {}
    try:
        deadline = time.time() + 60
        while p.poll() is None and time.time() < deadline:
            fr = read_shm(shm)
            if fr is not None and fr[0] not in pos:
                ws, img = fr[0], fr[1][C - 1]
                yy, xx = np.mgrid[0:32, 0:32]
                off = (xx - 16) ** 2 + (yy - 16) ** 2 > 4
                oy, ox = np.unravel_index(int(np.argmax(np.where(off, img, -np.inf))), img.shape)
                pos[ws] = (int(ox), int(oy))
            time.sleep(0.05)
        out, err = p.communicate(timeout=60)
    finally:
        if p.poll() is None:
            p.kill()
            p.communicate()
    sys.stdout.write(err[-1200:])
    assert p.returncode == 0, err[-2000:]
    seq = [pos[k] for k in sorted(pos)]
    print("planet pixels by frame:", seq, flush=True)
    assert len(seq) >= 8, seq
    r = np.array([np.hypot(x - 16, y - 16) for x, y in seq])
    assert np.mean(np.abs(r - 16 * rho) <= 1.5) >= 0.8, r
    ang = np.unwrap([np.arctan2(y - 16, x - 16) for x, y in seq])
    steps = np.diff(ang)
    assert np.mean(steps > 0) >= 0.7, steps
    assert ang[-1] - ang[0] > np.pi, ang
