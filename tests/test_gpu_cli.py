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
