"""ASCII view of one channel of the published image cube (the Corr32 region in /dev/shm).

Reads what the reference's peek32 viewer reads (src/peek32.cpp:13-64): header
{magic 0x43523343, w, h, nch, write_seq}, then img[nch][h][w], vmax[nch], vmin[nch], all f32.
Unlike peek32 it takes the image size from the header (32 or 64) and honours the writer flag
(bit 63 of write_seq, set while a frame is being copied in) so it never shows a torn frame.

    python -m mixer_amd.peek [channel] [--shm /corr32]

channel defaults to the one with the brightest off-centre pixel (the planet).
"""
from __future__ import annotations

import argparse
import mmap
import os
import struct
import sys
import time

import numpy as np

MAGIC = 0x43523343
HDR = 24
LUT = " .:-=+*#%@"


def read_frame(buf, retries: int = 1000):
    """(write_seq, img [nch][h][w], vmax, vmin) from a Corr32 region, or None before any frame."""
    for _ in range(retries):
        magic, w, h, nch = struct.unpack_from("<4I", buf, 0)
        if magic != MAGIC:
            raise ValueError(f"not a Corr32 image cube (magic {magic:08x})")
        ws1 = struct.unpack_from("<Q", buf, 16)[0]
        if ws1 == 0:
            return None
        if ws1 >> 63:  # writer busy
            time.sleep(1e-4)
            continue
        npix = nch * h * w
        data = np.frombuffer(buf, np.float32, npix + 2 * nch, HDR).copy()
        ws2 = struct.unpack_from("<Q", buf, 16)[0]
        if ws1 == ws2:
            img = data[:npix].reshape(nch, h, w)
            return ws1, img, data[npix:npix + nch], data[npix + nch:]
    raise TimeoutError("writer never released the frame")


def pick_channel(img: np.ndarray) -> int:
    """The channel with the brightest pixel more than 2 pixels from the centre (the planet)."""
    nch, h, w = img.shape
    y, x = np.mgrid[0:h, 0:w]
    off = (x - w // 2) ** 2 + (y - h // 2) ** 2 > 4
    return int(np.argmax(np.where(off[None], img, -np.inf).reshape(nch, -1).max(axis=1)))


def render(ws: int, img: np.ndarray, vmax: np.ndarray, vmin: np.ndarray, ch: int | None = None) -> str:
    nch, h, w = img.shape
    if ch is None or ch < 0:
        ch = pick_channel(img)
    ch = min(max(ch, 0), nch - 1)
    im, vx, vn = img[ch], float(vmax[ch]), float(vmin[ch])
    lines = [f"corr cube seq={ws}  channel {ch}/{nch}  vmax={vx:.3g} vmin={vn:.3g}  [{w}x{h}]"]
    n = (im - vn) / (vx - vn) if vx > vn else np.zeros_like(im)
    li = np.clip((n * 9.0).astype(int), 0, 9)
    for y in range(h):
        lines.append("".join(LUT[v] * 2 for v in li[y]))
    py, px = np.unravel_index(int(np.argmax(im)), im.shape)
    yy, xx = np.mgrid[0:h, 0:w]
    off = (xx - w // 2) ** 2 + (yy - h // 2) ** 2 > 4
    oy, ox = np.unravel_index(int(np.argmax(np.where(off, im, -np.inf))), im.shape)
    pk, opk = float(im[py, px]), float(im[oy, ox])
    lines.append(f"star peak (x={px},y={py}) val={pk:.3g}  [center=({w // 2},{h // 2})]")
    lines.append(f"planet    (x={ox},y={oy}) val={opk:.3g}  ({100.0 * opk / pk if pk else 0.0:.1f}% of star)")
    return "\n".join(lines)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("channel", nargs="?", type=int, default=-1)
    ap.add_argument("--shm", default="/corr32")
    a = ap.parse_args(argv)
    path = "/dev/shm/" + a.shm.lstrip("/")
    try:
        fd = os.open(path, os.O_RDONLY)
    except OSError as e:
        print(f"open {path}: {e}", file=sys.stderr)
        return 1
    try:
        buf = mmap.mmap(fd, 0, mmap.MAP_SHARED, mmap.PROT_READ)
    finally:
        os.close(fd)
    try:
        fr = read_frame(buf)
    except ValueError as e:
        print(e)
        return 2
    if fr is None:
        print("no frames yet")
        return 2
    print(render(*fr, ch=a.channel))
    return 0


if __name__ == "__main__":
    sys.exit(main())
