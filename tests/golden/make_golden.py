#!/usr/bin/env python3
"""Writes tests/golden/xengine_small.npz: a small end-to-end case computed with numpy only
(independent of the oracle's C++ code), in the reference's layouts.

Case: 64 elements on an 8x8 grid, 3 channels, 2 time blocks (64 samples), complex int8 in
[-127, 127] plus real-part -128 samples (which do not wrap), 32x32 image.
  heaps      uint8 [2][64][2*3*32]    header-stripped config.h payloads (planar re | im)
  vre, vim   int64 [3][64][64]        V = X X^H (src/cgemm.cuh:59-62 for this domain)
  v_f32      f32   [3][2][64][64]     the --dump record of one integration
  grid       c128  [3][32][32]        sum of V over baselines with (qa-qb, pa-pb) mod 32
  img        f32   [3][32][32]        Re of the fftshifted forward unnormalised 2-D DFT
  vmax, vmin f32   [3]
Run from the repository root: python tests/golden/make_golden.py
"""
import os

import numpy as np

N, GX, C, NTB, T, NG = 64, 8, 3, 2, 32, 32


def main():
    rng = np.random.default_rng(0x4D495830)
    x = rng.integers(-127, 128, size=(NTB, N, 2, C, T), dtype=np.int64)
    x[0, 5, 0, 1, :4] = -128  # real part -128: exact in the reference (no negation of re)
    heaps = x.astype(np.int8).reshape(NTB, N, 2 * C * T).view(np.uint8)
    re = x[:, :, 0].transpose(2, 1, 0, 3).reshape(C, N, NTB * T)
    im = x[:, :, 1].transpose(2, 1, 0, 3).reshape(C, N, NTB * T)
    X = re + 1j * im
    V = np.matmul(X, np.conj(X).transpose(0, 2, 1))
    vre, vim = V.real.astype(np.int64), V.imag.astype(np.int64)
    v_f32 = np.stack([vre, vim], axis=1).astype(np.float32)
    a = np.arange(N)
    px, qy = a % GX, a // GX
    cell = (((qy[:, None] - qy[None, :]) % NG) * NG + (px[:, None] - px[None, :]) % NG).ravel()
    grid = np.zeros((C, NG * NG), complex)
    for c in range(C):
        np.add.at(grid[c], cell, V[c].ravel())
    grid = grid.reshape(C, NG, NG)
    F = np.fft.fft2(grid)
    img = np.real(np.fft.fftshift(F, axes=(1, 2))).astype(np.float32)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "xengine_small.npz")
    np.savez_compressed(out, heaps=heaps, vre=vre, vim=vim, v_f32=v_f32, grid=grid, img=img,
                        vmax=img.reshape(C, -1).max(1), vmin=img.reshape(C, -1).min(1),
                        geometry=np.array([N, GX, C, NTB, T, NG]))
    print("wrote", out)


if __name__ == "__main__":
    main()
