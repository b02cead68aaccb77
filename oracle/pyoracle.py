"""ctypes binding of the oracle (oracle.cpp): the CPU restatement of the reference X-engine path.

TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import this module; the product library never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def _cpu_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
        return all(x in flags for x in ("avx512f", "avx512bw", "avx512vl", "avx512dq", "avx512cd"))
    except OSError:
        return False


def lib():
    global _lib
    if _lib is not None:
        return _lib
    name = "liboracle_v4.so" if _cpu_has_avx512() else "liboracle.so"
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library {path} is not built (run `make -C oracle`)")
    L = C.CDLL(path)
    P = C.c_void_p
    i = C.c_int
    L.orc_write_header.argtypes = [P, i, i, C.c_uint32]
    L.orc_read_seq.argtypes = [P, i]
    L.orc_read_seq.restype = C.c_uint32
    L.orc_unpack.argtypes = [P, i, i, i, i, P, P, P, P]
    L.orc_cgemm.argtypes = [i, i, i, i, P, P, P, P, P, P, i]
    L.orc_cgemm.restype = C.c_long
    L.orc_xengine.argtypes = [P, i, i, i, i, i, i, P, P, i]
    L.orc_xengine.restype = C.c_long
    L.orc_flush.argtypes = [P, P, P, i, i]
    L.orc_diff.argtypes = [P, P, P, C.c_size_t]
    L.orc_grid.argtypes = [P, i, i, i, i, P]
    L.orc_dft2.argtypes = [P, i, i, P]
    L.orc_publish.argtypes = [P, i, i, P, P, P]
    L.orc_sky_payload.argtypes = [i, i, i, i, i, C.c_double, C.c_double, P]
    _lib = L
    return L


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def write_header(pkt: np.ndarray, seq: int, hdr_bytes=64, seq_byte=48):
    lib().orc_write_header(_p(pkt), hdr_bytes, seq_byte, seq)


def read_seq(pkt: np.ndarray, seq_byte=48) -> int:
    return int(lib().orc_read_seq(_p(pkt), seq_byte))


def unpack(bb: np.ndarray, n_tb: int, N: int, C_: int, T: int):
    """Corner-turn restatement: returns (are, aim, bre, bim), each int8 [C][N][K]."""
    K = n_tb * T
    outs = [np.empty((C_, N, K), np.int8) for _ in range(4)]
    lib().orc_unpack(_p(bb), n_tb, N, C_, T, *[_p(o) for o in outs])
    return outs


def cgemm(are, aim, bre, bim, acc_re=None, acc_im=None):
    """Reference cgemm math, full square. Inputs [batch][M][K]; returns (cre, cim, overflow)."""
    batch, M, K = are.shape
    N = bre.shape[1]
    if acc_re is None:
        cre = np.zeros((batch, M, N), np.int32)
        cim = np.zeros((batch, M, N), np.int32)
        acc = 0
    else:
        cre, cim, acc = acc_re.copy(), acc_im.copy(), 1
    of = lib().orc_cgemm(M, N, K, batch, _p(are), _p(aim), _p(bre), _p(bim), _p(cre), _p(cim), acc)
    return cre, cim, int(of)


def xengine(bb: np.ndarray, n_tb: int, N: int, C_: int, T: int, c0=0, c1=None, nthreads=0):
    """CPU correlator: heaps -> full-square int32 V for channels [c0, c1)."""
    c1 = C_ if c1 is None else c1
    vre = np.empty((c1 - c0, N, N), np.int32)
    vim = np.empty((c1 - c0, N, N), np.int32)
    of = lib().orc_xengine(_p(bb), n_tb, N, C_, T, c0, c1, _p(vre), _p(vim), nthreads)
    return vre, vim, int(of)


def flush(vcube: np.ndarray, vire: np.ndarray, viim: np.ndarray):
    Cc, N, _ = vire.shape
    lib().orc_flush(_p(vcube), _p(vire), _p(viim), Cc, N)


def diff(vs: np.ndarray, vp: np.ndarray):
    dv = np.empty_like(vs)
    lib().orc_diff(_p(vs), _p(vp), _p(dv), vs.size)
    return dv


def grid(dv: np.ndarray, grid_x: int, ng: int):
    """dv: f32 planar cube [C][2][N][N] -> complex128 G [C][ng][ng]."""
    Cc, _, N, _ = dv.shape
    g = np.empty((Cc, ng, ng, 2), np.float64)
    lib().orc_grid(_p(dv), Cc, N, grid_x, ng, _p(g))
    return g[..., 0] + 1j * g[..., 1]


def dft2(g: np.ndarray):
    """FP64 direct forward 2-D DFT (unnormalised) of complex [C][ng][ng]."""
    Cc, ng, _ = g.shape
    gi = np.ascontiguousarray(np.stack([g.real, g.imag], axis=-1))
    f = np.empty_like(gi)
    lib().orc_dft2(_p(gi), Cc, ng, _p(f))
    return f[..., 0] + 1j * f[..., 1]


def publish(f: np.ndarray):
    """fftshift + Re + per-channel max/min of complex [C][ng][ng]; returns (img f32, vmax, vmin)."""
    Cc, ng, _ = f.shape
    fi = np.ascontiguousarray(np.stack([f.real, f.imag], axis=-1))
    img = np.empty((Cc, ng, ng), np.float32)
    vmax = np.empty(Cc, np.float32)
    vmin = np.empty(Cc, np.float32)
    lib().orc_publish(_p(fi), Cc, ng, _p(img), _p(vmax), _p(vmin))
    return img, vmax, vmin


def sky_payload(a: int, grid_x: int, C_: int, T: int, l: float, m: float, nch_gen=256):
    pay = np.empty(2 * C_ * T, np.uint8)
    lib().orc_sky_payload(a, grid_x, C_, T, nch_gen, l, m, _p(pay))
    return pay


def v_full_from_heaps(bb: np.ndarray, n_tb: int, N: int, C_: int, T: int):
    """Convenience: heaps -> (vre, vim) int32 [C][N][N] through unpack + cgemm (reference path)."""
    are, aim, bre, bim = unpack(bb, n_tb, N, C_, T)
    cre, cim, of = cgemm(are, aim, bre, bim)
    assert of == 0, "int32 overflow in oracle cgemm"
    return cre, cim


def image_from_v(vre: np.ndarray, vim: np.ndarray, grid_x: int, ng: int):
    """Exact V -> FP64 grid -> FP64 DFT -> publish (f32 img, vmax, vmin), and the complex F."""
    Cc, N, _ = vre.shape
    g = np.zeros((Cc, ng, ng), np.complex128)
    a = np.arange(N)
    px, qy = a % grid_x, a // grid_x
    dy = (qy[:, None] - qy[None, :]) % ng
    dx = (px[:, None] - px[None, :]) % ng
    cell = (dy * ng + dx).ravel()
    for c in range(Cc):
        w = vre[c].astype(np.float64).ravel() + 1j * vim[c].astype(np.float64).ravel()
        g[c] = np.bincount(cell, weights=w.real, minlength=ng * ng).reshape(ng, ng) + 1j * np.bincount(
            cell, weights=w.imag, minlength=ng * ng).reshape(ng, ng)
    f = dft2(g)
    img, vmax, vmin = publish(f)
    return img, vmax, vmin, f, g
