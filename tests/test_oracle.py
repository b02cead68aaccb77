"""Pins the oracle (oracle/oracle.cpp) before anything is checked against it.

Each check is chosen so that a plausible mistake in the restatement (a dropped term, a wrong
sign, a transposed operand, an off-by-one cell) fails: numpy's complex128 matmul and FFT are
independent formulations, the k_sky signal gives rank-1 closed forms, and the -128 wrap case is
computed by hand from src/cgemm.cuh:52,59-62 and src/rx_host_corr.cu:71.
"""
import numpy as np
import pytest

import pyoracle as O

RNG = np.random.default_rng(0x4D495830)


def rand_heaps(n_tb, N, C, T, lo=-127, hi=127, rng=RNG):
    return rng.integers(lo, hi + 1, size=(n_tb, N, 2 * C * T), dtype=np.int16).astype(np.int8).view(np.uint8)


def complex_x(bb, n_tb, N, C, T):
    """numpy formulation of the corner-turn: [tb][N][plane][C][T] -> complex [C][N][K]."""
    b = bb.view(np.int8).reshape(n_tb, N, 2, C, T).astype(np.int64)
    re = b[:, :, 0].transpose(2, 1, 0, 3).reshape(C, N, n_tb * T)
    im = b[:, :, 1].transpose(2, 1, 0, 3).reshape(C, N, n_tb * T)
    return re, im


# ---------------------------------------------------------------- a1 wire
def test_header_roundtrip_and_offsets():
    pkt = np.full(8256, 0xAB, np.uint8)
    O.write_header(pkt, 0x01020304)
    # config.h:10-14: bytes 0..47 zero (eth/ip/udp are NIC/TX business), seq BE at 48..51, 52..63 zero
    assert (pkt[:48] == 0).all() and (pkt[52:64] == 0).all()
    assert list(pkt[48:52]) == [1, 2, 3, 4]
    assert (pkt[64:] == 0xAB).all()  # payload untouched
    for s in (0, 1, 255, 256, 0xFFFFFFFF, 123456789):
        O.write_header(pkt, s)
        assert O.read_seq(pkt) == s
        assert int.from_bytes(bytes(pkt[48:52]), "big") == s


# ---------------------------------------------------------------- a3 corner-turn
@pytest.mark.parametrize("n_tb,N,C,T", [(3, 32, 4, 32), (2, 64, 8, 32), (1, 16, 128, 32)])
def test_unpack_matches_numpy_and_is_bijection(n_tb, N, C, T):
    bb = rand_heaps(n_tb, N, C, T, -128, 127)
    are, aim, bre, bim = O.unpack(bb, n_tb, N, C, T)
    re, im = complex_x(bb, n_tb, N, C, T)
    assert (are == re).all() and (aim == im).all() and (bre == re).all()
    assert (bim == np.negative(im.astype(np.int8))).all()  # int8 negation wraps -128 -> -128
    # bijection: every (ant, heap, sl) lands in exactly one output slot
    tag = np.arange(n_tb * N * C * T, dtype=np.int64)
    bb2 = np.zeros((n_tb, N, 2 * C * T), np.uint8)
    # encode a unique 8-bit pattern per position in 4 passes of the byte
    seen = np.zeros((C, N, n_tb * T), np.int64)
    for shift in range(0, 32, 8):
        bb2[:, :, : C * T] = ((tag >> shift) & 0xFF).reshape(n_tb, N, C * T).astype(np.uint8)
        a2, _, _, _ = O.unpack(bb2, n_tb, N, C, T)
        seen |= (a2.view(np.uint8).astype(np.int64)) << shift
    assert np.array_equal(np.sort(seen.ravel()), tag)


# ---------------------------------------------------------------- a4 GEMM
@pytest.mark.parametrize("n_tb,N,C", [(2, 32, 3), (4, 64, 2), (1, 96, 1)])
def test_cgemm_equals_complex_matmul(n_tb, N, C):
    T = 32
    bb = rand_heaps(n_tb, N, C, T)
    are, aim, bre, bim = O.unpack(bb, n_tb, N, C, T)
    cre, cim, of = O.cgemm(are, aim, bre, bim)
    assert of == 0
    re, im = complex_x(bb, n_tb, N, C, T)
    x = re + 1j * im
    v = np.einsum("cak,cbk->cab", x, np.conj(x))  # V = X X^H in complex128, exact here
    assert np.array_equal(cre, v.real.astype(np.int64))
    assert np.array_equal(cim, v.imag.astype(np.int64))
    # Hermitian, real diagonal = sum |x|^2
    assert np.array_equal(cre, cre.transpose(0, 2, 1))
    assert np.array_equal(cim, -cim.transpose(0, 2, 1))
    d = np.einsum("cak,cak->ca", re, re) + np.einsum("cak,cak->ca", im, im)
    assert np.array_equal(np.diagonal(cre, axis1=1, axis2=2), d)
    assert (np.diagonal(cim, axis1=1, axis2=2) == 0).all()
    # xengine (the timed CPU correlator) is the same math through a different loop structure
    vre, vim, of2 = O.xengine(bb, n_tb, N, C, T)
    assert of2 == 0 and np.array_equal(vre, cre) and np.array_equal(vim, cim)


def test_cgemm_accumulate_is_sum_of_updates():
    n_tb, N, C, T = 2, 32, 2, 32
    b1, b2 = rand_heaps(n_tb, N, C, T), rand_heaps(n_tb, N, C, T)
    u1 = O.unpack(b1, n_tb, N, C, T)
    u2 = O.unpack(b2, n_tb, N, C, T)
    r1, i1, _ = O.cgemm(*u1)
    r12, i12, _ = O.cgemm(*u2, acc_re=r1, acc_im=i1)
    both = np.concatenate([b1, b2], axis=0)
    rb, ib, _ = O.cgemm(*O.unpack(both, 2 * n_tb, N, C, T))
    assert np.array_equal(r12, rb) and np.array_equal(i12, ib)


def test_cgemm_minus128_wrap_by_hand():
    """x_a = (0, -128), x_b = (0, 1), one sample. cgemm.cuh:52 narrows -aim to int8 and
    rx_host_corr.cu:71 stores Bim = (int8)(-im), so wrap8(-(-128)) = -128:
      Vr[a][b] = re_a re_b + wrap(-im_a) wrap(-im_b) = 0 + (-128)(-1) = 128   (true value -128)
      Vr[a][a] = (-128)(-128) = 16384
      Vi[a][b] = re_a wrap(-im_b) + im_a re_b = 0 ; Vi[b][a] = 0
      Vr[b][b] = 1."""
    N, K = 2, 1
    are = np.zeros((1, N, K), np.int8)
    aim = np.array([[[-128], [1]]], np.int8)
    bre = are.copy()
    bim = np.negative(aim)
    cre, cim, of = O.cgemm(are, aim, bre, bim)
    assert of == 0
    assert cre[0].tolist() == [[16384, 128], [128, 1]]
    assert cim[0].tolist() == [[0, 0], [0, 0]]
    # -128 in the real part does not wrap: x_a = (-128, 5), x_b = (3, -7)
    are = np.array([[[-128], [3]]], np.int8)
    aim = np.array([[[5], [-7]]], np.int8)
    cre, cim, _ = O.cgemm(are, aim, are.copy(), np.negative(aim))
    assert cre[0].tolist() == [[16384 + 25, -384 - 35], [-384 - 35, 9 + 49]]
    # Vi[a][b] = re_a*(-im_b) + im_a*re_b = -128*7 + 5*3
    assert cim[0].tolist() == [[0, -128 * 7 + 15], [128 * 7 - 15, 0]]


def test_rank1_sky_closed_form():
    """k_sky signals are constant over t and noise-free, so V = K x x^H exactly:
    Vr_ab = K (qr_a qr_b + qi_a qi_b), Vi_ab = K (qi_a qr_b - qr_a qi_b)."""
    N, gx, C, T, n_tb = 64, 8, 8, 32, 3
    l, m = 0.3, -0.2
    bb = np.stack([np.stack([O.sky_payload(a, gx, C, T, l, m) for a in range(N)])] * n_tb)
    vre, vim = O.v_full_from_heaps(bb, n_tb, N, C, T)
    pay = bb[0].view(np.int8).reshape(N, 2, C, T)[:, :, :, 0].astype(np.int64)  # [N][2][C]
    qr, qi = pay[:, 0, :].T, pay[:, 1, :].T  # [C][N]
    K = n_tb * T
    er = K * (qr[:, :, None] * qr[:, None, :] + qi[:, :, None] * qi[:, None, :])
    ei = K * (qi[:, :, None] * qr[:, None, :] - qr[:, :, None] * qi[:, None, :])
    assert np.array_equal(vre, er) and np.array_equal(vim, ei)
    # channel 0 has no planet (Apl = 0): qr = rint(2*25) = 50, qi = 0 for every element
    assert (qr[0] == 50).all() and (qi[0] == 0).all()


# ---------------------------------------------------------------- a5/a6 flush, diff
def test_flush_and_diff():
    C, N = 2, 32
    vire = RNG.integers(-2**30, 2**30, size=(C, N, N), dtype=np.int64).astype(np.int32)
    viim = RNG.integers(-2**30, 2**30, size=(C, N, N), dtype=np.int64).astype(np.int32)
    cube = RNG.standard_normal((C, 2, N, N)).astype(np.float32)
    exp = cube.copy()
    exp[:, 0] += vire.astype(np.float32)
    exp[:, 1] += viim.astype(np.float32)
    r, i = vire.copy(), viim.copy()
    O.flush(cube, r, i)
    assert np.array_equal(cube, exp) and (r == 0).all() and (i == 0).all()
    vs = RNG.standard_normal(1000).astype(np.float32)
    vp = RNG.standard_normal(1000).astype(np.float32)
    vp0 = vp.copy()
    dv = O.diff(vs, vp)
    assert np.array_equal(dv, vs - vp0) and np.array_equal(vp, vs)


# ---------------------------------------------------------------- a7 grid
def test_grid_all_ones_closed_form():
    """All-ones V on the 16x16 array: cell (dy,dx) collects (16-|dx|)(16-|dy|) baselines
    for |dx|,|dy| <= 15 (rx_host_corr.cu:82); index 16 is never hit."""
    C, N, gx, ng = 1, 256, 16, 32
    dv = np.zeros((C, 2, N, N), np.float32)
    dv[:, 0] = 1.0
    g = O.grid(dv, gx, ng)[0]
    exp = np.zeros((ng, ng))
    for ddy in range(-15, 16):
        for ddx in range(-15, 16):
            exp[ddy % ng, ddx % ng] = (16 - abs(ddx)) * (16 - abs(ddy))
    assert np.array_equal(g.real, exp) and (g.imag == 0).all()
    assert (g[16, :] == 0).all() and (g[:, 16] == 0).all()
    assert g.real.sum() == N * N


def test_grid_hermitian_and_bincount():
    n_tb, N, C, T, gx, ng = 2, 64, 3, 32, 8, 32
    bb = rand_heaps(n_tb, N, C, T)
    vre, vim = O.v_full_from_heaps(bb, n_tb, N, C, T)
    dv = np.stack([vre, vim], axis=1).astype(np.float32)
    g = O.grid(dv, gx, ng)
    neg = (-np.arange(ng)) % ng
    assert np.allclose(g[:, neg][:, :, neg], np.conj(g), rtol=0, atol=1e-6 * np.abs(g).max())
    _, _, _, _, g2 = O.image_from_v(vre, vim, gx, ng)
    assert np.allclose(g, g2, rtol=0, atol=1e-6 * np.abs(g).max())


# ---------------------------------------------------------------- a8 FFT
def test_dft2_matches_numpy_fft2():
    for ng in (32, 64):
        g = RNG.standard_normal((3, ng, ng)) + 1j * RNG.standard_normal((3, ng, ng))
        f = O.dft2(g)
        ref = np.fft.fft2(g)  # forward, unnormalised: cuFFT CUFFT_FORWARD convention
        assert np.abs(f - ref).max() <= 1e-10 * np.abs(ref).max()


def test_dft2_closed_forms():
    ng = 32
    g = np.zeros((2, ng, ng), complex)
    g[0, 0, 0] = 3.0  # delta at the origin -> flat
    y, x = np.mgrid[0:ng, 0:ng]
    g[1] = np.exp(2j * np.pi * (5 * y + 7 * x) / ng)  # plane wave -> single peak at (5,7)
    f = O.dft2(g)
    assert np.allclose(f[0], 3.0, atol=1e-12)
    peak = np.zeros((ng, ng), complex)
    peak[5, 7] = ng * ng
    assert np.allclose(f[1], peak, atol=1e-9)


# ---------------------------------------------------------------- a9 publish
def test_publish_fftshift_and_minmax():
    ng = 32
    f = np.zeros((2, ng, ng), complex)
    f[0, 0, 0] = 7.0      # DC -> image centre (16,16)
    f[1, 3, 30] = -2.0    # (ky,kx)=(3,30) -> (y,x) = ((3+16)%32, (30+16)%32) = (19,14)
    f[1, 0, 0] = 1.0 + 9j  # imaginary part is dropped
    img, vmax, vmin = O.publish(f)
    assert img[0, 16, 16] == 7.0 and np.count_nonzero(img[0]) == 1
    assert img[1, 19, 14] == -2.0 and img[1, 16, 16] == 1.0
    assert vmax.tolist() == [7.0, 1.0] and vmin.tolist() == [0.0, -2.0]


def test_star_planet_image_positions():
    """Star at l=m=0 images to the centre (16,16) in every channel; a planet at (l,m)
    = (0.5, 0.25) lands at x = 16 + 16 l, y = 16 + 16 m; its brightness grows with channel."""
    N, gx, C, T, n_tb, ng = 256, 16, 16, 32, 1, 32
    l, m = 0.5, 0.25
    bb = np.stack([np.stack([O.sky_payload(a, gx, C, T, l, m) for a in range(N)])] * n_tb)
    vre, vim = O.v_full_from_heaps(bb, n_tb, N, C, T)
    img, vmax, vmin, f, g = O.image_from_v(vre, vim, gx, ng)
    for c in range(C):
        assert np.unravel_index(np.argmax(img[c]), img[c].shape) == (16, 16)
    # planet pixel grows with channel (Apl^2 ~ c, tx_host.cu:48); int8 rounding makes steps uneven
    pl = img[:, 16 + 4, 16 + 8].astype(np.float64)
    assert pl[-1] > 3 * pl[2] and np.corrcoef(np.arange(C), pl)[0, 1] > 0.95
