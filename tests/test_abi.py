"""CPU checks of the C-ABI boundary: the library loads, exports every symbol include/mixer.h
declares, and its host-only entry points (config, wire sizes, synthetic source, Corr32 size)
behave; the synthetic source is checked against the oracle's k_sky restatement."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import mixer_amd as M
import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "mixer.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(mixer_[a-z0-9_]+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = M.lib()
    names = declared_functions()
    assert len(names) >= 40
    for n in names:
        assert hasattr(L, n), n
    bound = {s[0] for s in M.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


def test_binding_constants_match_the_header():
    """Status codes (enum) and config flags (#define) of the binding are the header's."""
    src = open(os.path.join(ROOT, "include", "mixer.h")).read()
    flags = {k: int(v) for k, v in re.findall(r"#define\s+MIXER_(CFG_[A-Z0-9_]+)\s+(\d+)", src)}
    assert len(flags) >= 6 and all(getattr(M, k) == v for k, v in flags.items()), flags
    status = {k: int(v) for k, v in re.findall(r"^\s*MIXER_(OK|ERR_[A-Z0-9_]+)\s*=\s*(-?\d+)", src, flags=re.M)}
    assert len(status) >= 8 and all(getattr(M, k) == v for k, v in status.items()), status


def test_default_cfg_is_reference_constants():
    c = M.default_cfg()
    # src/config.h:26-38 and src/rx_host_corr.cu:47,51
    assert (c.n_elem, c.grid_x, c.n_chan, c.t_pkt, c.hdr_bytes, c.seq_byte, c.ng, c.pkt_pitch) == (
        256, 16, 128, 32, 64, 48, 32, 16384)
    assert M.lib().mixer_payload_bytes(C.byref(c)) == 8192
    assert M.lib().mixer_wire_bytes(C.byref(c)) == 8256
    assert M.lib().mixer_validate_cfg(C.byref(c)) == M.OK
    assert M.corr32_bytes(c) == 525336  # sizeof(Corr32) at the reference geometry


@pytest.mark.parametrize("field,value,status", [
    ("n_elem", 48, M.ERR_ARG), ("n_elem", 1056, M.ERR_ARG), ("n_elem", 288, M.ERR_ARG),
    ("t_pkt", 16, M.ERR_ARG), ("k_integ", 65536, M.ERR_OVERFLOW), ("k_update", 100, M.ERR_ARG),
    ("ng", 16, M.ERR_ARG), ("ng", 48, M.ERR_ARG), ("grid_x", 7, M.ERR_ARG), ("pkt_pitch", 8000, M.ERR_ARG),
])
def test_validate_rejects(field, value, status):
    c = M.default_cfg()
    setattr(c, field, value)
    if field == "k_integ":
        c.k_update = 1024
    assert M.lib().mixer_validate_cfg(C.byref(c)) == status


def test_validate_accepts_configs():
    for n, gx, ch, k, ng in [(256, 16, 128, 8192, 32), (1024, 32, 256, 4096, 64), (64, 8, 16, 1024, 32),
                             (512, 16, 200, 16384, 64), (128, 16, 64, 2048, 32)]:
        c = M.default_cfg(n_elem=n, grid_x=gx, n_chan=ch, k_integ=k, k_update=k, ng=ng,
                          pkt_pitch=max(16384, 64 + 64 * ch))
        assert M.lib().mixer_validate_cfg(C.byref(c)) == M.OK, (n, ch)


def test_synth_noise_free_matches_k_sky_restatement():
    """sigma = 0: every payload equals the oracle's k_sky restatement (src/tx_host.cu:41-60)."""
    c = M.default_cfg(n_elem=64, grid_x=8, n_chan=16)
    sky = M.default_sky(sigma=0.0, planet_l=0.3, planet_m=-0.45)
    pk = M.synth_packets(c, sky, seed=1, seq0=64 * 5, npkts=128)
    for i in range(128):
        seq = 64 * 5 + i
        assert O.read_seq(np.ascontiguousarray(pk[i])) == seq
        assert (pk[i, 42:48] == 0).all() and (pk[i, 52:64] == 0).all()  # reserved bytes
        exp = O.sky_payload(seq % 64, 8, 16, 32, 0.3, -0.45)
        assert np.array_equal(pk[i, 64:], exp), i


def test_synth_noise_is_seeded_and_gaussian():
    c = M.default_cfg(n_elem=32, grid_x=8, n_chan=64)
    sky = M.default_sky(sigma=0.5, planet_scale=0.0, star_amp=0.0)
    a = M.synth_packets(c, sky, seed=7, seq0=0, npkts=256)
    b = M.synth_packets(c, sky, seed=7, seq0=0, npkts=256, nthreads=3)
    d = M.synth_packets(c, sky, seed=8, seq0=0, npkts=256)
    assert np.array_equal(a, b) and not np.array_equal(a, d)
    v = a[:, 64:].view(np.int8).astype(np.float64)
    assert abs(v.mean()) < 0.2 and abs(v.std() - 12.5) < 0.3  # 25 * sigma LSB
    # noise is keyed by absolute (element, channel, sample): a later start reproduces the same packets
    e = M.synth_packets(c, sky, seed=7, seq0=32, npkts=224)
    assert np.array_equal(e, a[32:])


def test_wire_header_is_the_transmitters():
    """eth/ip/udp fields of src/tx_host.cu:69-79 with the src/config.h:53-58 defaults; the IPv4
    checksum is pinned by its defining property (one's-complement sum of the header = 0xFFFF)."""
    c = M.default_cfg()
    h = M.build_header(c, 0x01020304)
    assert bytes(h[0:6]) == bytes([0xa0, 0x88, 0xc2, 0x0d, 0x5e, 0x28]) and not h[6:12].any()
    assert bytes(h[12:14]) == b"\x08\x00"
    ip = h[14:34]
    assert ip[0] == 0x45 and ip[8] == 64 and ip[9] == 17
    assert int.from_bytes(bytes(ip[2:4]), "big") == 8256 - 14
    assert bytes(ip[12:16]) == bytes([10, 0, 0, 1]) and bytes(ip[16:20]) == bytes([10, 0, 0, 2])
    s = sum(int.from_bytes(bytes(ip[i:i + 2]), "big") for i in range(0, 20, 2))
    while s >> 16:
        s = (s & 0xFFFF) + (s >> 16)
    assert s == 0xFFFF
    udp = h[34:42]
    assert int.from_bytes(bytes(udp[0:2]), "big") == 4096 == int.from_bytes(bytes(udp[2:4]), "big")
    assert int.from_bytes(bytes(udp[4:6]), "big") == 8256 - 34 and not udp[6:8].any()
    assert bytes(h[48:52]) == bytes([1, 2, 3, 4])
    assert not h[42:48].any() and not h[52:64].any()
    # a planted mistake in the checksum is caught by the same property
    bad = ip.copy()
    bad[10] ^= 1
    s = sum(int.from_bytes(bytes(bad[i:i + 2]), "big") for i in range(0, 20, 2))
    while s >> 16:
        s = (s & 0xFFFF) + (s >> 16)
    assert s != 0xFFFF
    # synthetic packets carry that header with their own seq
    sky = M.default_sky(sigma=0.0)
    pk = M.synth_packets(c, sky, seed=1, seq0=513, npkts=3)
    for k in range(3):
        exp = M.build_header(c, 513 + k)
        assert np.array_equal(pk[k, :64], exp)


def test_orbit_sky_quarter_periods_and_payload():
    """The transmitter's orbiting planet (src/tx_host.cu:116-120): at t = 0, T/4, T/2, 3T/4 the
    planet sits at (rho, 0), (0, rho), (-rho, 0), (0, -rho), counter-clockwise, always at radius
    rho; the rest of the sky is untouched. A slot regenerated at an orbit time (sigma = 0)
    carries the oracle's k_sky payload for that planet, and its image puts the planet at pixel
    (16 + 16 l, 16 + 16 m)."""
    rho, T = 0.5, 3.0
    base = M.default_sky(sigma=0.0, star_amp=2.0)
    want = {0.0: (rho, 0.0), T / 4: (0.0, rho), T / 2: (-rho, 0.0), 3 * T / 4: (0.0, -rho), T: (rho, 0.0)}
    for t, (l, m) in want.items():
        s = M.orbit_sky(base, rho, T, t)
        assert abs(s.planet_l - l) < 1e-12 and abs(s.planet_m - m) < 1e-12, t
        assert s.star_amp == base.star_amp and s.sigma == base.sigma and s.planet_scale == base.planet_scale
    prev = None
    for t in np.linspace(0.01, T / 4 - 0.01, 7):  # first quadrant, angle increasing
        s = M.orbit_sky(base, rho, T, float(t))
        assert abs(np.hypot(s.planet_l, s.planet_m) - rho) < 1e-12
        assert s.planet_l > 0 and s.planet_m > 0
        ang = np.arctan2(s.planet_m, s.planet_l)
        assert prev is None or ang > prev
        prev = ang
    with pytest.raises(M.MixerError):
        M.orbit_sky(base, rho, 0.0, 1.0)
    N, gx, Cn = 256, 16, 128
    c = M.default_cfg(n_elem=N, grid_x=gx, n_chan=Cn)
    for t, (px, py) in ((T / 4, (16, 24)), (T / 2, (8, 16))):
        s = M.orbit_sky(base, rho, T, t)
        pk = M.synth_packets(c, s, seed=3, seq0=0, npkts=N)
        for a in (0, 17, 255):
            assert np.array_equal(pk[a, 64:], O.sky_payload(a, gx, Cn, 32, s.planet_l, s.planet_m)), (t, a)
        bb = np.ascontiguousarray(pk[:, 64:]).reshape(1, N, 64 * Cn)
        vre, vim = O.v_full_from_heaps(bb, 1, N, Cn, 32)
        img = O.image_from_v(vre, vim, gx, 32)[0]
        ch = Cn - 1  # the planet grows with channel (src/tx_host.cu:48): brightest in the last
        yy, xx = np.mgrid[0:32, 0:32]
        off = (xx - 16) ** 2 + (yy - 16) ** 2 > 4
        oy, ox = np.unravel_index(int(np.argmax(np.where(off, img[ch], -np.inf))), (32, 32))
        assert (ox, oy) == (px, py), (t, ox, oy)
