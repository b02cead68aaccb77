"""ctypes binding of libmixer (include/mixer.h), the gfx950 X-engine.

The binding mirrors the reference's operator surface (src/rx_host_corr.cu: k_packAB, cgemm_blk_t,
k_flush_cube, k_grid_cube + cuFFT, the --dump record and the Corr32 image cube). It loads the
in-tree library daqiri-mixer_amd/lib/libmixer.so and raises if it is missing: there is no CPU
fallback. Device memory, streams and events come from the library's own plumbing entry points,
so exactly one HIP runtime is loaded in the process.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libmixer.so")

OK, ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_STATE, ERR_BUDGET, ERR_NODEV, ERR_OVERFLOW = 0, -1, -2, -3, -4, -5, -6, -7
CFG_CUMULATIVE, CFG_NO_IMAGE, CFG_NOFLOP, CFG_AUTOTUNE, CFG_FRAME_DV, CFG_NO_WRAP_CHECK = 1, 2, 4, 8, 16, 32
H2D, D2H, D2D = 0, 1, 2


class MixerError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: {strerror(status)} ({status})")


class ArrayCfg(C.Structure):
    _fields_ = [
        ("n_elem", C.c_int32), ("grid_x", C.c_int32), ("n_chan", C.c_int32), ("t_pkt", C.c_int32),
        ("hdr_bytes", C.c_int32), ("seq_byte", C.c_int32), ("k_integ", C.c_int32),
        ("k_update", C.c_int32), ("ng", C.c_int32), ("flags", C.c_int32), ("pkt_pitch", C.c_int64),
    ]

    @property
    def payload(self) -> int:
        return 2 * self.n_chan * self.t_pkt

    @property
    def wire(self) -> int:
        return self.hdr_bytes + self.payload


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("heaps_in", "time_blocks", "vupdates", "integrations",
                                           "imissed", "misalign", "seq_gaps", "wrap128")] + [
        ("last_dump_latency_ms", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("slots", C.c_int32), ("prefetch", C.c_int32), ("ct_chan", C.c_int32), ("staging", C.c_int32),
                ("mfma", C.c_int32), ("integrate", C.c_int32)]

    def as_dict(self):
        return {"slots": self.slots, "prefetch": self.prefetch, "ct_chan": self.ct_chan,
                "staging": self.staging, "mfma": self.mfma, "integrate": self.integrate}


class BatchTuning(C.Structure):
    _fields_ = [("pack", C.c_int32), ("tb_target", C.c_int32), ("use_graph", C.c_int32),
                ("staging", C.c_int32), ("mfma", C.c_int32)]

    def as_dict(self):
        return {"pack": self.pack, "tb_target": self.tb_target, "use_graph": self.use_graph,
                "staging": self.staging, "mfma": self.mfma}


class BatchInfo(C.Structure):
    _fields_ = [("n_jobs", C.c_int32), ("n_items", C.c_int32), ("n_wgs", C.c_int32), ("graph", C.c_int32),
                ("payload_bytes", C.c_uint64), ("exec_ops", C.c_uint64), ("device_bytes", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("hbm_budget", "hbm_used", "admitted", "rejected", "chunks",
                                          "bytes_in", "integrations", "lat_n")] + \
               [(n, C.c_double) for n in ("lat_p50_ms", "lat_p99_ms", "lat_max_ms", "lat_mean_ms")] + \
               [("launches", C.c_uint64)]


class ServeSpec(C.Structure):
    _fields_ = [("duration_s", C.c_double), ("aggregate_gbps", C.c_double), ("seed", C.c_uint64)]


class ServeReport(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("chunks", "integrations", "late_submits")] + \
               [(n, C.c_double) for n in ("offered_gbps", "achieved_gbps", "wall_s", "lat_p50_ms",
                                          "lat_p99_ms", "lat_max_ms")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Wire(C.Structure):
    _fields_ = [("eth_dst", C.c_uint8 * 6), ("eth_src", C.c_uint8 * 6), ("ip_src", C.c_uint32),
                ("ip_dst", C.c_uint32), ("udp_port", C.c_uint16), ("reserved", C.c_uint16)]


class Sky(C.Structure):
    _fields_ = [("planet_l", C.c_double), ("planet_m", C.c_double), ("sigma", C.c_double),
                ("star_amp", C.c_double), ("planet_scale", C.c_double), ("nch_gen", C.c_int32),
                ("n_src", C.c_int32), ("src_l", C.c_double * 4), ("src_m", C.c_double * 4),
                ("src_amp", C.c_double * 4)]


_lib = None
P, I32, SZ, U64 = C.c_void_p, C.c_int, C.c_size_t, C.c_uint64

# (name, restype, argtypes) for every entry point in include/mixer.h
SIGNATURES = [
    ("mixer_default_cfg", None, [P]),
    ("mixer_validate_cfg", I32, [P]),
    ("mixer_payload_bytes", SZ, [P]),
    ("mixer_wire_bytes", SZ, [P]),
    ("mixer_version", C.c_char_p, []),
    ("mixer_strerror", C.c_char_p, [I32]),
    ("mixer_device_count", I32, [P]),
    ("mixer_tiles_bytes", SZ, [P, I32]),
    ("mixer_corner_turn", I32, [P, P, P, I32, P, P]),
    ("mixer_acc_bytes", SZ, [P]),
    ("mixer_acc_slots", I32, [P]),
    ("mixer_xcorr", I32, [P, P, P, I32, P, I32]),
    ("mixer_integrate", I32, [P, P, P, P, I32]),
    ("mixer_xcorr_heaps", I32, [P, P, P, I32, P, I32, P]),
    ("mixer_image", I32, [P, P, P, P, P, P]),
    ("mixer_image_noalloc", I32, [P, P, P, P, P, P]),
    ("mixer_set_tuning", I32, [I32, P]),
    ("mixer_get_tuning", I32, [P, P]),
    ("mixer_autotune", I32, [P, I32, P]),
    ("mixer_xcorr_plain", I32, [P, P, P, I32, P, P, P]),
    ("mixer_create", I32, [P, I32, P]),
    ("mixer_destroy", I32, [P]),
    ("mixer_get_cfg", I32, [P, P]),
    ("mixer_submit_heaps", I32, [P, P, SZ, I32, U64, P]),
    ("mixer_submit_device", I32, [P, P, I32, U64, P]),
    ("mixer_set_ingest_slots", I32, [P, I32]),
    ("mixer_token_done", I32, [P, U64]),
    ("mixer_token_wait", I32, [P, U64]),
    ("mixer_sync", I32, [P]),
    ("mixer_dump", I32, [P, P, I32, P]),
    ("mixer_read_image", I32, [P, P, P, P, P]),
    ("mixer_read_frame", I32, [P, P, P, P, P]),
    ("mixer_corr32_bytes", SZ, [P]),
    ("mixer_corr32_init", I32, [P, P]),
    ("mixer_publish_image", I32, [P, P, SZ]),
    ("mixer_get_stats", I32, [P, P]),
    ("mixer_set_timing", I32, [P, I32]),
    ("mixer_get_timing", I32, [P, I32, P, P]),
    ("mixer_compute_stream", P, [P]),
    ("mixer_batch_default_tuning", None, [P]),
    ("mixer_batch_create", I32, [P, I32, I32, P, P]),
    ("mixer_batch_destroy", I32, [P]),
    ("mixer_batch_run", I32, [P, P]),
    ("mixer_batch_sync", I32, [P]),
    ("mixer_batch_stream", P, [P]),
    ("mixer_batch_get_info", I32, [P, P]),
    ("mixer_batch_dump", I32, [P, I32, P]),
    ("mixer_batch_read_image", I32, [P, I32, P, P, P]),
    ("mixer_batch_acc", I32, [P, I32, P, P]),
    ("mixer_batch_autotune", I32, [P, I32, P]),
    ("mixer_batch_set_tuning", I32, [P, P]),
    ("mixer_batch_get_tuning", I32, [P, P]),
    ("mixer_fill_random", I32, [P, P, SZ, U64]),
    ("mixer_server_create", I32, [I32, U64, I32, P]),
    ("mixer_server_destroy", I32, [P]),
    ("mixer_server_footprint", U64, [P]),
    ("mixer_server_admit", I32, [P, P, P]),
    ("mixer_server_release", I32, [P, I32]),
    ("mixer_server_input", I32, [P, I32, P, P, P]),
    ("mixer_server_submit", I32, [P, I32, P, I32, U64, P]),
    ("mixer_server_token_wait", I32, [P, I32, U64]),
    ("mixer_server_drain", I32, [P]),
    ("mixer_server_set_coalescing", I32, [P, I32]),
    ("mixer_server_get_stats", I32, [P, P]),
    ("mixer_server_reset_latency", None, [P]),
    ("mixer_server_read_image", I32, [P, I32, P, P, P, P]),
    ("mixer_server_dump", I32, [P, I32, P, P]),
    ("mixer_serve_run", I32, [P, P, P]),
    ("mixer_set_device", I32, [I32]),
    ("mixer_dev_alloc", I32, [SZ, P]),
    ("mixer_dev_free", I32, [P]),
    ("mixer_host_alloc", I32, [SZ, P]),
    ("mixer_host_free", I32, [P]),
    ("mixer_memcpy", I32, [P, P, SZ, I32, P]),
    ("mixer_memset", I32, [P, I32, SZ, P]),
    ("mixer_stream_create", I32, [P]),
    ("mixer_stream_destroy", I32, [P]),
    ("mixer_stream_sync", I32, [P]),
    ("mixer_event_create", I32, [P]),
    ("mixer_event_destroy", I32, [P]),
    ("mixer_event_record", I32, [P, P]),
    ("mixer_event_elapsed_ms", I32, [P, P, P]),
    ("mixer_device_sync", I32, []),
    ("mixer_default_sky", None, [P]),
    ("mixer_default_wire", None, [P]),
    ("mixer_build_header", I32, [P, P, C.c_uint32, P]),
    ("mixer_synth_set_wire", I32, [P]),
    ("mixer_orbit_sky", I32, [P, C.c_double, C.c_double, C.c_double, P]),
    ("mixer_synth_packets", I32, [P, P, U64, U64, I32, P, SZ, I32]),
]


def lib():
    """Load libmixer.so (raises if it has not been built: the HIP path has no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libmixer is not built ({LIB_PATH} missing): run `make -C daqiri-mixer_amd`")
    L = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def strerror(status: int) -> str:
    return lib().mixer_strerror(status).decode()


def check(status: int, what: str = "libmixer"):
    if status != OK:
        raise MixerError(status, what)
    return status


def ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def default_cfg(**kw) -> ArrayCfg:
    c = ArrayCfg()
    lib().mixer_default_cfg(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def default_sky(**kw) -> Sky:
    s = Sky()
    lib().mixer_default_sky(C.byref(s))
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def orbit_sky(sky: Sky, rho: float, torbit_s: float, t_s: float) -> Sky:
    """The sky with its planet moved along the transmitter's orbit to time t_s."""
    out = Sky()
    check(lib().mixer_orbit_sky(C.byref(sky), rho, torbit_s, t_s, C.byref(out)), "mixer_orbit_sky")
    return out


def default_wire() -> Wire:
    w = Wire()
    lib().mixer_default_wire(C.byref(w))
    return w


def build_header(cfg, seq: int, wire: Wire | None = None) -> np.ndarray:
    """The transmitter's wire header (eth/ip/udp, seq big-endian at seq_byte)."""
    h = np.zeros(cfg.hdr_bytes, np.uint8)
    w = wire if wire is not None else default_wire()
    check(lib().mixer_build_header(C.byref(cfg), C.byref(w), seq, ptr(h)), "mixer_build_header")
    return h


def device_count() -> int:
    n = C.c_int(0)
    check(lib().mixer_device_count(C.byref(n)))
    return n.value


def synth_packets(cfg: ArrayCfg, sky: Sky, seed: int, seq0: int, npkts: int, pitch: int | None = None,
                  nthreads: int = 0) -> np.ndarray:
    """Host packets [npkts][pitch] in the wire format (header + planar payload)."""
    pitch = cfg.wire if pitch is None else pitch
    out = np.zeros((npkts, pitch), np.uint8)
    check(lib().mixer_synth_packets(C.byref(cfg), C.byref(sky), seed, seq0, npkts, ptr(out), pitch, nthreads),
          "mixer_synth_packets")
    return out


def download_ptr(addr: int, dtype, shape) -> np.ndarray:
    """Copy device memory at a raw device address to a new host array."""
    out = np.empty(shape, dtype)
    check(lib().mixer_memcpy(ptr(out), C.c_void_p(addr), out.nbytes, D2H, None), "mixer_memcpy")
    check(lib().mixer_stream_sync(None), "mixer_stream_sync")
    return out


class DeviceBuffer:
    """Device memory owned by libmixer's HIP runtime."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().mixer_dev_alloc(max(self.nbytes, 1), C.byref(p)), "mixer_dev_alloc")
        self.ptr = p

    @classmethod
    def from_array(cls, a: np.ndarray, stream=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a, stream)
        return b

    def upload(self, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().mixer_memcpy(self.ptr, ptr(a), a.nbytes, H2D, stream), "mixer_memcpy")
        check(lib().mixer_stream_sync(stream), "mixer_stream_sync")

    def download(self, dtype, shape, stream=None) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        check(lib().mixer_stream_sync(stream), "mixer_stream_sync")
        check(lib().mixer_memcpy(ptr(out), self.ptr, out.nbytes, D2H, stream), "mixer_memcpy")
        check(lib().mixer_stream_sync(stream), "mixer_stream_sync")
        return out

    def zero(self, stream=None):
        check(lib().mixer_memset(self.ptr, 0, self.nbytes, stream), "mixer_memset")

    def free(self):
        if self.ptr:
            check(lib().mixer_dev_free(self.ptr), "mixer_dev_free")
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.mixer_dev_free(self.ptr)
        except Exception:
            pass


# ---- operators on device buffers (the kernel seam) ------------------------------------------
def corner_turn(cfg, d_heaps: DeviceBuffer, n_tb: int, d_tiles: DeviceBuffer, d_wrap=None, stream=None):
    check(lib().mixer_corner_turn(stream, C.byref(cfg), d_heaps.ptr, n_tb, d_tiles.ptr,
                                  d_wrap.ptr if d_wrap else None), "mixer_corner_turn")


def xcorr(cfg, d_tiles: DeviceBuffer, n_tb: int, d_acc: DeviceBuffer, accumulate=False, stream=None):
    check(lib().mixer_xcorr(stream, C.byref(cfg), d_tiles.ptr, n_tb, d_acc.ptr, 1 if accumulate else 0),
          "mixer_xcorr")


def xcorr_heaps(cfg, d_heaps: DeviceBuffer, n_tb: int, d_acc: DeviceBuffer, accumulate=False, d_wrap=None,
                stream=None):
    """Fused corner-turn + correlate straight from header-free device heaps."""
    check(lib().mixer_xcorr_heaps(stream, C.byref(cfg), d_heaps.ptr, n_tb, d_acc.ptr, 1 if accumulate else 0,
                                  d_wrap.ptr if d_wrap is not None else None), "mixer_xcorr_heaps")


def integrate(cfg, d_acc: DeviceBuffer, d_v: DeviceBuffer, accumulate=False, stream=None):
    check(lib().mixer_integrate(stream, C.byref(cfg), d_acc.ptr, d_v.ptr, 1 if accumulate else 0),
          "mixer_integrate")


def image(cfg, d_acc: DeviceBuffer, d_img: DeviceBuffer, d_vmax: DeviceBuffer, d_vmin: DeviceBuffer, stream=None):
    check(lib().mixer_image(stream, C.byref(cfg), d_acc.ptr, d_img.ptr, d_vmax.ptr, d_vmin.ptr), "mixer_image")


def image_noalloc(cfg, d_acc: DeviceBuffer, d_img: DeviceBuffer, d_vmax: DeviceBuffer, d_vmin: DeviceBuffer,
                  stream=None):
    """mixer_image without device allocation (static workspace of the library)."""
    check(lib().mixer_image_noalloc(stream, C.byref(cfg), d_acc.ptr, d_img.ptr, d_vmax.ptr, d_vmin.ptr),
          "mixer_image_noalloc")


def xcorr_plain(cfg, d_heaps, n_tb, d_tiles, d_vre, d_vim, stream=None):
    check(lib().mixer_xcorr_plain(stream, C.byref(cfg), d_heaps.ptr, n_tb, d_tiles.ptr, d_vre.ptr, d_vim.ptr),
          "mixer_xcorr_plain")


def set_tuning(n_elem: int, slots=0, prefetch=0, ct_chan=0, staging=0, mfma=0, integrate=0):
    """mfma: 32 = v_mfma_i32_32x32x32_i8 tiles, 16 = v_mfma_i32_16x16x64_i8 (<= 256 elements), 0 = auto.
    integrate: 2 = fused into the blocked correlator's epilogue where one launch covers an
    integration, 1 = the separate integrate kernel, 0 = auto (2)."""
    t = Tuning(slots, prefetch, ct_chan, staging, mfma, integrate)
    check(lib().mixer_set_tuning(n_elem, C.byref(t)), "mixer_set_tuning")


def get_tuning(cfg) -> Tuning:
    t = Tuning()
    check(lib().mixer_get_tuning(C.byref(cfg), C.byref(t)), "mixer_get_tuning")
    return t


def autotune(cfg, device: int = 0) -> Tuning:
    t = Tuning()
    check(lib().mixer_autotune(C.byref(cfg), device, C.byref(t)), "mixer_autotune")
    return t


def tiles_bytes(cfg, n_tb: int) -> int:
    return int(lib().mixer_tiles_bytes(C.byref(cfg), n_tb))


def acc_bytes(cfg) -> int:
    return int(lib().mixer_acc_bytes(C.byref(cfg)))


def acc_slots(cfg) -> int:
    return int(lib().mixer_acc_slots(C.byref(cfg)))


def unpack_acc(acc: np.ndarray, cfg) -> tuple[np.ndarray, np.ndarray]:
    """Tile-packed int32 accumulator [slots][C][tiles][2][32][32] -> full-square (Vre, Vim)
    [C][N][N] in int64: slots summed, upper triangle = conj(lower). Host-side view for tests."""
    N, Cc = cfg.n_elem, cfg.n_chan
    nb = N // 32
    a = acc.reshape(-1, Cc, nb * (nb + 1) // 2, 2, 32, 32).astype(np.int64).sum(axis=0)
    vre = np.zeros((Cc, N, N), np.int64)
    vim = np.zeros((Cc, N, N), np.int64)
    t = 0
    for I in range(nb):
        for J in range(I + 1):
            vre[:, I * 32:(I + 1) * 32, J * 32:(J + 1) * 32] = a[:, t, 0]
            vim[:, I * 32:(I + 1) * 32, J * 32:(J + 1) * 32] = a[:, t, 1]
            if I != J:
                vre[:, J * 32:(J + 1) * 32, I * 32:(I + 1) * 32] = a[:, t, 0].transpose(0, 2, 1)
                vim[:, J * 32:(J + 1) * 32, I * 32:(I + 1) * 32] = -a[:, t, 1].transpose(0, 2, 1)
            t += 1
    return vre, vim


# ---- engine context ---------------------------------------------------------------------------
class Engine:
    """The X-engine context (the reference's gpu_thread + publish loop)."""

    def __init__(self, cfg: ArrayCfg, device: int = 0):
        self.cfg = cfg
        h = C.c_void_p()
        check(lib().mixer_create(C.byref(cfg), device, C.byref(h)), "mixer_create")
        self.h = h

    def close(self):
        if self.h:
            check(lib().mixer_destroy(self.h), "mixer_destroy")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit_heaps(self, pkts: np.ndarray, pitch: int, npkts: int, arrival_ns: int = 0) -> int:
        tok = C.c_uint64(0)
        check(lib().mixer_submit_heaps(self.h, ptr(pkts), pitch, npkts, arrival_ns, C.byref(tok)),
              "mixer_submit_heaps")
        return tok.value

    def submit_device(self, d: DeviceBuffer, n_tb: int, arrival_ns: int = 0, offset: int = 0,
                      token: bool = True) -> int:
        """token=False: no completion token (and no event record on the compute stream); 0."""
        tok = C.c_uint64(0)
        p = C.c_void_p(d.ptr.value + offset)
        check(lib().mixer_submit_device(self.h, p, n_tb, arrival_ns, C.byref(tok) if token else None),
              "mixer_submit_device")
        return tok.value

    def token_wait(self, tok: int):
        check(lib().mixer_token_wait(self.h, tok), "mixer_token_wait")

    def token_done(self, tok: int) -> bool:
        r = lib().mixer_token_done(self.h, tok)
        if r < 0:
            check(r, "mixer_token_done")
        return r == 1

    def set_ingest_slots(self, n_slots: int):
        """Bounded ingest: a time block without a free slot is dropped and counted (imissed)."""
        check(lib().mixer_set_ingest_slots(self.h, n_slots), "mixer_set_ingest_slots")

    def sync(self):
        check(lib().mixer_sync(self.h), "mixer_sync")

    def dump(self, cumulative=False):
        c = self.cfg
        out = np.empty((c.n_chan, 2, c.n_elem, c.n_elem), np.float32)
        idx = C.c_uint64(0)
        check(lib().mixer_dump(self.h, ptr(out), 1 if cumulative else 0, C.byref(idx)), "mixer_dump")
        return out, idx.value

    def read_image(self):
        c = self.cfg
        img = np.empty((c.n_chan, c.ng, c.ng), np.float32)
        vmax = np.empty(c.n_chan, np.float32)
        vmin = np.empty(c.n_chan, np.float32)
        idx = C.c_uint64(0)
        check(lib().mixer_read_image(self.h, ptr(img), ptr(vmax), ptr(vmin), C.byref(idx)), "mixer_read_image")
        return img, vmax, vmin, idx.value

    def read_frame(self):
        """(img, vmax, vmin, n): the image of everything integrated since the previous frame."""
        c = self.cfg
        img = np.empty((c.n_chan, c.ng, c.ng), np.float32)
        vmax = np.empty(c.n_chan, np.float32)
        vmin = np.empty(c.n_chan, np.float32)
        n = C.c_uint64(0)
        check(lib().mixer_read_frame(self.h, ptr(img), ptr(vmax), ptr(vmin), C.byref(n)), "mixer_read_frame")
        return img, vmax, vmin, n.value

    def publish(self, region: np.ndarray):
        check(lib().mixer_publish_image(self.h, ptr(region), region.nbytes), "mixer_publish_image")

    def stats(self) -> Stats:
        s = Stats()
        check(lib().mixer_get_stats(self.h, C.byref(s)), "mixer_get_stats")
        return s

    def set_timing(self, on, stages: int = 0xF):
        """Time every launch of the given stages (bit mask: 1 corner-turn, 2 GEMM, 4 integrate,
        8 image)."""
        check(lib().mixer_set_timing(self.h, stages if on else 0), "mixer_set_timing")

    def timing(self, stage: int):
        ms = C.c_double(0)
        n = C.c_uint64(0)
        check(lib().mixer_get_timing(self.h, stage, C.byref(ms), C.byref(n)), "mixer_get_timing")
        return ms.value, n.value


# ---- batch of arrays ----------------------------------------------------------------------------
def batch_default_tuning() -> BatchTuning:
    t = BatchTuning()
    lib().mixer_batch_default_tuning(C.byref(t))
    return t


def fill_random(d: DeviceBuffer, seed: int, stream=None, nbytes: int | None = None):
    """Fill device memory with int8 values uniform in [-127, 127] (keyed by seed and offset)."""
    check(lib().mixer_fill_random(stream, d.ptr, d.nbytes if nbytes is None else nbytes, seed),
          "mixer_fill_random")


class Batch:
    """Many independent arrays correlated together: one integration per array per run()."""

    def __init__(self, cfgs, device: int = 0, tuning: BatchTuning | None = None):
        self.cfgs = list(cfgs)
        arr = (ArrayCfg * len(self.cfgs))(*self.cfgs)
        h = C.c_void_p()
        check(lib().mixer_batch_create(arr, len(self.cfgs), device,
                                       C.byref(tuning) if tuning is not None else None, C.byref(h)),
              "mixer_batch_create")
        self.h = h

    def close(self):
        if self.h:
            check(lib().mixer_batch_destroy(self.h), "mixer_batch_destroy")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def run(self, heaps):
        """heaps: one device pointer (DeviceBuffer or int address) per array, header-free
        [n_tb][n_elem][payload] layout."""
        ptrs = (C.c_void_p * len(heaps))(*[h.ptr.value if isinstance(h, DeviceBuffer) else int(h)
                                          for h in heaps])
        check(lib().mixer_batch_run(self.h, ptrs), "mixer_batch_run")

    def sync(self):
        check(lib().mixer_batch_sync(self.h), "mixer_batch_sync")

    def stream(self):
        return lib().mixer_batch_stream(self.h)

    def info(self) -> BatchInfo:
        i = BatchInfo()
        check(lib().mixer_batch_get_info(self.h, C.byref(i)), "mixer_batch_get_info")
        return i

    def dump(self, job: int) -> np.ndarray:
        c = self.cfgs[job]
        out = np.empty((c.n_chan, 2, c.n_elem, c.n_elem), np.float32)
        check(lib().mixer_batch_dump(self.h, job, ptr(out)), "mixer_batch_dump")
        return out

    def set_tuning(self, t: BatchTuning):
        check(lib().mixer_batch_set_tuning(self.h, C.byref(t)), "mixer_batch_set_tuning")

    def tuning(self) -> BatchTuning:
        t = BatchTuning()
        check(lib().mixer_batch_get_tuning(self.h, C.byref(t)), "mixer_batch_get_tuning")
        return t

    def acc(self, job: int):
        """(device address, time slots) of job's int32 accumulator of the last run."""
        p, n = C.c_void_p(), C.c_int(0)
        check(lib().mixer_batch_acc(self.h, job, C.byref(p), C.byref(n)), "mixer_batch_acc")
        return p.value, n.value

    def read_image(self, job: int):
        c = self.cfgs[job]
        img = np.empty((c.n_chan, c.ng, c.ng), np.float32)
        vmax = np.empty(c.n_chan, np.float32)
        vmin = np.empty(c.n_chan, np.float32)
        check(lib().mixer_batch_read_image(self.h, job, ptr(img), ptr(vmax), ptr(vmin)),
              "mixer_batch_read_image")
        return img, vmax, vmin

    def autotune(self, reps: int = 3) -> BatchTuning:
        t = BatchTuning()
        check(lib().mixer_batch_autotune(self.h, reps, C.byref(t)), "mixer_batch_autotune")
        return t


# ---- serving loop ---------------------------------------------------------------------------------
def server_footprint(cfg) -> int:
    return int(lib().mixer_server_footprint(C.byref(cfg)))


class Server:
    """Arrays admitted under an HBM budget; chunks correlated as they arrive; per-integration
    images dumped to pinned host memory with arrival-to-dump latency."""

    def __init__(self, hbm_budget: int, device: int = 0, n_streams: int = 0):
        h = C.c_void_p()
        check(lib().mixer_server_create(device, int(hbm_budget), n_streams, C.byref(h)), "mixer_server_create")
        self.h = h
        self.cfgs = {}

    def close(self):
        if self.h:
            check(lib().mixer_server_destroy(self.h), "mixer_server_destroy")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def admit(self, cfg) -> int:
        """Array id; raises MixerError(MIXER_ERR_BUDGET) when it does not fit."""
        i = C.c_int(0)
        check(lib().mixer_server_admit(self.h, C.byref(cfg), C.byref(i)), "mixer_server_admit")
        self.cfgs[i.value] = cfg
        return i.value

    def try_admit(self, cfg):
        i = C.c_int(0)
        st = lib().mixer_server_admit(self.h, C.byref(cfg), C.byref(i))
        if st == ERR_BUDGET:
            return None
        check(st, "mixer_server_admit")
        self.cfgs[i.value] = cfg
        return i.value

    def release(self, aid: int):
        check(lib().mixer_server_release(self.h, aid), "mixer_server_release")
        self.cfgs.pop(aid, None)

    def input(self, aid: int):
        p, n, cb = C.c_void_p(), C.c_int(0), C.c_size_t(0)
        check(lib().mixer_server_input(self.h, aid, C.byref(p), C.byref(n), C.byref(cb)), "mixer_server_input")
        return p.value, n.value, cb.value

    def submit(self, aid: int, d_ptr, n_tb: int, arrival_ns: int = 0) -> int:
        tok = C.c_uint64(0)
        addr = d_ptr.ptr.value if isinstance(d_ptr, DeviceBuffer) else int(d_ptr)
        check(lib().mixer_server_submit(self.h, aid, C.c_void_p(addr), n_tb, arrival_ns, C.byref(tok)),
              "mixer_server_submit")
        return tok.value

    def token_wait(self, aid: int, tok: int):
        check(lib().mixer_server_token_wait(self.h, aid, tok), "mixer_server_token_wait")

    def drain(self):
        check(lib().mixer_server_drain(self.h), "mixer_server_drain")

    def set_coalescing(self, on: bool):
        """Hold chunks that arrive while their array's previous launch runs and correlate the
        contiguous ones in one launch (default on)."""
        check(lib().mixer_server_set_coalescing(self.h, 1 if on else 0), "mixer_server_set_coalescing")

    def stats(self) -> ServerStats:
        s = ServerStats()
        check(lib().mixer_server_get_stats(self.h, C.byref(s)), "mixer_server_get_stats")
        return s

    def reset_latency(self):
        lib().mixer_server_reset_latency(self.h)

    def read_image(self, aid: int):
        c = self.cfgs[aid]
        img = np.empty((c.n_chan, c.ng, c.ng), np.float32)
        vmax = np.empty(c.n_chan, np.float32)
        vmin = np.empty(c.n_chan, np.float32)
        idx = C.c_uint64(0)
        check(lib().mixer_server_read_image(self.h, aid, ptr(img), ptr(vmax), ptr(vmin), C.byref(idx)),
              "mixer_server_read_image")
        return img, vmax, vmin, idx.value

    def dump(self, aid: int):
        c = self.cfgs[aid]
        out = np.empty((c.n_chan, 2, c.n_elem, c.n_elem), np.float32)
        idx = C.c_uint64(0)
        check(lib().mixer_server_dump(self.h, aid, ptr(out), C.byref(idx)), "mixer_server_dump")
        return out, idx.value

    def run(self, duration_s: float, aggregate_gbps: float, seed: int = 0x4D495835) -> ServeReport:
        spec = ServeSpec(duration_s, aggregate_gbps, seed)
        rep = ServeReport()
        check(lib().mixer_serve_run(self.h, C.byref(spec), C.byref(rep)), "mixer_serve_run")
        return rep


def corr32_bytes(cfg) -> int:
    return int(lib().mixer_corr32_bytes(C.byref(cfg)))
