"""ctypes binding of libvfrb.so (the C ABI declared in include/vfrb.h).

Used by the Python tests, the benchmark and the smoke entry point.  The
library is loaded from ``_build/libvfrb.so`` next to this file; importing
this module fails loudly when it has not been built, and every call raises
VfrbError on a non-zero status (there is no fallback path).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libvfrb.so")

STORE_F32, STORE_F16, STORE_U8 = 0, 1, 2
ACC_F64, ACC_F32C = 0, 1
STORAGE_NAMES = {"f32": STORE_F32, "f16": STORE_F16, "u8": STORE_U8}
(OK, ERR_INVALID, ERR_UNSUPPORTED, ERR_NOMEM, ERR_HIP, ERR_NODEVICE, ERR_IO, ERR_FORMAT, ERR_EOF,
 ERR_CAPACITY) = range(10)


class VfrbError(RuntimeError):
    def __init__(self, func: str, status: int, message: str):
        super().__init__(f"{func} failed with status {status}: {message}")
        self.status = status


if not os.path.exists(LIB_PATH):
    raise ImportError(f"libvfrb.so not built at {LIB_PATH}; run `make -C vlite-frb_amd` "
                      "or __graft_entry__.build()")
lib = C.CDLL(LIB_PATH)

vp = C.c_void_p
i32, i64, f32, f64, u64 = C.c_int32, C.c_int64, C.c_float, C.c_double, C.c_uint64
sz = C.c_size_t
P = C.POINTER


class Params(C.Structure):
    _fields_ = [(n, i32) for n in ("twonchan", "nchan", "nfreq", "nave", "nbundle", "nloop",
                                   "nchanout", "nskip", "maxnt", "nhalfcyc")] + \
               [(n, f64) for n in ("chanbw", "foff", "fch1", "tchan", "tbundle", "maxtime")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SearchDesc(C.Structure):
    _fields_ = [("nchans", i32), ("tsamp", f64), ("fch1", f64), ("chanbw", f64), ("ndm", i32),
                ("dms", P(f32)), ("max_rows", i32), ("storage", i32), ("tile_config", i32)]


class VdifHeader(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "seconds", "legacymode", "invalid", "frame", "epoch", "unassigned", "framelength8", "nchan",
        "version", "stationid", "threadid", "nbits", "iscomplex", "eversion", "extended1",
        "extended2", "extended3", "extended4")]


MAXP, MAXT = 64, 64


class SynthSpec(C.Structure):
    _fields_ = [("seed", u64), ("nfft", i32), ("mean", f64), ("sigma", f64), ("f_top_mhz", f64),
                ("bw_mhz", f64), ("npulse", i32), ("pulse_dm", f64 * MAXP), ("pulse_t0", f64 * MAXP),
                ("pulse_fwhm", f64 * MAXP), ("pulse_amp", f64 * MAXP), ("ntone", i32),
                ("tone_bin", f64 * MAXT), ("tone_amp", f64 * MAXT), ("nthreads", i32)]


class ServerDesc(C.Structure):
    _fields_ = [("twonchan", i32), ("nave", i32), ("nchans", i32), ("nskip", i32), ("fch1", f64),
                ("chanbw", f64), ("tsamp", f64), ("storage", i32), ("acc", i32), ("threshold", f32),
                ("max_cand", i32), ("max_streams", i32), ("max_segment_bytes", i64),
                ("arena_bytes", i64), ("max_batch", i32), ("batched", i32), ("tile_config", i32)]


class ServerOutput(C.Structure):
    _fields_ = [("peak", P(f32)), ("peak_t", P(i32)), ("ave", P(f32)), ("sigma", P(f32)),
                ("cand_dm", P(i32)), ("cand_t", P(i32)), ("cand_snr", P(f32))]


class ServerResult(C.Structure):
    _fields_ = [("tag", i64), ("stream_id", i32), ("status", i32), ("rows", i64), ("sample0", i64),
                ("nt", i32), ("ncand", i32), ("best_dm", i32), ("best_t", i32), ("best_snr", f32),
                ("batch_size", i32), ("arrival_s", f64), ("start_s", f64), ("done_s", f64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", i64), ("completed", i64), ("iterations", i64),
                ("batched_segments", i64), ("arena_peak_bytes", i64), ("device_busy_s", f64)]


LOG_FN = C.CFUNCTYPE(None, C.c_char_p, vp)

_SIGS = {
    "vfrb_strerror": (C.c_char_p, [i32]),
    "vfrb_last_error": (C.c_char_p, []),
    "vfrb_version": (C.c_char_p, []),
    "vfrb_params_derive": (i32, [i32, i32, i32, f64, i32, i32, P(Params)]),
    "vfrb_delay_table": (i32, [P(f32), i32, i32, f64, f64, f64, P(i32), P(i32), P(f64)]),
    "vfrb_device_count": (i32, [P(i32)]),
    "vfrb_set_device": (i32, [i32]),
    "vfrb_device_name": (i32, [C.c_char_p, i32]),
    "vfrb_mem_info": (i32, [P(sz), P(sz)]),
    "vfrb_malloc": (i32, [P(vp), sz]),
    "vfrb_free": (i32, [vp]),
    "vfrb_host_alloc": (i32, [P(vp), sz]),
    "vfrb_host_free": (i32, [vp]),
    "vfrb_memcpy_h2d": (i32, [vp, vp, sz, vp]),
    "vfrb_memcpy_d2h": (i32, [vp, vp, sz, vp]),
    "vfrb_memcpy_d2d": (i32, [vp, vp, sz, vp]),
    "vfrb_memset": (i32, [vp, i32, sz, vp]),
    "vfrb_stream_create": (i32, [P(vp)]),
    "vfrb_stream_destroy": (i32, [vp]),
    "vfrb_stream_synchronize": (i32, [vp]),
    "vfrb_device_synchronize": (i32, []),
    "vfrb_event_create": (i32, [P(vp)]),
    "vfrb_event_destroy": (i32, [vp]),
    "vfrb_event_record": (i32, [vp, vp]),
    "vfrb_event_synchronize": (i32, [vp]),
    "vfrb_event_elapsed_ms": (i32, [P(f32), vp, vp]),
    "vfrb_stream_wait_event": (i32, [vp, vp]),
    "vfrb_graph_begin": (i32, [vp]),
    "vfrb_graph_end": (i32, [vp, P(vp)]),
    "vfrb_graph_launch": (i32, [vp, vp]),
    "vfrb_graph_destroy": (i32, [vp]),
    "vfrb_fb_plan_create": (i32, [P(vp), i32, i32, i32, i32]),
    "vfrb_fb_plan_destroy": (i32, [vp]),
    "vfrb_fb_execute": (i32, [vp, vp, vp, i64, vp, i64, i64, vp]),
    "vfrb_fb_execute_frames": (i32, [vp, vp, vp, i64, i64, i64, vp, i64, i64, vp]),
    "vfrb_ingest_create": (i32, [P(vp), i64, i32]),
    "vfrb_ingest_destroy": (i32, [vp]),
    "vfrb_ingest_reset": (i32, [vp, i32, i32]),
    "vfrb_ingest_frames": (i32, [vp, vp, i64, vp, i64, P(i64), vp]),
    "vfrb_ingest_gather": (i32, [vp, vp, i64, i32, vp, vp]),
    "vfrb_ingest_frames_noalloc": (i32, [vp, i64, i32, P(i32), P(i32), vp, i64, P(i64), vp]),
    "vfrb_bandpass_inplace": (i32, [vp, i32, i32, vp]),
    "vfrb_chan_stats": (i32, [vp, i32, i32, i32, i32, vp, vp, vp, vp]),
    "vfrb_bandpass_chan_stats": (i32, [vp, i32, i32, i32, i32, vp, vp, vp, vp]),
    "vfrb_search_plan_create": (i32, [P(vp), P(SearchDesc)]),
    "vfrb_search_plan_destroy": (i32, [vp]),
    "vfrb_search_max_delay": (i32, [vp]),
    "vfrb_search_dm_count": (i32, [vp]),
    "vfrb_search_get_offsets": (i32, [vp, P(i32)]),
    "vfrb_search_load": (i32, [vp, vp, i64, i64, i32, i32, vp]),
    "vfrb_search_dedisperse": (i32, [vp, i32, vp, i64, vp]),
    "vfrb_search_execute": (i32, [vp, i32, vp, vp, i64, vp]),
    "vfrb_search_dedisperse_stats": (i32, [vp, i32, vp, vp, i64, vp]),
    "vfrb_search_tile_count": (i32, []),
    "vfrb_search_tile_name": (C.c_char_p, [i32]),
    "vfrb_search_set_tile": (i32, [vp, i32]),
    "vfrb_search_get_tile": (i32, [vp]),
    "vfrb_search_autotune": (i32, [vp, i32, vp, i64, i32, P(f32), vp]),
    "vfrb_batch_create": (i32, [P(vp), i32, i32, i32, i32]),
    "vfrb_batch_destroy": (i32, [vp]),
    "vfrb_batch_set": (i32, [vp, i32, P(vp), P(i32), P(vp), P(i64), P(i32), vp]),
    "vfrb_batch_dedisperse": (i32, [vp, vp]),
    "vfrb_batch_graph_launch": (i32, [vp, vp]),
    "vfrb_batch_autotune": (i32, [vp, i32, P(f32), vp]),
    "vfrb_batch_get_tile": (i32, [vp]),
    "vfrb_batch_grid": (i32, [vp]),
    "vfrb_server_create": (i32, [P(vp), P(ServerDesc)]),
    "vfrb_server_destroy": (i32, [vp]),
    "vfrb_server_open_stream": (i32, [vp, P(f32), i32, P(i32)]),
    "vfrb_server_max_delay": (i32, [vp, i32]),
    "vfrb_server_submit": (i32, [vp, i32, vp, vp, i64, i64, P(ServerOutput)]),
    "vfrb_server_submit_device": (i32, [vp, i32, vp, vp, i64, i64, P(ServerOutput)]),
    "vfrb_server_poll": (i32, [vp, P(ServerResult), i32, i32, P(i32)]),
    "vfrb_server_drain": (i32, [vp]),
    "vfrb_server_get_stats": (i32, [vp, P(ServerStats)]),
    "vfrb_server_now": (f64, [vp]),
    "vfrb_server_save_stream": (i32, [vp, i32, vp, i64, P(i64)]),
    "vfrb_server_restore_stream": (i32, [vp, i32, vp, i64]),
    "vfrb_dm_stats": (i32, [vp, i64, i32, i32, i32, vp, vp, vp, vp, vp, vp]),
    "vfrb_candidates": (i32, [vp, i64, i32, i32, vp, vp, f32, vp, vp, vp, i32, vp, vp]),
    "vfrb_vdif_decode": (i32, [P(C.c_uint8), P(VdifHeader)]),
    "vfrb_vdif_encode": (i32, [P(VdifHeader), P(C.c_uint8)]),
    "vfrb_vdif_dmjd": (f64, [P(C.c_uint8), i32]),
    "vfrb_seq_open_file": (i32, [P(vp), C.c_char_p]),
    "vfrb_seq_open_memory": (i32, [P(vp), P(C.c_uint8), sz]),
    "vfrb_seq_set_log": (i32, [vp, LOG_FN, vp]),
    "vfrb_seq_set_seconds_check": (i32, [vp, i32]),
    "vfrb_seq_skip_pairs": (i32, [vp, i64]),
    "vfrb_seq_init": (i32, [vp]),
    "vfrb_seq_header": (i32, [vp, P(C.c_uint8)]),
    "vfrb_seq_fill": (i32, [vp, i32, P(C.c_uint8), P(C.c_uint8)]),
    "vfrb_seq_counters": (i32, [vp, P(i64), P(i64)]),
    "vfrb_seq_consumed": (i32, [vp, P(i64)]),
    "vfrb_seq_close": (i32, [vp]),
    "vfrb_sigproc_header": (i32, [P(C.c_uint8), i32, P(i32), f64, f64, i32, i32, f64, f64]),
    "vfrb_presto_inf": (i32, [C.c_char_p, i32, C.c_char_p, f64, i32, f64, f32, f64, f64, i32, f64]),
    "vfrb_synth_payload": (i32, [P(SynthSpec), i32, i64, i64, P(C.c_uint8)]),
    "vfrb_synth_amp_for_snr": (f64, [f64, i32, i32]),
}
for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = sorted(_SIGS)


def check(status: int, func: str = "vfrb"):
    if status != 0:
        msg = lib.vfrb_last_error().decode(errors="replace")
        raise VfrbError(func, status, msg or lib.vfrb_strerror(status).decode())


def call(name, *args):
    check(getattr(lib, name)(*args), name)


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(P(ctype))


# --------------------------------------------------------------------------
# Host-side helpers
# --------------------------------------------------------------------------
def params_derive(twonchan, nave, nbundle, maxtime, fullband=False, nchanout_override=0) -> Params:
    p = Params()
    call("vfrb_params_derive", twonchan, nave, nbundle, float(maxtime), int(bool(fullband)),
         int(nchanout_override), C.byref(p))
    return p


def delay_table(dms, nchanout, fch1, chanbw, tchan):
    dms = np.ascontiguousarray(dms, dtype=np.float32)
    off = np.empty((len(dms), nchanout), dtype=np.int32)
    maxd = i32()
    margin = f64()
    call("vfrb_delay_table", _ptr(dms, f32), len(dms), nchanout, fch1, chanbw, tchan,
         _ptr(off, i32), C.byref(maxd), C.byref(margin))
    return off, maxd.value, margin.value


def vdif_decode(raw: bytes) -> dict:
    a = np.frombuffer(raw[:32], dtype=np.uint8).copy()
    h = VdifHeader()
    call("vfrb_vdif_decode", _ptr(a, C.c_uint8), C.byref(h))
    return {k: getattr(h, k) for k, _ in h._fields_}


def vdif_encode(**fields) -> bytes:
    h = VdifHeader()
    for k, v in fields.items():
        setattr(h, k, v)
    out = np.zeros(32, dtype=np.uint8)
    call("vfrb_vdif_encode", C.byref(h), _ptr(out, C.c_uint8))
    return out.tobytes()


def vdif_dmjd(raw: bytes, fps=25600) -> float:
    a = np.frombuffer(raw[:32], dtype=np.uint8).copy()
    return lib.vfrb_vdif_dmjd(_ptr(a, C.c_uint8), fps)


def sigproc_header(fch1, foff, nchans, nbits, tstart, tsamp) -> bytes:
    buf = np.zeros(1024, dtype=np.uint8)
    n = i32()
    call("vfrb_sigproc_header", _ptr(buf, C.c_uint8), 1024, C.byref(n), fch1, foff, nchans, nbits,
         tstart, tsamp)
    return buf[: n.value].tobytes()


def presto_inf(base, mjdstart, ntimes, tsamp, dm, fchanlo, bw, nchan, chanbw) -> str:
    buf = C.create_string_buffer(4096)
    call("vfrb_presto_inf", buf, 4096, base.encode(), mjdstart, ntimes, tsamp, dm, fchanlo, bw,
         nchan, chanbw)
    return buf.value.decode()


class Sequencer:
    """Frame sequencer over an in-memory VDIF byte string."""

    def __init__(self, data: bytes | None = None, path: str | None = None, seconds_check=False):
        self.h = vp()
        self.log_lines: list[str] = []
        if path is not None:
            call("vfrb_seq_open_file", C.byref(self.h), path.encode())
        else:
            self._buf = np.frombuffer(data, dtype=np.uint8).copy()
            call("vfrb_seq_open_memory", C.byref(self.h), _ptr(self._buf, C.c_uint8), self._buf.size)
        self._cb = LOG_FN(lambda line, user: self.log_lines.append(line.decode()))
        call("vfrb_seq_set_log", self.h, self._cb, None)
        call("vfrb_seq_set_seconds_check", self.h, int(bool(seconds_check)))

    def init(self) -> int:
        return lib.vfrb_seq_init(self.h)

    def fill(self, nloop):
        a = np.empty(nloop, dtype=np.uint8)
        b = np.empty(nloop, dtype=np.uint8)
        st = lib.vfrb_seq_fill(self.h, nloop, _ptr(a, C.c_uint8), _ptr(b, C.c_uint8))
        return st, a, b

    def header(self) -> bytes:
        a = np.zeros(32, dtype=np.uint8)
        call("vfrb_seq_header", self.h, _ptr(a, C.c_uint8))
        return a.tobytes()

    def counters(self):
        r, f = i64(), i64()
        call("vfrb_seq_counters", self.h, C.byref(r), C.byref(f))
        return r.value, f.value

    def close(self):
        if self.h:
            lib.vfrb_seq_close(self.h)
            self.h = vp()

    def __del__(self):
        self.close()


def synth_spec(nfft, seed=0, pulses=(), tones=(), mean=128.0, sigma=8.0, nthreads=0) -> SynthSpec:
    """pulses: (dm, t0, fwhm, amp); tones: (bin, amp)."""
    s = SynthSpec()
    s.seed, s.nfft, s.mean, s.sigma = seed, nfft, mean, sigma
    s.f_top_mhz, s.bw_mhz, s.nthreads = 384.0, 64.0, nthreads
    s.npulse = len(pulses)
    for i, (dm, t0, w, a) in enumerate(pulses):
        s.pulse_dm[i], s.pulse_t0[i], s.pulse_fwhm[i], s.pulse_amp[i] = dm, t0, w, a
    s.ntone = len(tones)
    for i, (b, a) in enumerate(tones):
        s.tone_bin[i], s.tone_amp[i] = b, a
    return s


def synth_payload(spec: SynthSpec, pol: int, block0: int, nblocks: int, out=None) -> np.ndarray:
    if out is None:
        out = np.empty(nblocks * spec.nfft, dtype=np.uint8)
    call("vfrb_synth_payload", C.byref(spec), pol, block0, nblocks, _ptr(out, C.c_uint8))
    return out


def amp_for_snr(snr, nave, nchan) -> float:
    return lib.vfrb_synth_amp_for_snr(snr, nave, nchan)


# --------------------------------------------------------------------------
# Device helpers
# --------------------------------------------------------------------------
def device_count() -> int:
    n = i32()
    call("vfrb_device_count", C.byref(n))
    return n.value


def device_name() -> str:
    buf = C.create_string_buffer(256)
    call("vfrb_device_name", buf, 256)
    return buf.value.decode()


def mem_info():
    f, t = sz(), sz()
    call("vfrb_mem_info", C.byref(f), C.byref(t))
    return f.value, t.value


class DeviceArray:
    """A device allocation with a numpy shape/dtype for host round trips."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = vp()
        call("vfrb_malloc", C.byref(self.ptr), max(self.nbytes, 1))

    @classmethod
    def from_host(cls, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a, stream)
        return d

    def upload(self, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        call("vfrb_memcpy_h2d", self.ptr, a.ctypes.data_as(vp), a.nbytes, stream)
        if stream:
            call("vfrb_stream_synchronize", stream)

    def download(self, stream=None) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        call("vfrb_memcpy_d2h", out.ctypes.data_as(vp), self.ptr, self.nbytes, stream)
        if stream:
            call("vfrb_stream_synchronize", stream)
        return out

    def zero(self, stream=None):
        call("vfrb_memset", self.ptr, 0, self.nbytes, stream)

    def offset(self, nbytes: int) -> vp:
        return vp(self.ptr.value + int(nbytes))

    def free(self):
        if self.ptr:
            lib.vfrb_free(self.ptr)
            self.ptr = vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.h = vp()
        call("vfrb_stream_create", C.byref(self.h))

    def sync(self):
        call("vfrb_stream_synchronize", self.h)

    def wait(self, event: "Event"):
        call("vfrb_stream_wait_event", self.h, event.h)

    def __del__(self):
        try:
            if self.h:
                lib.vfrb_stream_destroy(self.h)
        except Exception:
            pass


class Event:
    def __init__(self):
        self.h = vp()
        call("vfrb_event_create", C.byref(self.h))

    def record(self, stream=None):
        call("vfrb_event_record", self.h, stream.h if isinstance(stream, Stream) else stream)

    def sync(self):
        call("vfrb_event_synchronize", self.h)

    def elapsed_ms(self, later: "Event") -> float:
        ms = f32()
        call("vfrb_event_elapsed_ms", C.byref(ms), self.h, later.h)
        return ms.value

    def __del__(self):
        try:
            if self.h:
                lib.vfrb_event_destroy(self.h)
        except Exception:
            pass


class Graph:
    """The library calls enqueued on `stream` between begin() and end(),
    captured into a hipGraph and replayed by launch()."""

    def __init__(self, stream: Stream):
        self.stream, self.h = stream, vp()

    def begin(self):
        call("vfrb_graph_begin", self.stream.h)

    def end(self):
        call("vfrb_graph_end", self.stream.h, C.byref(self.h))

    def launch(self, stream=None):
        call("vfrb_graph_launch", self.h, _s(stream) if stream is not None else self.stream.h)

    def __del__(self):
        try:
            if self.h:
                lib.vfrb_graph_destroy(self.h)
        except Exception:
            pass


def sync():
    call("vfrb_device_synchronize")


def _s(stream):
    return stream.h if isinstance(stream, Stream) else stream


class FilterbankPlan:
    def __init__(self, twonchan, nave, nskip, nchanout):
        self.h = vp()
        self.twonchan, self.nave, self.nskip, self.nchanout = twonchan, nave, nskip, nchanout
        call("vfrb_fb_plan_create", C.byref(self.h), twonchan, nave, nskip, nchanout)

    def execute(self, pol0: DeviceArray, pol1: DeviceArray, nrows, fb: DeviceArray, ring_rows=0,
                row0=0, stream=None, pol_offset=0):
        call("vfrb_fb_execute", self.h, pol0.offset(pol_offset), pol1.offset(pol_offset), nrows,
             fb.ptr, ring_rows, row0, _s(stream))

    def execute_frames(self, frames: DeviceArray, table: DeviceArray, nslots, byte0, nrows, fb: DeviceArray,
                       ring_rows=0, row0=0, stream=None):
        """Filterbank straight from raw frame pairs through an ingest slot table."""
        call("vfrb_fb_execute_frames", self.h, frames.ptr, table.ptr, nslots, byte0, nrows, fb.ptr, ring_rows,
             row0, _s(stream))

    def __del__(self):
        try:
            if self.h:
                lib.vfrb_fb_plan_destroy(self.h)
        except Exception:
            pass


class Ingest:
    """Device-side VDIF frame ingest (vfrb_ingest_*): raw frame pairs in
    device memory -> payload slot table with the reference's zero fill."""

    def __init__(self, max_pairs, seconds_check=False):
        self.h = vp()
        call("vfrb_ingest_create", C.byref(self.h), max_pairs, int(bool(seconds_check)))

    def reset(self, expected_frame=-1, expected_second=0):
        call("vfrb_ingest_reset", self.h, expected_frame, expected_second)

    def frames(self, frames: DeviceArray, npairs, table: DeviceArray, stream=None, offset=0):
        """Builds the table of npairs pairs starting `offset` bytes into
        `frames`; returns the payload slots they produce."""
        n = i64(0)
        call("vfrb_ingest_frames", self.h, frames.offset(offset), npairs, table.ptr, table.nbytes // 4,
             C.byref(n), _s(stream))
        return n.value

    @staticmethod
    def gather(frames: DeviceArray, table: DeviceArray, nslots, pol, out: DeviceArray, stream=None, offset=0):
        call("vfrb_ingest_gather", frames.offset(offset), table.ptr, nslots, pol, out.ptr, _s(stream))

    def __del__(self):
        try:
            if self.h:
                lib.vfrb_ingest_destroy(self.h)
        except Exception:
            pass


def ingest_frames_noalloc(frames: DeviceArray, npairs, state, table: DeviceArray, seconds_check=False,
                          stream=None, offset=0):
    """vfrb_ingest_frames_noalloc: the ingest with no allocation; `state` is
    the caller's [expected_frame, expected_second] list, updated in place on
    success.  Returns the payload slots."""
    ef, es, n = i32(state[0]), i32(state[1]), i64(0)
    call("vfrb_ingest_frames_noalloc", frames.offset(offset), npairs, int(bool(seconds_check)), C.byref(ef),
         C.byref(es), table.ptr, table.nbytes // 4, C.byref(n), _s(stream))
    state[0], state[1] = ef.value, es.value
    return n.value


class SearchPlan:
    def __init__(self, nchans, tsamp, fch1, chanbw, dms, max_rows, storage=STORE_F32, tile=-1):
        self._dms = np.ascontiguousarray(dms, dtype=np.float32)
        d = SearchDesc(nchans, tsamp, fch1, chanbw, len(self._dms), _ptr(self._dms, f32), max_rows,
                       storage, tile)
        self.h = vp()
        self.nchans, self.ndm, self.storage = nchans, len(self._dms), storage
        call("vfrb_search_plan_create", C.byref(self.h), C.byref(d))
        self.max_delay = lib.vfrb_search_max_delay(self.h)

    def offsets(self) -> np.ndarray:
        off = np.empty((self.ndm, self.nchans), dtype=np.int32)
        call("vfrb_search_get_offsets", self.h, _ptr(off, i32))
        return off

    def load(self, fb: DeviceArray, nrows, ring_rows=0, row0=0, bandpass=True, stream=None):
        call("vfrb_search_load", self.h, fb.ptr, ring_rows, row0, nrows, int(bool(bandpass)), _s(stream))

    def dedisperse(self, nrows, out: DeviceArray, ld_out, stream=None):
        call("vfrb_search_dedisperse", self.h, nrows, out.ptr, ld_out, _s(stream))

    def execute(self, nrows, fb_tm: DeviceArray, out: DeviceArray, ld_out, stream=None):
        call("vfrb_search_execute", self.h, nrows, fb_tm.ptr, out.ptr, ld_out, _s(stream))

    def dedisperse_stats(self, nrows, out: DeviceArray, sig: DeviceArray, ld_out, stream=None):
        call("vfrb_search_dedisperse_stats", self.h, nrows, out.ptr, sig.ptr, ld_out, _s(stream))

    def set_tile(self, i):
        call("vfrb_search_set_tile", self.h, i)

    def autotune(self, nrows, out: DeviceArray, ld_out, iters=2, stream=None):
        """Times every tile configuration on the loaded window, keeps the
        fastest; returns {tile name: ms} (-1 = does not fit)."""
        n = lib.vfrb_search_tile_count()
        ms = np.zeros(n, np.float32)
        call("vfrb_search_autotune", self.h, nrows, out.ptr, ld_out, iters, _ptr(ms, f32), _s(stream))
        return dict(zip(tile_names(), ms.tolist()))

    @property
    def tile(self):
        return lib.vfrb_search_get_tile(self.h)

    def __del__(self):
        try:
            if self.h:
                lib.vfrb_search_plan_destroy(self.h)
        except Exception:
            pass


class Batch:
    """Persistent batched dedispersion over many search plans
    (vfrb_batch_*): set() the jobs, then dedisperse() eagerly or
    graph_launch() a captured hipGraph replay."""

    def __init__(self, storage=STORE_F32, tile=-1, max_jobs=64, max_tiles=1 << 20):
        self.h = vp()
        call("vfrb_batch_create", C.byref(self.h), storage, tile, max_jobs, max_tiles)
        self._keep = ()

    def set(self, jobs, stream=None):
        """jobs: sequence of (SearchPlan, nrows, out DeviceArray, ld_out)."""
        n = len(jobs)
        plans = (vp * max(n, 1))(*[j[0].h for j in jobs])
        nrows = np.array([j[1] for j in jobs] or [0], dtype=np.int32)
        outs = (vp * max(n, 1))(*[j[2].ptr for j in jobs])
        lds = np.array([j[3] for j in jobs] or [0], dtype=np.int64)
        nt = i32(0)
        call("vfrb_batch_set", self.h, n, plans, _ptr(nrows, i32), outs, _ptr(lds, i64),
             C.byref(nt), _s(stream))
        self._keep = tuple(jobs)  # plans and outputs must outlive the work list
        return nt.value

    def dedisperse(self, stream=None):
        call("vfrb_batch_dedisperse", self.h, _s(stream))

    def graph_launch(self, stream):
        call("vfrb_batch_graph_launch", self.h, _s(stream))

    def autotune(self, iters=2, stream=None):
        n = lib.vfrb_search_tile_count()
        ms = np.zeros(n, np.float32)
        call("vfrb_batch_autotune", self.h, iters, _ptr(ms, f32), _s(stream))
        return dict(zip(tile_names(), ms.tolist()))

    @property
    def tile(self):
        return lib.vfrb_batch_get_tile(self.h)

    @property
    def grid(self):
        return lib.vfrb_batch_grid(self.h)

    def __del__(self):
        try:
            if self.h:
                lib.vfrb_batch_destroy(self.h)
        except Exception:
            pass


class PinnedBytes:
    """Pinned host buffer (vfrb_host_alloc) exposed as a numpy uint8 array."""

    def __init__(self, nbytes):
        self.ptr = vp()
        call("vfrb_host_alloc", C.byref(self.ptr), max(int(nbytes), 1))
        self.array = np.ctypeslib.as_array(C.cast(self.ptr, P(C.c_uint8)), shape=(int(nbytes),))

    def __del__(self):
        try:
            if self.ptr:
                lib.vfrb_host_free(self.ptr)
        except Exception:
            pass


class Server:
    """Segment server (vfrb_server_*).  submit() returns at once; poll()
    hands back finished segments as dicts with their per-DM results and
    candidates (numpy arrays owned by the result)."""

    def __init__(self, twonchan, nave, nchans, nskip, fch1, chanbw, tsamp, storage=STORE_F16,
                 acc=ACC_F32C, threshold=7.0, max_cand=4096, max_streams=16,
                 max_segment_bytes=1 << 28, arena_bytes=1 << 32, max_batch=16, batched=True,
                 tile=-1):
        d = ServerDesc(twonchan, nave, nchans, nskip, fch1, chanbw, tsamp, storage, acc, threshold,
                       max_cand, max_streams, max_segment_bytes, arena_bytes, max_batch,
                       int(bool(batched)), tile)
        self.max_cand = max_cand
        self.h = vp()
        call("vfrb_server_create", C.byref(self.h), C.byref(d))
        self._ndm = {}
        self._live = {}  # tag -> (payloads, outputs) kept until polled

    def open_stream(self, dms):
        dms = np.ascontiguousarray(dms, dtype=np.float32)
        sid = i32(0)
        call("vfrb_server_open_stream", self.h, _ptr(dms, f32), len(dms), C.byref(sid))
        self._ndm[sid.value] = len(dms)
        return sid.value

    def max_delay(self, sid):
        return lib.vfrb_server_max_delay(self.h, sid)

    def submit(self, sid, pol0, pol1, tag, outputs=True):
        """pol0/pol1: contiguous uint8 arrays of the same length (views of
        PinnedBytes memory copy fastest); referenced until polled."""
        assert pol0.dtype == np.uint8 and pol1.dtype == np.uint8 and len(pol0) == len(pol1)
        assert pol0.flags.c_contiguous and pol1.flags.c_contiguous
        self._submit("vfrb_server_submit", sid, pol0.ctypes.data, pol1.ctypes.data, len(pol0),
                     (pol0, pol1), tag, outputs)

    def submit_device(self, sid, dev0: "DeviceArray", dev1: "DeviceArray", offset, nbytes, tag,
                      outputs=True):
        """Payload already in device memory: bytes [offset, offset+nbytes) of
        dev0/dev1 (kept alive until polled)."""
        assert offset >= 0 and offset + nbytes <= min(dev0.nbytes, dev1.nbytes)
        self._submit("vfrb_server_submit_device", sid, dev0.ptr.value + offset,
                     dev1.ptr.value + offset, nbytes, (dev0, dev1), tag, outputs)

    def _submit(self, fn, sid, a0, a1, n, keep, tag, outputs):
        k0, k1 = keep
        ndm = self._ndm.get(sid, 1)  # unknown ids are rejected by the library
        outs = None
        o = None
        if outputs:
            outs = {"peak": np.zeros(ndm, np.float32), "peak_t": np.zeros(ndm, np.int32),
                    "ave": np.zeros(ndm, np.float32), "sigma": np.zeros(ndm, np.float32),
                    "cand_dm": np.zeros(self.max_cand, np.int32),
                    "cand_t": np.zeros(self.max_cand, np.int32),
                    "cand_snr": np.zeros(self.max_cand, np.float32)}
            o = ServerOutput(*[_ptr(outs[k], f32 if outs[k].dtype == np.float32 else i32)
                               for k in ("peak", "peak_t", "ave", "sigma", "cand_dm", "cand_t", "cand_snr")])
        self._live[tag] = (k0, k1, outs, o)
        try:
            call(fn, self.h, sid, vp(a0), vp(a1), n, tag, C.byref(o) if o is not None else None)
        except VfrbError:
            del self._live[tag]
            raise

    def poll(self, wait_ms=-1, max_results=64):
        res = (ServerResult * max_results)()
        n = i32(0)
        call("vfrb_server_poll", self.h, res, max_results, wait_ms, C.byref(n))
        out = []
        for r in res[: n.value]:
            d = {k: getattr(r, k) for k, _ in ServerResult._fields_}
            _, _, outs, _ = self._live.pop(r.tag)
            if outs is not None:
                k = min(r.ncand, self.max_cand)
                d.update({"peak": outs["peak"], "peak_t": outs["peak_t"], "ave": outs["ave"],
                          "sigma": outs["sigma"], "cand_dm": outs["cand_dm"][:k],
                          "cand_t": outs["cand_t"][:k], "cand_snr": outs["cand_snr"][:k]})
            out.append(d)
        return out

    def drain(self):
        call("vfrb_server_drain", self.h)

    def save_stream(self, sid) -> bytes:
        """The stream's state (overlap rows, partial row, row count) once its
        submitted segments have finished."""
        n = i64(0)
        call("vfrb_server_save_stream", self.h, sid, None, 0, C.byref(n))
        buf = C.create_string_buffer(n.value)
        call("vfrb_server_save_stream", self.h, sid, buf, n.value, C.byref(n))
        return buf.raw[: n.value]

    def restore_stream(self, sid, state: bytes):
        call("vfrb_server_restore_stream", self.h, sid, state, len(state))

    def stats(self):
        st = ServerStats()
        call("vfrb_server_get_stats", self.h, C.byref(st))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def now(self):
        return lib.vfrb_server_now(self.h)

    def close(self):
        if self.h:
            lib.vfrb_server_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tile_names():
    return [lib.vfrb_search_tile_name(i).decode() for i in range(lib.vfrb_search_tile_count())]


def bandpass_inplace(fb: DeviceArray, nrows, nch, stream=None):
    call("vfrb_bandpass_inplace", fb.ptr, nrows, nch, _s(stream))


def chan_stats(fb: DeviceArray, nrows, nch, ncol, nblk, stream=None):
    ave, sig, out = (DeviceArray((nblk, nch), np.float32) for _ in range(3))
    call("vfrb_chan_stats", fb.ptr, nrows, nch, ncol, nblk, ave.ptr, sig.ptr, out.ptr, _s(stream))
    return ave.download(), sig.download(), out.download()


def bandpass_chan_stats(fb: DeviceArray, nrows, nch, ncol, nblk, stream=None, outs=None):
    """bandpass_inplace + chan_stats in one pass; fb is left unchanged.
    outs: three device arrays [nblk][nch] to fill (then nothing is
    downloaded and they are returned), else new ones are downloaded."""
    arrs = outs or tuple(DeviceArray((nblk, nch), np.float32) for _ in range(3))
    ave, sig, out = arrs
    call("vfrb_bandpass_chan_stats", fb.ptr, nrows, nch, ncol, nblk, ave.ptr, sig.ptr, out.ptr, _s(stream))
    if outs:
        return outs
    return ave.download(), sig.download(), out.download()


@dataclass
class DmStats:
    frb: np.ndarray
    ave: np.ndarray
    sigma: np.ndarray
    peak: np.ndarray
    peak_t: np.ndarray


def dm_stats(out: DeviceArray, ld, ndm, nt, acc=ACC_F64, stream=None) -> DmStats:
    frb = DeviceArray(ndm, np.int32)
    ave, sigma, peak = (DeviceArray(ndm, np.float32) for _ in range(3))
    pt = DeviceArray(ndm, np.int32)
    call("vfrb_dm_stats", out.ptr, ld, ndm, nt, acc, frb.ptr, ave.ptr, sigma.ptr, peak.ptr, pt.ptr,
         _s(stream))
    return DmStats(frb.download(), ave.download(), sigma.download(), peak.download(), pt.download())


def candidates(out: DeviceArray, ld, ndm, nt, ave: np.ndarray, sigma: np.ndarray, threshold,
               max_cand=1 << 16, stream=None):
    d_ave = DeviceArray.from_host(np.asarray(ave, np.float32))
    d_sig = DeviceArray.from_host(np.asarray(sigma, np.float32))
    cdm, ct = DeviceArray(max_cand, np.int32), DeviceArray(max_cand, np.int32)
    cs = DeviceArray(max_cand, np.float32)
    cnt = DeviceArray(1, np.int32)
    call("vfrb_candidates", out.ptr, ld, ndm, nt, d_ave.ptr, d_sig.ptr, float(threshold), cdm.ptr,
         ct.ptr, cs.ptr, max_cand, cnt.ptr, _s(stream))
    n = int(cnt.download()[0])
    k = min(n, max_cand)
    return n, cdm.download()[:k], ct.download()[:k], cs.download()[:k]
