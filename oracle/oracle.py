"""ctypes binding of liboracle.so (vfrb_oracle.c) plus pure-Python oracles.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker (or the timed CPU
baseline), never by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


if not os.path.exists(LIB_PATH):
    build()
lib = C.CDLL(LIB_PATH)

P = C.POINTER
vp = C.c_void_p
i32, i64, f64 = C.c_int32, C.c_int64, C.c_double

_SIGS = {
    "or_params": (i32, [i32, i32, i32, f64, i32, i32, vp, vp]),
    "or_delay_table": (i32, [vp, i32, i32, f64, f64, f64, vp, vp]),
    "or_fft_f64": (None, [i32, vp, vp, vp, vp]),
    "or_dft_f64": (None, [i32, vp, vp, vp, vp]),
    "or_filterbank_f64": (None, [vp, vp, i64, i32, i32, i32, i32, vp]),
    "or_filterbank_f32": (None, [vp, vp, i64, i32, i32, i32, i32, vp]),
    "or_bandpass_f32": (None, [vp, i32, i32, vp]),
    "or_bandpass_f64": (None, [vp, i32, i32]),
    "or_dedisperse_f32": (None, [vp, i32, i32, vp, i32, i32, vp]),
    "or_dedisperse_cm_f32": (None, [vp, i64, i32, vp, i32, i32, vp]),
    "or_dedisperse_f64": (None, [vp, i32, i32, vp, i32, i32, vp]),
    "or_dedisperse_stats_f32": (None, [vp, i32, vp, i32, i32, vp, vp]),
    "or_dm_stats": (None, [vp, i64, i32, i32, vp, vp, vp, vp, vp]),
    "or_dm_stats_f64": (None, [vp, i64, i32, i32, vp, vp, vp, vp]),
    "or_chan_stats": (None, [vp, i32, i32, i32, i32, vp, vp, vp]),
    "or_num_threads": (i32, []),
    "or_set_threads": (None, [i32]),
}
for _n, (_r, _a) in _SIGS.items():
    _f = getattr(lib, _n)
    _f.restype = _r
    _f.argtypes = _a


def _p(a):
    return a.ctypes.data_as(vp)


def params(twonchan, nave, nbundle, maxtime, fullband=False, nchanout_override=0) -> dict:
    oi = np.zeros(10, np.int32)
    od = np.zeros(5, np.float64)
    lib.or_params(twonchan, nave, nbundle, maxtime, int(fullband), nchanout_override, _p(oi), _p(od))
    keys_i = ["twonchan", "nchan", "nfreq", "nave", "nbundle", "nloop", "nchanout", "nskip",
              "maxnt", "nhalfcyc"]
    keys_d = ["chanbw", "foff", "fch1", "tchan", "tbundle"]
    d = {k: int(v) for k, v in zip(keys_i, oi)}
    d.update({k: float(v) for k, v in zip(keys_d, od)})
    return d


def delay_table(dms, nchanout, fch1, chanbw, tchan):
    dms = np.ascontiguousarray(dms, np.float32)
    off = np.empty((len(dms), nchanout), np.int32)
    md = np.zeros(1, np.int32)
    lib.or_delay_table(_p(dms), len(dms), nchanout, fch1, chanbw, tchan, _p(off), _p(md))
    return off, int(md[0])


def fft(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, np.complex128)
    re, im = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
    ore, oim = np.empty_like(re), np.empty_like(re)
    lib.or_fft_f64(len(x), _p(re), _p(im), _p(ore), _p(oim))
    return ore + 1j * oim


def dft(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, np.complex128)
    re, im = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
    ore, oim = np.empty_like(re), np.empty_like(re)
    lib.or_dft_f64(len(x), _p(re), _p(im), _p(ore), _p(oim))
    return ore + 1j * oim


def filterbank_f64(p0, p1, nrows, n, nave, nskip, nchanout) -> np.ndarray:
    p0, p1 = np.ascontiguousarray(p0, np.uint8), np.ascontiguousarray(p1, np.uint8)
    assert p0.size >= nrows * nave * n and p1.size >= nrows * nave * n
    out = np.empty((nrows, nchanout), np.float64)
    lib.or_filterbank_f64(_p(p0), _p(p1), nrows, n, nave, nskip, nchanout, _p(out))
    return out


def filterbank_f32(p0, p1, nrows, n, nave, nskip, nchanout) -> np.ndarray:
    p0, p1 = np.ascontiguousarray(p0, np.uint8), np.ascontiguousarray(p1, np.uint8)
    assert p0.size >= nrows * nave * n and p1.size >= nrows * nave * n
    assert n & (n - 1) == 0
    out = np.empty((nrows, nchanout), np.float32)
    lib.or_filterbank_f32(_p(p0), _p(p1), nrows, n, nave, nskip, nchanout, _p(out))
    return out


def bandpass_f32(fb: np.ndarray, with_margin=False):
    fb = np.array(fb, np.float32, order="C", copy=True)
    nrows, nch = fb.shape
    margin = np.empty(((nrows + 127) // 128, nch), np.float32)
    lib.or_bandpass_f32(_p(fb), nrows, nch, _p(margin))
    return (fb, margin) if with_margin else fb


def bandpass_f64(fb: np.ndarray) -> np.ndarray:
    fb = np.array(fb, np.float64, order="C", copy=True)
    lib.or_bandpass_f64(_p(fb), fb.shape[0], fb.shape[1])
    return fb


def dedisperse_f32(fb: np.ndarray, off: np.ndarray, nt: int) -> np.ndarray:
    fb = np.ascontiguousarray(fb, np.float32)
    off = np.ascontiguousarray(off, np.int32)
    ndm, nch = off.shape
    assert fb.shape[1] == nch and nt + off.max() <= fb.shape[0]
    out = np.empty((ndm, nt), np.float32)
    lib.or_dedisperse_f32(_p(fb), fb.shape[0], nch, _p(off), ndm, nt, _p(out))
    return out


def dedisperse_cm_f32(fbcm: np.ndarray, off: np.ndarray, nt: int) -> np.ndarray:
    fbcm = np.ascontiguousarray(fbcm, np.float32)
    off = np.ascontiguousarray(off, np.int32)
    ndm, nch = off.shape
    assert fbcm.shape[0] == nch and nt + off.max() <= fbcm.shape[1]
    out = np.empty((ndm, nt), np.float32)
    lib.or_dedisperse_cm_f32(_p(fbcm), fbcm.shape[1], nch, _p(off), ndm, nt, _p(out))
    return out


def dedisperse_f64(fb: np.ndarray, off: np.ndarray, nt: int) -> np.ndarray:
    fb = np.ascontiguousarray(fb, np.float64)
    off = np.ascontiguousarray(off, np.int32)
    ndm, nch = off.shape
    assert fb.shape[1] == nch and nt + off.max() <= fb.shape[0]
    out = np.empty((ndm, nt), np.float64)
    lib.or_dedisperse_f64(_p(fb), fb.shape[0], nch, _p(off), ndm, nt, _p(out))
    return out


def dedisperse_stats_f32(fb: np.ndarray, off: np.ndarray, nt: int):
    fb = np.ascontiguousarray(fb, np.float32)
    off = np.ascontiguousarray(off, np.int32)
    ndm, nch = off.shape
    assert nt + off.max() <= fb.shape[0]
    out = np.empty((ndm, nt), np.float32)
    sig = np.empty((ndm, nt), np.float32)
    lib.or_dedisperse_stats_f32(_p(fb), nch, _p(off), ndm, nt, _p(out), _p(sig))
    return out, sig


def dm_stats(out: np.ndarray):
    out = np.ascontiguousarray(out, np.float32)
    ndm, nt = out.shape
    frb = np.empty(ndm, np.int32)
    ave, sig, peak = (np.empty(ndm, np.float32) for _ in range(3))
    pt = np.empty(ndm, np.int32)
    lib.or_dm_stats(_p(out), nt, ndm, nt, _p(frb), _p(ave), _p(sig), _p(peak), _p(pt))
    return frb, ave, sig, peak, pt


def dm_stats_f64(out: np.ndarray):
    out = np.ascontiguousarray(out, np.float64)
    ndm, nt = out.shape
    ave, sig, peak = (np.empty(ndm, np.float64) for _ in range(3))
    pt = np.empty(ndm, np.int32)
    lib.or_dm_stats_f64(_p(out), nt, ndm, nt, _p(ave), _p(sig), _p(peak), _p(pt))
    return ave, sig, peak, pt


def chan_stats(fb: np.ndarray, ncol: int, nblk: int):
    fb = np.ascontiguousarray(fb, np.float32)
    nrows, nch = fb.shape
    ave, sig, out = (np.empty((nblk, nch), np.float32) for _ in range(3))
    lib.or_chan_stats(_p(fb), nrows, nch, ncol, nblk, _p(ave), _p(sig), _p(out))
    return ave, sig, out


def set_threads(n: int):
    lib.or_set_threads(n)


def num_threads() -> int:
    return lib.or_num_threads()
