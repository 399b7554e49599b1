"""Helpers shared by the CLI tests: synthetic VDIF files and the reference's
window/overlap bookkeeping (PIPELINE:626-704) replayed on the oracle."""
import os

import numpy as np

import vdif_oracle as vo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "vlite-frb_amd", "_build")


def write_vdif(path, vfrb, nfft, nblocks, seed=0, pulses=(), start_frame=0, drop=()):
    """Both pols' payload streams cut into 5000-byte frames, alternating
    thread 0/1, frame numbers from start_frame (wrapping at 25600)."""
    spec = vfrb.synth_spec(nfft, seed=seed, pulses=list(pulses))
    s0 = vfrb.synth_payload(spec, 0, 0, nblocks)
    s1 = vfrb.synth_payload(spec, 1, 0, nblocks)
    nfr = s0.size // 5000
    with open(path, "wb") as f:
        for i in range(nfr):
            if i in drop:
                continue
            fn = (start_frame + i) % 25600
            sec = 1000 + (start_frame + i) // 25600
            f.write(vo.make_header(seconds=sec, frame=fn, thread=0))
            f.write(s0[i * 5000:(i + 1) * 5000].tobytes())
            f.write(vo.make_header(seconds=sec, frame=fn, thread=1))
            f.write(s1[i * 5000:(i + 1) * 5000].tobytes())
    return s0, s1


def reference_windows(data, nfft, nave, nbundle, nskip, nch, maxnt, max_delay, orc):
    """Raw (pre-bandpass) filterbank of every window the reference main loop
    processes: bundles from the restated sequencer, NBUNDLE rows each, the
    window full at MAXNT/NBUNDLE bundles, then the tail rows
    [MAXNT-max_delay, MAXNT) moved to the front and filling restarting at
    row (max_delay/NBUNDLE)*NBUNDLE."""
    nloop = nbundle * nave * nfft
    st, bundles = vo.reference_bundles(data, nloop)
    assert st == 0
    fb = np.zeros((maxnt, nch), np.float64)
    windows = []
    loopcnt = 0
    for a, b in bundles:
        rows = orc.filterbank_f64(np.frombuffer(a, np.uint8), np.frombuffer(b, np.uint8), nbundle, nfft,
                                  nave, nskip, nch)
        fb[loopcnt * nbundle:(loopcnt + 1) * nbundle] = rows
        loopcnt += 1
        if loopcnt == maxnt // nbundle:
            windows.append(fb.copy())
            tail = fb[maxnt - max_delay:maxnt].copy()
            fb[:max_delay] = tail
            loopcnt = max_delay // nbundle
    return windows


def reference_rows(data, nfft, nave, nbundle, nskip, nch, orc, seconds=False):
    """Raw filterbank rows (FP64) of every bundle the reference sequencer
    delivers, in order: the input of windows that do not overlap (STATS)."""
    nloop = nbundle * nave * nfft
    st, bundles = vo.reference_bundles(data, nloop, seconds=seconds)
    assert st == 0
    rows = [orc.filterbank_f64(np.frombuffer(a, np.uint8), np.frombuffer(b, np.uint8), nbundle, nfft, nave,
                               nskip, nch) for a, b in bundles]
    return np.concatenate(rows) if rows else np.zeros((0, nch))
