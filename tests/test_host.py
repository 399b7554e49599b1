"""Host-side logic of libvfrb (no GPU): parameter derivation, delay table,
VDIF codec and frame sequencer, output formats, synthetic generator, and
the exported C ABI."""
import os
import re
import struct

import numpy as np
import pytest

import vdif_oracle as vo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---------------------------------------------------------------- ABI ----
def test_library_exports_every_declared_symbol(vfrb):
    hdr = open(os.path.join(ROOT, "include", "vfrb.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"\b(vfrb_[a-z0-9_]+)\s*\(", hdr))
    names -= {"vfrb_log_fn"}
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(vfrb.lib, n)]
    assert not missing, missing
    assert set(vfrb.EXPORTED) <= names | set()  # the binding declares only real symbols


def test_status_strings(vfrb):
    assert vfrb.lib.vfrb_strerror(0) == b"success"
    assert vfrb.lib.vfrb_strerror(8) == b"end of input"
    with pytest.raises(vfrb.VfrbError) as e:
        vfrb.params_derive(7, 4, 32, 70.0)
    assert "twonchan" in str(e.value)


# --------------------------------------------------------- parameters ----
@pytest.mark.parametrize("args", [
    (12500, 4, 32, 70.0, False, 0),    # reference -d defaults
    (12500, 4, 32, 30.0, False, 3840),  # TIMESERIES (NCHANOUT hard-coded)
    (2048, 64, 16, 1.0, True, 0),      # 1024-channel workload
    (8192, 16, 16, 32.38, True, 0),    # 4096-channel window
    (4096, 8, 7, 3.3, False, 0),
])
def test_params_match_oracle(vfrb, orc, args):
    got = vfrb.params_derive(*args).as_dict()
    ref = orc.params(*args)
    for k, v in ref.items():
        assert got[k] == v, (k, got[k], v)


def test_params_reference_defaults_known_values(vfrb):
    p = vfrb.params_derive(12500, 4, 32, 70.0)
    assert (p.nfreq, p.nloop, p.nchanout, p.nskip, p.maxnt, p.nhalfcyc) == (6251, 1600000, 3908, 2343, 179200, 1400)
    assert p.tchan == pytest.approx(0.390625e-3, rel=1e-15)
    p = vfrb.params_derive(2048, 64, 16, 1.0, fullband=True)
    assert (p.nchanout, p.nskip, p.maxnt) == (1024, 1, 976) and p.fch1 == 383.96875
    p = vfrb.params_derive(8192, 16, 16, 20.0, fullband=True)
    assert p.fch1 == 383.9921875 and p.chanbw == 0.015625


def test_delay_table_bit_exact_with_oracle(vfrb, orc):
    for nch, n, nave, dms in [(1024, 2048, 64, np.arange(256) * 0.1),
                              (4096, 8192, 16, np.linspace(0, 1000, 2048)),
                              (3840, 12500, 4, np.array([0.0, 119.3]))]:
        chanbw = 64.0 / (n // 2)
        fch1 = 320.0 + nch * chanbw - 0.5 * chanbw
        tchan = nave * n / 128e6
        dms = dms.astype(np.float32)
        off, maxd, margin = vfrb.delay_table(dms, nch, fch1, chanbw, tchan)
        roff, rmaxd = orc.delay_table(dms, nch, fch1, chanbw, tchan)
        assert np.array_equal(off, roff) and maxd == rmaxd
        assert margin > 1e-9   # no delay sits within libm rounding of an integer


def test_delay_table_rejects_bad_input(vfrb):
    with pytest.raises(vfrb.VfrbError):
        vfrb.delay_table(np.array([1.0, np.nan], np.float32), 16, 383.0, 0.0625, 1e-3)
    with pytest.raises(vfrb.VfrbError):
        vfrb.delay_table(np.array([1.0], np.float32), 100000, 383.0, 0.0625, 1e-3)


# -------------------------------------------------------------- VDIF -----
def test_vdif_header_roundtrip(vfrb):
    rng = np.random.default_rng(0)
    for _ in range(200):
        f = dict(seconds=int(rng.integers(0, 2**30)), frame=int(rng.integers(0, 2**24)),
                 thread=int(rng.integers(0, 1024)), epoch=int(rng.integers(0, 64)),
                 invalid=int(rng.integers(0, 2)), framelength8=629, nbits=7)
        raw = vo.make_header(**f)
        d = vfrb.vdif_decode(raw)
        assert (d["seconds"], d["frame"], d["threadid"], d["epoch"], d["invalid"]) == \
            (f["seconds"], f["frame"], f["thread"], f["epoch"], f["invalid"])
        assert d["framelength8"] * 8 == 5032 and d["nbits"] + 1 == 8
        back = vfrb.vdif_encode(**d)
        assert back == raw


def test_vdif_dmjd_closed_forms(vfrb):
    # epoch 30 = 2015-01-01 (MJD 57023), epoch 31 = 2015-07-01 (MJD 57204),
    # epoch 0 = 2000-01-01 (MJD 51544)
    assert vfrb.vdif_dmjd(vo.make_header(epoch=30)) == 57023.0
    assert vfrb.vdif_dmjd(vo.make_header(epoch=31)) == 57204.0
    assert vfrb.vdif_dmjd(vo.make_header(epoch=0)) == 51544.0
    v = vfrb.vdif_dmjd(vo.make_header(epoch=30, seconds=86400 * 3 + 43200, frame=12800))
    assert v == pytest.approx(57026.5 + 0.5 / 86400, abs=1e-11)


def _stream(frames_per_pol, start_frame=0, start_sec=100, drop=(), first_thread=0, seed=0,
            mismatch=False, truncate=0):
    """Synthetic alternating pol-0/pol-1 frames; `drop` holds frame indices
    missing from both pols."""
    rng = np.random.default_rng(seed)
    fr = []
    if first_thread == 1:
        fr.append((start_sec, start_frame, 1, rng.integers(0, 256, 5000, dtype=np.uint8)))
    for i in range(frames_per_pol):
        fnum = (start_frame + i) % 25600
        sec = start_sec + (start_frame + i) // 25600
        if i in drop:
            continue
        fr.append((sec, fnum, 0, rng.integers(0, 256, 5000, dtype=np.uint8)))
        f1 = fnum + 1 if (mismatch and i == 0) else fnum
        fr.append((sec, f1, 1, rng.integers(0, 256, 5000, dtype=np.uint8)))
    data = vo.build_stream(fr)
    return data[: len(data) - truncate] if truncate else data


@pytest.mark.parametrize("case", [
    dict(),
    dict(drop=(5,)),
    dict(drop=(3, 4, 5, 9)),
    dict(start_frame=25590, drop=(12, 13)),          # gap across the second boundary
    dict(first_thread=1),                            # starts on thread 1: skip one frame
    dict(start_frame=2555),                          # starts 5 frames before a cal boundary
    dict(truncate=1000),                             # truncated final frame
])
@pytest.mark.parametrize("nloop", [5000, 4096 * 3, 2048 * 16])
def test_sequencer_matches_reference_restatement(vfrb, case, nloop):
    data = _stream(40, **case)
    st, ref = vo.reference_bundles(data, nloop)
    assert st == 0
    seq = vfrb.Sequencer(data)
    assert seq.init() == 0
    got = []
    while True:
        s, a, b = seq.fill(nloop)
        if s != 0:
            assert s == 8  # VFRB_ERR_EOF: bundle dropped
            break
        got.append((a.tobytes(), b.tobytes()))
    assert len(got) == len(ref) and len(ref) > 0
    for (a, b), (ra, rb) in zip(got, ref):
        assert a == ra and b == rb
    if case.get("drop"):
        assert any("FRAME SKIP" in line for line in seq.log_lines)


@pytest.mark.parametrize("case", [dict(), dict(drop=(5,)), dict(start_frame=25590, drop=(12, 13)),
                                  dict(start_frame=2555)])
def test_sequencer_seconds_aware_matches_restatement(vfrb, case):
    data = _stream(40, **case)
    nloop = 4096 * 3
    st, ref = vo.reference_bundles(data, nloop, seconds=True)
    seq = vfrb.Sequencer(data, seconds_check=True)
    assert st == 0 and seq.init() == 0
    got = []
    while True:
        s, a, b = seq.fill(nloop)
        if s != 0:
            break
        got.append((a.tobytes(), b.tobytes()))
    assert [(a, b) for a, b in got] == ref


def test_sequencer_seconds_aware_sees_a_whole_second_gap(vfrb):
    # 25600 frames missing: same frame number, next second.  The frame-number
    # rule sees no gap, the seconds-aware rule inserts a second of zeros.
    rng = np.random.default_rng(1)
    fr = []
    for i in range(6):
        sec = 100 + (1 if i >= 3 else 0)
        for th in (0, 1):
            fr.append((sec, i, th, rng.integers(1, 256, 5000, dtype=np.uint8)))
    data = vo.build_stream(fr)
    plain = vfrb.Sequencer(data)
    aware = vfrb.Sequencer(data, seconds_check=True)
    assert plain.init() == 0 and aware.init() == 0
    st, a, _ = plain.fill(5000 * 4)
    assert st == 0 and np.all(a != 0)
    st, a, _ = aware.fill(5000 * 4)
    assert st == 0 and np.all(a[3 * 5000:] == 0)
    _, ref = vo.reference_bundles(data, 5000 * 4, 1, seconds=True)
    assert a.tobytes() == ref[0][0]


def test_sequencer_gap_is_zero_filled(vfrb):
    data = _stream(12, drop=(4, 5))
    seq = vfrb.Sequencer(data)
    assert seq.init() == 0
    st, a, b = seq.fill(5000 * 10)
    assert st == 0
    assert np.all(a[4 * 5000:6 * 5000] == 0) and np.all(b[4 * 5000:6 * 5000] == 0)
    assert np.any(a[6 * 5000:7 * 5000] != 0)
    read, filled = seq.counters()
    assert filled == 4  # 2 frames x 2 pols


def test_sequencer_alignment_errors(vfrb):
    seq = vfrb.Sequencer(_stream(10, mismatch=True))
    assert seq.init() == 7  # VFRB_ERR_FORMAT: pol frame numbers differ
    assert seq.init() != 0 or True
    seq = vfrb.Sequencer(b"")
    assert seq.init() == 8  # end of input
    # two thread-1 frames in a row: thread alignment fails
    bad = vo.build_stream([(0, 0, 1, np.zeros(5000, np.uint8)), (0, 0, 1, np.zeros(5000, np.uint8)),
                           (0, 0, 1, np.zeros(5000, np.uint8))])
    seq = vfrb.Sequencer(bad)
    assert seq.init() == 7


def test_sequencer_cal_alignment_header(vfrb):
    data = _stream(30, start_frame=2550)
    seq = vfrb.Sequencer(data)
    assert seq.init() == 0
    assert vo.parse_header(seq.header())["frame"] == 2560
    assert any("Skipped 10 frames" in line for line in seq.log_lines)


def test_sequencer_reads_file(vfrb, tmp_path):
    data = _stream(15)
    p = tmp_path / "x.vdif"
    p.write_bytes(data)
    seq = vfrb.Sequencer(path=str(p))
    assert seq.init() == 0
    st, a, b = seq.fill(5000 * 4)
    _, ref = vo.reference_bundles(data, 5000 * 4, 1)
    assert st == 0 and a.tobytes() == ref[0][0] and b.tobytes() == ref[0][1]


# ------------------------------------------------------------ formats ----
def test_sigproc_header_bytes(vfrb):
    got = vfrb.sigproc_header(383.96875, -0.0625, 1024, 32, 57023.5, 0.001024)

    def s(x):
        return struct.pack("<i", len(x)) + x.encode()

    want = (s("HEADER_START") + s("barycentric") + struct.pack("<i", 1) + s("telescope_id") +
            struct.pack("<i", 12) + s("data_type") + struct.pack("<i", 1) + s("fch1") +
            struct.pack("<d", 383.96875) + s("foff") + struct.pack("<d", -0.0625) + s("nchans") +
            struct.pack("<i", 1024) + s("nbits") + struct.pack("<i", 32) + s("tstart") +
            struct.pack("<d", 57023.5) + s("tsamp") + struct.pack("<d", 0.001024) + s("nifs") +
            struct.pack("<i", 1) + s("HEADER_END"))
    assert got == want


def test_presto_inf_text(vfrb):
    txt = vfrb.presto_inf("DM0.0_2024-01-01_00:00:00", 57023.25, 1000, 0.000390625, 0.0, 320.005120,
                          39.3216, 3840, 0.01024)
    lines = txt.split("\n")
    assert lines[0] == " Data file name without suffix          = DM0.0_2024-01-01_00:00:00"
    assert " Epoch of observation (MJD)             = 57023.250000000000000" in lines
    assert " Number of bins in the time series      = 1000" in lines
    assert " Width of each time series bin (sec)    = 0.000390625" in lines
    assert " Dispersion measure (cm-3 pc)           = 0.000000" in lines
    assert " Number of channels                     = 3840" in lines
    assert txt.endswith(" Any additional notes:\n    \n\n")


# -------------------------------------------------------------- synth ----
def test_synth_noise_statistics_and_determinism(vfrb):
    spec = vfrb.synth_spec(2048, seed=3)
    a = vfrb.synth_payload(spec, 0, 0, 200).astype(np.float64)
    assert abs(a.mean() - 128.0) < 0.05 and abs(a.std() - np.sqrt(64 + 1 / 12)) < 0.05
    b = vfrb.synth_payload(spec, 0, 50, 10).astype(np.float64)
    assert np.array_equal(a[50 * 2048:60 * 2048], b)
    c = vfrb.synth_payload(spec, 1, 0, 200).astype(np.float64)
    assert abs(np.corrcoef(a, c)[0, 1]) < 0.01
    spec2 = vfrb.synth_spec(2048, seed=3, nthreads=1)
    assert np.array_equal(vfrb.synth_payload(spec2, 0, 0, 200), a.astype(np.uint8))


def test_synth_pulse_lands_on_the_dispersion_curve(vfrb, orc):
    n, nave = 256, 8
    dm, t0 = 1.0, 0.02   # 12 ms sweep across the band, 4000 rows = 64 ms
    spec = vfrb.synth_spec(n, seed=1, pulses=[(dm, t0, 0.002, 2.0)])
    rows = 4000
    p0 = vfrb.synth_payload(spec, 0, 0, rows * nave)
    p1 = vfrb.synth_payload(spec, 1, 0, rows * nave)
    fb = orc.filterbank_f64(p0, p1, rows, n, nave, 1, n // 2)
    tchan = nave * n / 128e6
    for c in (0, 40, 100, 127):
        f = 384.0 - (1 + c) * (64.0 / (n // 2))
        d = 4.148808e3 * dm * (f ** -2 - 384.0 ** -2)
        r = int((t0 + d) / tchan)
        assert fb[r - 1:r + 2, c].mean() > 1.8 * np.median(fb[:, c]), c
