"""The drop-in programs end to end on the GPU, against the oracle replay of
the reference main loops: vdifFILE2dedisp_PIPELINE (windows, overlap carry,
the dumped window's filterbank and dedispersed series, SIGPROC header),
vdifDADA2timeseries (dedisperseSTATS series and sigma per window, PRESTO
files) and vdifDADA2stats (per-channel statistics per block)."""
import glob
import os
import struct
import subprocess

import numpy as np
import pytest

from cli_support import BUILD, reference_rows, reference_windows, write_vdif

pytestmark = pytest.mark.gpu
PIPE = os.path.join(BUILD, "vdifFILE2dedisp_PIPELINE")
TS = os.path.join(BUILD, "vdifDADA2timeseries")
ST = os.path.join(BUILD, "vdifDADA2stats")


def test_pipeline_cli_matches_oracle(gpu, orc, tmp_path):
    vfrb = gpu
    nchan, nave, nb, window = 128, 8, 5, 0.05
    nfft = 2 * nchan
    p = vfrb.params_derive(nfft, nave, nb, window, fullband=True)
    dms = np.array([0.0, 0.5, 1.0, 1.5, 2.0], np.float32)
    with open(tmp_path / "DM101.txt", "w") as f:
        for d in dms:
            f.write(f"{d:f}\n")
    off, maxd, _ = vfrb.delay_table(dms, p.nchanout, p.fch1, p.chanbw, p.tchan)
    ntime = p.maxnt - maxd
    # enough payload for two windows
    rows = p.maxnt + (p.maxnt - (maxd // nb) * nb) + nb
    nblocks = rows * nave + 8
    spec_pulse = [(1.0, 0.03, 0.001, vfrb.amp_for_snr(25, nave, p.nchanout))]
    write_vdif(tmp_path / "in.vdif", vfrb, nfft, nblocks, seed=4, pulses=spec_pulse)
    print(f"[case] CLI maxnt={p.maxnt} max_delay={maxd} ntime={ntime}", flush=True)
    r = subprocess.run([PIPE, "-d", "-nchan", str(nchan), "-nave", str(nave), "-nb", str(nb), "-i", "in.vdif",
                        "-o", "run", "--fullband", "--window", str(window), "--dump-window", "1",
                        "--cands", "run.cands", "--fil-stream", "run.fil"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr + open(tmp_path / "run.log").read()
    assert "TOTAL TIME" in r.stdout
    data = open(tmp_path / "in.vdif", "rb").read()
    wins = reference_windows(data, nfft, nave, nb, p.nskip, p.nchanout, p.maxnt, maxd, orc)
    assert len(wins) >= 2
    bp = orc.bandpass_f64(wins[1])
    fil = np.fromfile(tmp_path / "run.nohead.fil", np.float32).reshape(p.maxnt, p.nchanout)
    assert np.max(np.abs(fil - bp)) / bp.std() < 1e-3
    ref = orc.dedisperse_f64(bp, off, ntime)
    dd = np.fromfile(tmp_path / "OUTDD.bin", np.float32).reshape(len(dms), 1 + ntime)
    assert np.array_equal(dd[:, 0], dms)
    assert np.max(np.abs(dd[:, 1:] - ref)) / ref.std() < 1e-3
    head = open(tmp_path / "run.head", "rb").read()
    assert head.startswith(struct.pack("<i", 12) + b"HEADER_START") and head.endswith(b"HEADER_END")
    assert struct.pack("<d", p.fch1) in head and struct.pack("<d", p.tchan) in head
    cands = open(tmp_path / "run.cands").read().splitlines()
    assert cands[0].startswith("#")
    # streamed .fil: the same header, then every window's bandpassed rows from
    # where its filling started (row 0, then (max_delay/NBUNDLE)*NBUNDLE)
    stream = open(tmp_path / "run.fil", "rb").read()
    assert stream.startswith(head)
    rows = np.frombuffer(stream[len(head):], np.float32).reshape(-1, p.nchanout)
    r0 = (maxd // nb) * nb
    ref = np.concatenate([orc.bandpass_f64(w)[0 if i == 0 else r0:] for i, w in enumerate(wins)])
    assert rows.shape == ref.shape
    assert np.max(np.abs(rows - ref)) / ref.std() < 1e-3


def test_timeseries_cli_matches_oracle(gpu, orc, tmp_path):
    vfrb = gpu
    # NCHANOUT = 3840 of 4096 channels (TIMESERIES:282), a short window and a
    # small DM list instead of 30 s and DM 119.3
    nchan, nave, nb, window = 4096, 1, 5, 0.05
    nfft = 2 * nchan
    p = vfrb.params_derive(nfft, nave, nb, window, False, 3840)
    dms = np.array([0.0, 0.3], np.float32)
    with open(tmp_path / "dms.txt", "w") as f:
        for d in dms:
            f.write(f"{d:f}\n")
    off, maxd, _ = vfrb.delay_table(dms, p.nchanout, p.fch1, p.chanbw, p.tchan)
    ntime = p.maxnt - maxd
    rows = p.maxnt + (p.maxnt - (maxd // nb) * nb) + nb
    nblocks = rows * nave + 8
    amp = vfrb.amp_for_snr(25, nave, p.nchanout)
    write_vdif(tmp_path / "in.vdif", vfrb, nfft, nblocks, seed=6, pulses=[(0.2, 0.02, 0.0005, amp)])
    print(f"[case] TIMESERIES CLI maxnt={p.maxnt} max_delay={maxd} ntime={ntime}", flush=True)
    r = subprocess.run([TS, "-d", "-nchan", str(nchan), "-nave", str(nave), "-nb", str(nb), "-o", "ts",
                        "--input-file", "in.vdif", "--skip-pairs", "0", "--window", str(window),
                        "--dmlist", "dms.txt", "--final-inf"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    logs = "".join(open(f).read() for f in glob.glob(str(tmp_path / "ts*.log")))
    assert r.returncode == 0, r.stderr + logs
    data = open(tmp_path / "in.vdif", "rb").read()
    wins = reference_windows(data, nfft, nave, nb, p.nskip, p.nchanout, p.maxnt, maxd, orc)
    assert len(wins) >= 2
    for k, dm in enumerate(dms):
        dat = [f for f in glob.glob(str(tmp_path / f"DM{dm:.1f}_*.dat")) if not f.endswith("_SIGMA.dat")]
        sig = glob.glob(str(tmp_path / f"DM{dm:.1f}_*_SIGMA.dat"))
        inf = glob.glob(str(tmp_path / f"DM{dm:.1f}_*.inf"))
        assert len(dat) == 1 and len(sig) == 1 and len(inf) == 1
        series = np.fromfile(dat[0], np.float32).reshape(-1, ntime)
        sigma = np.fromfile(sig[0], np.float32).reshape(-1, ntime)
        assert series.shape[0] == len(wins) and sigma.shape[0] == len(wins)
        for w, win in enumerate(wins):
            ref = orc.dedisperse_f64(orc.bandpass_f64(win), off, ntime)[k]
            assert np.max(np.abs(series[w] - ref)) / ref.std() < 1e-3, (k, w)
            _, rsig = orc.dedisperse_stats_f32(orc.bandpass_f32(win.astype(np.float32)), off, ntime)
            assert np.max(np.abs(sigma[w] / rsig[k] - 1)) < 1e-3, (k, w)


def test_stats_cli_matches_oracle(gpu, orc, tmp_path):
    vfrb = gpu
    nchan, nave, nb, window, statsec = 128, 8, 5, 0.05, 0.01
    nfft = 2 * nchan
    p = vfrb.params_derive(nfft, nave, nb, window, fullband=True)
    tsamp = nave * nfft / 128e6
    # STATS:209-232: float MAXTIME/TIMESTATS; TIMESTATS (float) / tsamp (double)
    nblk = int(np.float32(window) / np.float32(statsec))
    ncol = int(float(np.float32(statsec)) / tsamp)
    rows = 2 * p.maxnt + nb
    write_vdif(tmp_path / "in.vdif", vfrb, nfft, rows * nave + 8, seed=9)
    print(f"[case] STATS CLI maxnt={p.maxnt} blocks={nblk}x{ncol}", flush=True)
    r = subprocess.run([ST, "--input-file", "in.vdif", "-o", "st", "--nchan", str(nchan), "--nave", str(nave),
                        "--nb", str(nb), "--window", str(window), "--statsec", str(statsec), "--fullband"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + open(tmp_path / "st.log").read()
    got = np.fromfile(tmp_path / "st.stats", np.float32).reshape(-1, 3, nblk, p.nchanout)
    data = open(tmp_path / "in.vdif", "rb").read()
    fb = reference_rows(data, nfft, nave, nb, p.nskip, p.nchanout, orc)
    nwin = fb.shape[0] // p.maxnt
    assert got.shape[0] == nwin >= 2
    for w in range(nwin):
        bp = orc.bandpass_f32(fb[w * p.maxnt:(w + 1) * p.maxnt].astype(np.float32))
        ave, sig, out = orc.chan_stats(bp, ncol, nblk)
        # per-block mean of bandpassed (zero-mean) data: compare on the noise scale
        scale = np.sqrt(np.abs(sig)) + 1e-30
        assert np.max(np.abs(got[w, 0] - ave) / scale) < 1e-3, w
        assert np.max(np.abs(got[w, 1] / sig - 1)) < 2e-3, w
        assert np.array_equal(got[w, 2], out), w


# A shape past the reference's launch limit at a tenth of its data rate cost:
# 64 channels, NAVE 1 (1 us rows), 0.15 s windows: MAXNT = 150 000 rows,
# NHALFCYC = 1172 > 1024, as at PIPELINE -d (1400) and STATS' defaults (1200).
LIMIT_SHAPE = dict(nchan=64, nave=1, nb=50, window=0.15)


def _limit_params(vfrb):
    s = LIMIT_SHAPE
    p = vfrb.params_derive(2 * s["nchan"], s["nave"], s["nb"], s["window"], fullband=True)
    assert p.maxnt > 131072 and p.nhalfcyc > 1024
    return p


@pytest.mark.parametrize("always", [False, True], ids=["reference-default", "bandpass-always"])
def test_pipeline_cli_at_reference_launch_limit(gpu, orc, tmp_path, always):
    """By default the program reproduces the reference's output when its
    subtractbpCAL<<<NCHANOUT,NHALFCYC>>> launch exceeds 1024 threads per
    block (PIPELINE:323-324,658): the window is searched without bandpass.
    --bandpass-always applies the bandpass at every shape."""
    vfrb = gpu
    s = LIMIT_SHAPE
    nfft = 2 * s["nchan"]
    p = _limit_params(vfrb)
    dms = np.array([0.0, 0.5, 1.0], np.float32)
    with open(tmp_path / "DM101.txt", "w") as f:
        for d in dms:
            f.write(f"{d:f}\n")
    off, maxd, _ = vfrb.delay_table(dms, p.nchanout, p.fch1, p.chanbw, p.tchan)
    ntime = p.maxnt - maxd
    rows = p.maxnt + s["nb"]
    write_vdif(tmp_path / "in.vdif", vfrb, nfft, rows * s["nave"] + 8, seed=12)
    print(f"[case] PIPELINE CLI at the launch limit, maxnt={p.maxnt} nhalfcyc={p.nhalfcyc} "
          f"bandpass_always={always}", flush=True)
    cmd = [PIPE, "-d", "-nchan", str(s["nchan"]), "-nave", str(s["nave"]), "-nb", str(s["nb"]), "-i", "in.vdif",
           "-o", "run", "--fullband", "--window", str(s["window"]), "--dump-window", "0"]
    r = subprocess.run(cmd + (["--bandpass-always"] if always else []), cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr + open(tmp_path / "run.log").read()
    data = open(tmp_path / "in.vdif", "rb").read()
    win = reference_windows(data, nfft, s["nave"], s["nb"], p.nskip, p.nchanout, p.maxnt, maxd, orc)[0]
    bp = orc.bandpass_f64(win)
    fil = np.fromfile(tmp_path / "run.nohead.fil", np.float32).reshape(p.maxnt, p.nchanout)
    dd = np.fromfile(tmp_path / "OUTDD.bin", np.float32).reshape(len(dms), 1 + ntime)[:, 1:]
    if always:
        assert np.max(np.abs(fil - bp)) / bp.std() < 1e-3
        ref = orc.dedisperse_f64(bp, off, ntime)
        assert np.max(np.abs(dd - ref)) / ref.std() < 1e-3
    else:
        # the raw filterbank, as the reference leaves it
        assert np.max(np.abs(fil - win)) / np.max(np.abs(win)) < 1e-5
        ref = orc.dedisperse_f64(win, off, ntime)
        assert np.max(np.abs(dd - ref)) / np.max(np.abs(ref)) < 1e-5
        assert np.max(np.abs(fil - bp)) / bp.std() > 1.0  # and visibly not bandpassed


@pytest.mark.parametrize("always", [False, True], ids=["reference-default", "bandpass-always"])
def test_stats_cli_at_reference_launch_limit(gpu, orc, tmp_path, always):
    """The STATS program's statistics at NHALFCYC > 1024: of the raw
    filterbank by default (the reference's failed subtractbpCAL launch,
    STATS:248-249,414), of the bandpassed one with --bandpass-always."""
    vfrb = gpu
    s = LIMIT_SHAPE
    nfft = 2 * s["nchan"]
    statsec = 0.01
    p = _limit_params(vfrb)
    tsamp = s["nave"] * nfft / 128e6
    nblk = int(np.float32(s["window"]) / np.float32(statsec))
    ncol = int(float(np.float32(statsec)) / tsamp)
    rows = p.maxnt + s["nb"]
    write_vdif(tmp_path / "in.vdif", vfrb, nfft, rows * s["nave"] + 8, seed=13)
    print(f"[case] STATS CLI at the launch limit, maxnt={p.maxnt} nhalfcyc={p.nhalfcyc} blocks={nblk}x{ncol} "
          f"bandpass_always={always}", flush=True)
    cmd = [ST, "--input-file", "in.vdif", "-o", "st", "--nchan", str(s["nchan"]), "--nave", str(s["nave"]),
           "--nb", str(s["nb"]), "--window", str(s["window"]), "--statsec", str(statsec), "--fullband"]
    r = subprocess.run(cmd + (["--bandpass-always"] if always else []), cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr + open(tmp_path / "st.log").read()
    got = np.fromfile(tmp_path / "st.stats", np.float32).reshape(-1, 3, nblk, p.nchanout)
    data = open(tmp_path / "in.vdif", "rb").read()
    fb = reference_rows(data, nfft, s["nave"], s["nb"], p.nskip, p.nchanout, orc)[:p.maxnt].astype(np.float32)
    assert got.shape[0] == 1
    x = orc.bandpass_f32(fb) if always else fb
    ave, sig, out = orc.chan_stats(x, ncol, nblk)
    if always:
        scale = np.sqrt(np.abs(sig)) + 1e-30
        assert np.max(np.abs(got[0, 0] - ave) / scale) < 1e-3
    else:
        assert np.max(np.abs(got[0, 0] / ave - 1)) < 1e-5
    assert np.max(np.abs(got[0, 1] / sig - 1)) < 2e-3
    assert np.array_equal(got[0, 2], out)


def _run_from_pipe(cmd, data, cwd, timeout=120):
    """Runs cmd with `data` fed through a pipe on its stdin (--input-file -),
    the way a shell pipeline or a ring reader hands over the VDIF stream."""
    p = subprocess.Popen(cmd, cwd=cwd, stdin=subprocess.PIPE, stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    out, err = p.communicate(data, timeout=timeout)
    return p.returncode, out.decode(errors="replace"), err.decode(errors="replace")


def _inf_fields(path):
    f = {}
    for line in open(path).read().splitlines():
        if "=" in line:
            k, v = line.split("=", 1)
            f[k.strip()] = v.strip()
    return f


def test_timeseries_cli_rotation_from_a_pipe(gpu, orc, tmp_path):
    """vdifDADA2timeseries reading its VDIF stream from a pipe, rotating its
    PRESTO files every 2 windows (--rotate-windows, the hourly rollover of
    TIMESERIES:694-739 on a window count): each file set holds its windows'
    series in order, each .inf its ntimes, and mjdstart advances by
    ntimes*tsamp/86400 from the first frame's MJD."""
    vfrb = gpu
    nchan, nave, nb, window = 4096, 1, 5, 0.05
    nfft = 2 * nchan
    p = vfrb.params_derive(nfft, nave, nb, window, False, 3840)
    dms = np.array([0.0, 0.3], np.float32)
    with open(tmp_path / "dms.txt", "w") as f:
        for d in dms:
            f.write(f"{d:f}\n")
    off, maxd, _ = vfrb.delay_table(dms, p.nchanout, p.fch1, p.chanbw, p.tchan)
    ntime = p.maxnt - maxd
    r0 = (maxd // nb) * nb
    rows = p.maxnt + 5 * (p.maxnt - r0) + nb
    write_vdif(tmp_path / "in.vdif", vfrb, nfft, rows * nave + 8, seed=14)
    data = open(tmp_path / "in.vdif", "rb").read()
    print(f"[case] TIMESERIES rotation from a pipe: maxnt={p.maxnt} max_delay={maxd} ntime={ntime}", flush=True)
    rc, _, err = _run_from_pipe([TS, "-d", "-nchan", str(nchan), "-nave", str(nave), "-nb", str(nb), "-o", "ts",
                                 "--input-file", "-", "--skip-pairs", "0", "--window", str(window),
                                 "--dmlist", "dms.txt", "--rotate-windows", "2", "--final-inf"], data, tmp_path)
    logs = "".join(open(f).read() for f in glob.glob(str(tmp_path / "ts*.log")))
    assert rc == 0, err + logs
    wins = reference_windows(data, nfft, nave, nb, p.nskip, p.nchanout, p.maxnt, maxd, orc)
    nwin = len(wins)
    assert nwin >= 6
    tsamp = nave * nfft / 128e6
    tstart = vfrb.vdif_dmjd(data[:32])
    for k, dm in enumerate(dms):
        def order(path):
            b = os.path.basename(path)[:-4]
            return int(b.rsplit("_w", 1)[1]) if "_w" in b else 0
        dat = sorted((f for f in glob.glob(str(tmp_path / f"DM{dm:.1f}_*.dat")) if not f.endswith("_SIGMA.dat")),
                     key=order)
        # a set per 2 windows, and the empty set opened after the last rotation
        assert len(dat) == nwin // 2 + 1, dat
        mjd = tstart
        w = 0
        for i, path in enumerate(dat):
            base = path[:-4]
            series = np.fromfile(path, np.float32)
            sigma = np.fromfile(base + "_SIGMA.dat", np.float32)
            nw = min(2, nwin - w)
            assert series.size == nw * ntime and sigma.size == nw * ntime, (k, i)
            for j in range(nw):
                ref = orc.dedisperse_f64(orc.bandpass_f64(wins[w + j]), off, ntime)[k]
                got = series[j * ntime:(j + 1) * ntime]
                assert np.max(np.abs(got - ref)) / ref.std() < 1e-3, (k, i, j)
            inf = _inf_fields(base + ".inf")
            assert inf["Data file name without suffix"] == os.path.basename(base)
            assert int(inf["Number of bins in the time series"]) == nw * ntime
            assert inf["Epoch of observation (MJD)"] == f"{mjd:.15f}", (k, i)
            assert float(inf["Dispersion measure (cm-3 pc)"]) == pytest.approx(float(dm), abs=1e-6)
            mjd = mjd + (nw * ntime * tsamp / 86400.0)
            w += nw
        assert w == nwin


def test_stats_cli_from_a_pipe_equals_file_input(gpu, tmp_path):
    """vdifDADA2stats with its VDIF stream on a pipe (--input-file -) writes
    the same .stats bytes as from the file."""
    vfrb = gpu
    nchan, nave, nb, window, statsec = 128, 8, 5, 0.05, 0.01
    nfft = 2 * nchan
    p = vfrb.params_derive(nfft, nave, nb, window, fullband=True)
    write_vdif(tmp_path / "in.vdif", vfrb, nfft, (3 * p.maxnt + nb) * nave + 8, seed=15)
    args = ["--nchan", str(nchan), "--nave", str(nave), "--nb", str(nb), "--window", str(window), "--statsec",
            str(statsec), "--fullband"]
    print("[case] STATS from a file and from a pipe", flush=True)
    r = subprocess.run([ST, "--input-file", "in.vdif", "-o", "file"] + args, cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    data = open(tmp_path / "in.vdif", "rb").read()
    rc, _, err = _run_from_pipe([ST, "--input-file", "-", "-o", "pipe"] + args, data, tmp_path)
    assert rc == 0, err
    a = open(tmp_path / "file.stats", "rb").read()
    b = open(tmp_path / "pipe.stats", "rb").read()
    per_window = 3 * 4 * p.nchanout * int(np.float32(window) / np.float32(statsec))
    assert len(a) % per_window == 0 and len(a) // per_window >= 2 and a == b


def _edit_vdif(src, dst, lead_thread1=False, swap_pair=None, cut_tail=0):
    """Copies a VDIF file of alternating thread 0/1 frames, optionally with a
    thread-1 frame in front (misaligned start), one pair's threads swapped
    (a stream the device ingest rejects), or the last cut_tail bytes removed."""
    import vdif_oracle as vo
    data = bytearray(open(src, "rb").read())
    if swap_pair is not None:
        o = swap_pair * 2 * 5032
        h0, h1 = bytes(data[o:o + 32]), bytes(data[o + 5032:o + 5064])
        data[o:o + 32], data[o + 5032:o + 5064] = h1, h0
    if lead_thread1:
        data = bytearray(vo.make_header(seconds=999, frame=25599, thread=1) + bytes(5000)) + data
    if cut_tail:
        data = data[:-cut_tail]
    open(dst, "wb").write(bytes(data))


@pytest.mark.timeout(600)
@pytest.mark.parametrize("case", ["drops", "offcal_misaligned", "thread_swap", "odd_bundle"])
def test_pipeline_fast_input_equals_host_sequencer(gpu, tmp_path, case):
    """The drop-in program's fast input path (large chunks of raw frame pairs,
    device ingest) writes exactly the bytes of its host frame-sequencer path,
    whose bundles the CLI oracle test pins to the reference's loop: across
    chunk boundaries (more than 4096 frame pairs), with dropped frames,
    a misaligned and off-cal start, a truncated last frame, a stream the
    device ingest rejects (thread ids swapped mid-stream: the program reruns
    on the host sequencer) and a bundle size that is not a multiple of 4
    bytes (host path from the start)."""
    vfrb = gpu
    nchan, nave, nb, window = (125, 1, 1, 0.01) if case == "odd_bundle" else (128, 8, 5, 0.05)
    nfft = 2 * nchan
    p = vfrb.params_derive(nfft, nave, nb, window, fullband=True)
    dms = np.array([0.0, 0.4, 0.8, 1.2], np.float32) * (0.25 if case == "odd_bundle" else 1.0)
    rows = 9 * p.maxnt if case != "odd_bundle" else 3 * p.maxnt
    drop = {100, 101, 102, 5000, 8191, 8192, 12000} if case == "drops" else ()
    start = 2556 if case == "offcal_misaligned" else 0
    write_vdif(tmp_path / "raw.vdif", vfrb, nfft, rows * nave + 64, seed=21, start_frame=start, drop=drop,
               pulses=[] if case == "odd_bundle" else [(0.8, 0.02, 0.001, vfrb.amp_for_snr(25, nave, p.nchanout))])
    _edit_vdif(tmp_path / "raw.vdif", tmp_path / "in.vdif", lead_thread1=case == "offcal_misaligned",
               swap_pair=9000 if case == "thread_swap" else None, cut_tail=1234 if case == "drops" else 0)
    size = os.path.getsize(tmp_path / "in.vdif")
    print(f"[case] fast vs host input, {case}: {size / 2 / 5032:.0f} pairs, maxnt={p.maxnt}", flush=True)
    outs = {}
    for mode in ("fast", "host"):
        d = tmp_path / mode
        d.mkdir()
        with open(d / "DM101.txt", "w") as f:
            for x in dms:
                f.write(f"{x:f}\n")
        cmd = [PIPE, "-d", "-nchan", str(nchan), "-nave", str(nave), "-nb", str(nb), "-i", str(tmp_path / "in.vdif"),
               "-o", "run", "--fullband", "--window", str(window), "--dump-window", "2", "--cands", "run.cands",
               "--fil-stream", "run.fil"] + (["--host-sequencer"] if mode == "host" else [])
        r = subprocess.run(cmd, cwd=d, capture_output=True, text=True, timeout=300)
        log = open(d / "run.log").read()
        assert r.returncode == 0, r.stderr + log
        outs[mode] = {f: open(d / f, "rb").read() for f in ("OUTDD.bin", "run.nohead.fil", "run.head", "run.fil",
                                                              "run.cands")}
        outs[mode]["log"] = log
    for f in ("OUTDD.bin", "run.nohead.fil", "run.head", "run.fil", "run.cands"):
        assert outs["fast"][f] == outs["host"][f], f
    nwin = (len(outs["host"]["run.fil"]) - len(outs["host"]["run.head"])) // (4 * p.nchanout)
    assert nwin >= p.maxnt  # at least a window's rows went through
    if case == "thread_swap":
        assert "rerunning on the host frame sequencer" in outs["fast"]["log"]
    elif case == "odd_bundle":
        assert "bundle size not a multiple of 4 bytes" in outs["fast"]["log"]
    else:
        assert "fast input path:" in outs["fast"]["log"]
