#!/usr/bin/env python3
"""Benchmark of the VLITE FRB search chain on one MI355X.

Workload (default): 4096 channels (8192-point FFT, 16 spectra averaged,
1.024 ms samples), 2048 DM trials 0..1000, 20 s segments of synthetic 8-bit
VDIF payload (N(128, 8^2) noise, 8 narrowband tones at 5 sigma, 5 dispersed
pulses) with the 12090-row maximum-delay overlap carried between segments
in an HBM ring.  One step = one segment through the whole chain, with its
payload already resident in HBM:
    filterbank (fused unpack/FFT/detect/average) -> bandpass + pack ->
    dedispersion -> per-DM statistics -> candidate list.
value = real-time factor (data-seconds per wall-second) of the configured
storage; the other storages are measured in the same run and reported.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vlite-frb_amd"))

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def workload(name):
    if name == "c2":
        return dict(workload="plain pipeline, 4096 ch x 2048 DM, 20 s segments", twonchan=8192,
                    nave=16, ndm=2048, dm_max=1000.0, seg_s=20.0, npulse=5, ntone=8,
                    pulse_dm=(50, 900), width_ms=(1, 8), snr=(8, 30), seed=1)
    if name == "c1":
        return dict(workload="plain pipeline, 1024 ch x 256 DM, 1 s segments", twonchan=2048,
                    nave=64, ndm=256, dm_max=25.5, seg_s=1.0, npulse=1, ntone=0,
                    pulse_dm=(12.3, 12.3), width_ms=(2, 2), snr=(20, 20), seed=0)
    raise SystemExit(f"unknown config {name}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=["c1", "c2"])
    ap.add_argument("--storage", default="f16", choices=["f32", "f16", "u8"])
    ap.add_argument("--compare", default="f32,f16,u8", help="storages also timed in this run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--tile", type=int, default=-1)
    ap.add_argument("--payload-cache", default="",
                    help="file to reuse the synthesised payload between runs (written if absent)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(sharding is described in DESIGN.md); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import vfrb

    assert vfrb.device_count() > 0, "no GPU"
    dev = vfrb.device_name()
    log("device", dev)
    W = workload(args.config)
    n, nave = W["twonchan"], W["nave"]
    p = vfrb.params_derive(n, nave, 16, 1.0, fullband=True)
    nch, tchan, fch1, chanbw = p.nchanout, p.tchan, p.fch1, p.chanbw
    dms = np.linspace(0.0, W["dm_max"], W["ndm"]).astype(np.float32)
    if args.config == "c1":
        dms = (np.arange(W["ndm"]) * 0.1).astype(np.float32)
    seg_rows = int(W["seg_s"] / tchan)
    seg_data_s = seg_rows * tchan            # data-seconds consumed per step
    off, max_delay, _ = vfrb.delay_table(dms, nch, fch1, chanbw, tchan)
    win_rows = seg_rows + max_delay
    ring_rows = win_rows
    log(f"{W['workload']}: nch={nch} ndm={len(dms)} seg_rows={seg_rows} max_delay={max_delay} "
        f"window={win_rows}")

    # ---- synthetic payload (host), then resident in HBM
    rng = np.random.default_rng(W["seed"])
    pulses = []
    for _ in range(W["npulse"]):
        dm = rng.uniform(*W["pulse_dm"])
        w = rng.uniform(*W["width_ms"]) * 1e-3
        snr = rng.uniform(*W["snr"])
        t0 = rng.uniform(0.1, W["seg_s"] * 0.6)
        pulses.append((dm, t0, w, vfrb.amp_for_snr(snr, nave, nch)))
    # Narrowband tones "at 5 sigma": the tone's channel power sits 5 filterbank
    # noise sigmas above the channel mean.  Per FFT frame and pol a tone of
    # amplitude a puts (a*N/2)^2 in its bin against N*sigma^2 of noise, and
    # the averaged noise power has relative sigma 1/sqrt(2*nave), so
    # a = 2*sigma*sqrt(5/(N*sqrt(2*nave))).  (5 sigma of the voltage, 40
    # counts, would put ~5e4 x the noise into the channel and swamp the
    # channel sum of the mean-subtracted filterbank.)
    tone_amp = 2.0 * 8.0 * np.sqrt(5.0 / (n * np.sqrt(2.0 * nave)))
    tones = [(float(rng.integers(1, n // 2)), tone_amp) for _ in range(W["ntone"])]
    spec = vfrb.synth_spec(n, seed=W["seed"], pulses=pulses, tones=tones)
    nblk = seg_rows * nave
    t = time.perf_counter()
    d0 = vfrb.DeviceArray(nblk * n, np.uint8)
    d1 = vfrb.DeviceArray(nblk * n, np.uint8)
    chunk = max(1, (1 << 28) // n)  # 256 MB host chunks
    host_first = {}
    cache = args.payload_cache
    reuse = bool(cache) and os.path.exists(cache) and os.path.getsize(cache) == 2 * nblk * n
    fcache = open(cache, "rb" if reuse else "wb") if cache else None
    for pol, d in ((0, d0), (1, d1)):
        for b0 in range(0, nblk, chunk):
            nb = min(chunk, nblk - b0)
            if reuse:
                h = np.fromfile(fcache, dtype=np.uint8, count=nb * n)
            else:
                h = vfrb.synth_payload(spec, pol, b0, nb)
                if fcache:
                    h.tofile(fcache)
            if b0 == 0:
                host_first[pol] = h[: min(h.size, (int(1.0 / tchan) + 1) * nave * n)].copy()
            vfrb.call("vfrb_memcpy_h2d", d.offset(b0 * n), h.ctypes.data_as(C.c_void_p), h.nbytes, None)
    if fcache:
        fcache.close()
    log(f"{'loaded' if reuse else 'synthesised'} + uploaded {2 * nblk * n / 1e9:.2f} GB in "
        f"{time.perf_counter() - t:.1f} s")
    # plain copy of known size: the calibration kernel for the traffic counters
    calib = vfrb.DeviceArray(1 << 28, np.uint8)
    vfrb.call("vfrb_memcpy_d2d", calib.ptr, d0.ptr, 1 << 28, None)
    vfrb.sync()
    del calib

    # ---- device buffers and plans
    ring = vfrb.DeviceArray((ring_rows, nch), np.float32)
    ring.zero()
    fbp = vfrb.FilterbankPlan(n, nave, p.nskip, nch)
    out = vfrb.DeviceArray((len(dms), seg_rows), np.float32)
    ndm = len(dms)
    frb = vfrb.DeviceArray(ndm, np.int32)
    ave, sig, peak = (vfrb.DeviceArray(ndm, np.float32) for _ in range(3))
    peak_t = vfrb.DeviceArray(ndm, np.int32)
    maxc = 1 << 20
    cdm, ct = vfrb.DeviceArray(maxc, np.int32), vfrb.DeviceArray(maxc, np.int32)
    csnr, ccount = vfrb.DeviceArray(maxc, np.float32), vfrb.DeviceArray(1, np.int32)
    stream = vfrb.Stream()
    s = stream.h

    storages = [args.storage] + [x for x in args.compare.split(",") if x and x != args.storage]
    results = {}
    for sname in storages:
        sp = vfrb.SearchPlan(nch, tchan, fch1, chanbw, dms, win_rows, vfrb.STORAGE_NAMES[sname], args.tile)
        acc = vfrb.ACC_F64 if sname == "f32" else vfrb.ACC_F32C
        tune = None
        if args.tile < 0:
            # autotune the dedispersion tile on a real window (not timed)
            fbp.execute(d0, d1, seg_rows, ring, ring_rows=ring_rows, row0=max_delay, stream=s)
            sp.load(ring, win_rows, ring_rows=ring_rows, row0=0, bandpass=True, stream=s)
            tune = sp.autotune(win_rows, out, seg_rows, iters=2, stream=s)
            log(f"{sname}: autotuned tile {vfrb.tile_names()[sp.tile]}: " +
                " ".join(f"{k}={v:.2f}" for k, v in tune.items()))
        ev = [vfrb.Event() for _ in range(6)]
        names = ["filterbank", "bandpass_pack", "dedisperse", "dm_stats", "candidates"]
        stage_ms = {k: [] for k in names}
        step_ms = []

        def step(i, timed):
            new0 = (max_delay + i * seg_rows) % ring_rows
            win0 = (i * seg_rows) % ring_rows
            ev[0].record(s)
            fbp.execute(d0, d1, seg_rows, ring, ring_rows=ring_rows, row0=new0, stream=s)
            ev[1].record(s)
            sp.load(ring, win_rows, ring_rows=ring_rows, row0=win0, bandpass=True, stream=s)
            ev[2].record(s)
            sp.dedisperse(win_rows, out, seg_rows, stream=s)
            ev[3].record(s)
            vfrb.call("vfrb_dm_stats", out.ptr, seg_rows, ndm, seg_rows, acc, frb.ptr, ave.ptr, sig.ptr,
                      peak.ptr, peak_t.ptr, s)
            ev[4].record(s)
            vfrb.call("vfrb_candidates", out.ptr, seg_rows, ndm, seg_rows, ave.ptr, sig.ptr, 7.0, cdm.ptr,
                      ct.ptr, csnr.ptr, maxc, ccount.ptr, s)
            ev[5].record(s)
            if timed:
                ev[5].sync()
                for k, nm in enumerate(names):
                    stage_ms[nm].append(ev[k].elapsed_ms(ev[k + 1]))
                step_ms.append(ev[0].elapsed_ms(ev[5]))

        for i in range(args.warmup):
            step(i, False)
        stream.sync()
        vfrb.sync()
        t0 = time.perf_counter()
        for i in range(args.steps):
            step(args.warmup + i, True)
        stream.sync()
        vfrb.sync()
        wall = time.perf_counter() - t0
        ms_step = wall * 1e3 / args.steps
        rtf = args.steps * seg_data_s / wall
        ncand = int(ccount.download()[0])
        pk = peak.download()
        best = int(np.argmax(pk))
        lat = np.array(step_ms)
        results[sname] = dict(rtf=rtf, ms_per_step=ms_step, wall_s=wall,
                              stage_ms={k: float(np.mean(v)) for k, v in stage_ms.items()},
                              p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                              candidates=ncand, best_dm=float(dms[best]), best_snr=float(pk[best]),
                              tile=vfrb.tile_names()[sp.tile], autotune_ms=tune)
        log(f"{sname}: RTF={rtf:.1f} ms/step={ms_step:.2f} stages="
            + " ".join(f"{k}={np.mean(v):.3f}" for k, v in stage_ms.items())
            + f" cand={ncand} best DM={dms[best]:.1f} SNR={pk[best]:.1f}")
        del sp

    main_r = results[args.storage]
    st = main_r["stage_ms"]
    # ---- algorithmic work per launch (per step) for the roofline
    in_bytes = 2.0 * seg_rows * nave * n
    esz = {"f32": 4, "f16": 2, "u8": 1}[args.storage]
    work = {
        "filterbank": dict(bytes=in_bytes + seg_rows * nch * 4.0,
                           flops=2 * 2.5 * n * np.log2(n) * seg_rows * nave),
        "bandpass_pack": dict(bytes=win_rows * nch * (4.0 * (1 if args.storage == "f32" else 2) + esz)),
        "dedisperse": dict(adds=float(ndm) * nch * seg_rows,
                           bytes=win_rows * nch * esz + ndm * seg_rows * 4.0),
        "dm_stats": dict(bytes=3.0 * ndm * seg_rows * 4),
        "candidates": dict(bytes=ndm * seg_rows * 4.0),
    }
    dom = max(st, key=lambda k: st[k])
    dur_s = st[dom] * 1e-3
    if dom == "dedisperse":
        ach = work[dom]["adds"] / dur_s / 1e12
        roof = dict(bound="mfma", achieved=ach, peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=ach / FP32_PEAK_TFLOPS, traffic=None, kernel="dedisp_kernel",
                    note="no MFMA: FP32 vector-add bound; 1 add = 1 flop; adds = ndm*nch*rows_out")
    else:
        ach = work[dom]["bytes"] / dur_s / 1e9
        roof = dict(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s", frac=ach / HBM_PEAK_GBS,
                    traffic=None, kernel=dom)
    stages_roof = {}
    for k, v in st.items():
        d = v * 1e-3
        e = {"ms": v}
        if "bytes" in work[k]:
            e["GBps"] = work[k]["bytes"] / d / 1e9
            e["hbm_frac"] = e["GBps"] / HBM_PEAK_GBS
        if "adds" in work[k]:
            e["TFLOPs"] = work[k]["adds"] / d / 1e12
        if "flops" in work[k]:
            e["TFLOPs"] = work[k]["flops"] / d / 1e12
        stages_roof[k] = e

    # ---- CPU baseline: the oracle's FP32 stage mirror on a 1 s sample
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, n, nave, nch, tchan, dms, off, max_delay, seg_rows, win_rows, host_first)

    line = {
        "metric": "real-time factor (data-s/s) + p99 latency, 4096ch×2048 DM, FP32 vs FP16",
        "value": main_r["rtf"],
        "unit": "data-s/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded VDIF payload: N(128,8^2) noise, 8 tones at 5 sigma, 5 dispersed pulses)",
        "config": {"workload": W["workload"], "nchan": nch, "ndm": len(dms), "dm_max": W["dm_max"],
                   "fft": n, "nave": nave, "tsamp_s": tchan, "segment_s": W["seg_s"],
                   "segment_rows": seg_rows, "max_delay_rows": max_delay, "window_rows": win_rows,
                   "storage": args.storage, "accumulate": "f32",
                   "input": "VDIF payload resident in HBM (device-resident RTF)",
                   "parallelism": "single GPU"},
        "p50_latency_ms": main_r["p50_ms"],
        "p99_latency_ms": main_r["p99_ms"],
        "segment_ms": seg_data_s * 1e3,
        "rtf_by_storage": {k: v["rtf"] for k, v in results.items()},
        "p99_ms_by_storage": {k: v["p99_ms"] for k, v in results.items()},
        "stage_ms_by_storage": {k: v["stage_ms"] for k, v in results.items()},
        "candidates_by_storage": {k: v["candidates"] for k, v in results.items()},
        "tile_by_storage": {k: v["tile"] for k, v in results.items()},
        "autotune_ms_by_storage": {k: v["autotune_ms"] for k, v in results.items()},
        "stages": stages_roof,
        "roofline": roof,
        "cpu_baseline": cpu,
        "device": dev,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(args, n, nave, nch, tchan, dms, off, max_delay, seg_rows, win_rows, host_first):
    """The oracle's FP32 stage mirror timed on this host on 1 s of the same
    workload: filterbank of 976 new rows (from the first second of the
    synthesised payload), bandpass over the rows a 1 s share of the window
    carries, dedispersion of 976 output samples over a (976 + max_delay)-row
    window, and the per-DM statistics."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    if args.cpu_threads > 0:
        orc.set_threads(args.cpu_threads)
    threads = orc.num_threads()
    rows = int(1.0 / tchan)
    p0, p1 = host_first[0], host_first[1]
    t0 = time.perf_counter()
    fb = orc.filterbank_f32(p0, p1, rows, n, nave, 1, nch)
    t_fb = time.perf_counter() - t0
    bp_rows = int(round(rows * win_rows / seg_rows))
    rng = np.random.default_rng(0)
    fbw = rng.normal(256, 45, (bp_rows, nch)).astype(np.float32)
    t0 = time.perf_counter()
    orc.bandpass_f32(fbw)
    t_bp = time.perf_counter() - t0
    wrows = rows + max_delay
    fbcm = rng.normal(0, 45, (nch, wrows)).astype(np.float32)
    t0 = time.perf_counter()
    dd = orc.dedisperse_cm_f32(fbcm, off, rows)
    t_dd = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.dm_stats(dd)
    t_st = time.perf_counter() - t0
    total = t_fb + t_bp + t_dd + t_st
    log(f"cpu baseline ({threads} threads): fb={t_fb:.2f}s bp={t_bp:.2f}s dedisp={t_dd:.2f}s "
        f"stats={t_st:.2f}s -> RTF {rows * tchan / total:.3f}")
    return {"value": rows * tchan / total, "unit": "data-s/s", "cores": threads, "kind": "port",
            "sample": f"1 s of the same workload ({rows} rows): FP32 stage-mirror filterbank of the first "
                      f"second of the payload, bandpass over {bp_rows} rows, dedispersion of {rows} samples x "
                      f"{len(dms)} DMs x {nch} ch over a {wrows}-row window, per-DM stats",
            "stage_s": {"filterbank": t_fb, "bandpass": t_bp, "dedisperse": t_dd, "dm_stats": t_st}}


if __name__ == "__main__":
    main()
