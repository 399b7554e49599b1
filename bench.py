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
storage (default FP32, the plain full-shape pipeline); FP16 and 8-bit
storage are measured in the same run and reported beside it.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes as C
import json
import os
import re
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
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "u8"])
    ap.add_argument("--compare", default="f32,f16,u8", help="storages also timed in this run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--tile", type=int, default=-1)
    ap.add_argument("--payload-cache", default="",
                    help="file to reuse the synthesised payload between runs (written if absent)")
    ap.add_argument("--sections", default="c1,precision,batching,serving,serving_full,stats,cli",
                    help="extra measurements reported in the same line (empty = none)")
    ap.add_argument("--input", default="frames", choices=["frames", "payload"],
                    help="device-resident input: raw VDIF frame pairs (ingested on the GPU) or de-framed payload")
    ap.add_argument("--no-e2e", action="store_true",
                    help="skip the PCIe-inclusive run (frames copied from pinned host memory every segment)")
    ap.add_argument("--serve-seconds", type=float, default=4.0,
                    help="wall seconds of each replayed arrival trace")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: libvfrb has no multi-GPU path (streams or DM ranges "
              "would shard across processes, one GPU each); rerun with --gpus 1", file=sys.stderr)
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
    frames_mode = args.input == "frames"
    npairs = -(-seg_rows * nave * n // 5000)      # VDIF frames per pol covering one segment
    nblk = -(-npairs * 5000 // n)                 # synthesised FFT blocks (covers both layouts)
    t = time.perf_counter()
    chunk = max(1, (1 << 28) // n)  # 256 MB host chunks
    keep = frames_mode or not args.no_cpu_baseline
    host_payload = {0: np.empty(nblk * n, np.uint8), 1: np.empty(nblk * n, np.uint8)} if keep else {}
    d0 = d1 = raw = None
    if not frames_mode:
        d0 = vfrb.DeviceArray(nblk * n, np.uint8)
        d1 = vfrb.DeviceArray(nblk * n, np.uint8)
    cache = args.payload_cache
    reuse = bool(cache) and os.path.exists(cache) and os.path.getsize(cache) == 2 * nblk * n
    fcache = open(cache, "rb" if reuse else "wb") if cache else None
    for pol in (0, 1):
        for b0 in range(0, nblk, chunk):
            nb = min(chunk, nblk - b0)
            if reuse:
                h = np.fromfile(fcache, dtype=np.uint8, count=nb * n)
            else:
                h = vfrb.synth_payload(spec, pol, b0, nb)
                if fcache:
                    h.tofile(fcache)
            if host_payload:
                host_payload[pol][b0 * n: (b0 + nb) * n] = h
            if not frames_mode:
                d = d0 if pol == 0 else d1
                vfrb.call("vfrb_memcpy_h2d", d.offset(b0 * n), h.ctypes.data_as(C.c_void_p), h.nbytes, None)
    if fcache:
        fcache.close()
    hraw = None
    if frames_mode:
        # VDIF frame pairs as the instrument sends them: 32-byte headers
        # (seconds from 1000, frame 0..25599, thread 0 then 1) + 5000 bytes,
        # also kept in pinned host memory for the PCIe-inclusive run
        raw = vfrb.DeviceArray(npairs * 2 * 5032, np.uint8)
        hraw = vfrb.PinnedBytes(npairs * 2 * 5032) if not args.no_e2e else None
        step = 65536
        for k0 in range(0, npairs, step):
            k = np.arange(k0, min(npairs, k0 + step))
            words = np.zeros((k.size, 8), np.uint32)
            words[:, 0] = 1000 + k // 25600
            words[:, 1] = (k % 25600) | (30 << 24)
            words[:, 2] = 629
            words[:, 3] = 0x5856 | (7 << 26)
            blk = np.empty((k.size, 2, 5032), np.uint8)
            blk[:, 0, :32] = words.view(np.uint8).reshape(k.size, 32)
            words[:, 3] |= 1 << 16
            blk[:, 1, :32] = words.view(np.uint8).reshape(k.size, 32)
            for pol in (0, 1):
                blk[:, pol, 32:] = host_payload[pol][k[0] * 5000:(k[-1] + 1) * 5000].reshape(k.size, 5000)
            vfrb.call("vfrb_memcpy_h2d", raw.offset(k0 * 2 * 5032), blk.ctypes.data_as(C.c_void_p), blk.nbytes,
                      None)
            if hraw is not None:
                hraw.array[k0 * 2 * 5032: k0 * 2 * 5032 + blk.nbytes] = blk.reshape(-1)
        ing = vfrb.Ingest(npairs)
        table = vfrb.DeviceArray(npairs + 4096, np.int32)
    log(f"{'loaded' if reuse else 'synthesised'} + uploaded {2 * nblk * n / 1e9:.2f} GB "
        f"({'VDIF frame pairs' if frames_mode else 'payload'}) in {time.perf_counter() - t:.1f} s")
    # plain copy of known size: the calibration kernel for the traffic
    # counters, and (a larger one, timed) the HBM bandwidth a copy achieves
    # (its own source buffer: the input of a small workload can be shorter)
    calib_src, calib = vfrb.DeviceArray(1 << 28, np.uint8), vfrb.DeviceArray(1 << 28, np.uint8)
    calib_src.zero()
    vfrb.call("vfrb_memcpy_d2d", calib.ptr, calib_src.ptr, 1 << 28, None)
    vfrb.sync()
    del calib, calib_src
    copy_gbs = measured_copy_bandwidth(vfrb)
    log(f"device-to-device copy: {copy_gbs:.0f} GB/s (read + write)")

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
    stream, stream2 = vfrb.Stream(), vfrb.Stream()
    s, s2 = stream.h, stream2.h

    names = (["ingest"] if frames_mode else []) + ["filterbank", "bandpass_pack", "dedisperse", "dm_stats",
                                                      "candidates"]

    def front(i, st, mark=None, frames=None, fb_ev=None):
        """Segment i's new rows into the ring: from the raw frame pairs
        through the device ingest's slot table, or from the payload.
        fb_ev: (start, end) events around the filterbank kernel."""
        row0 = (max_delay + i * seg_rows) % ring_rows
        if frames_mode:
            # every step replays the same frames: restart the parse state
            # (a live stream carries it from one batch to the next)
            fr = raw if frames is None else frames
            ing.reset()
            nsl = ing.frames(fr, npairs, table, stream=st)
            if mark is not None:
                mark.record(st)
            fb_ev and fb_ev[0].record(st)
            fbp.execute_frames(fr, table, nsl, 0, seg_rows, ring, ring_rows=ring_rows, row0=row0, stream=st)
        else:
            fb_ev and fb_ev[0].record(st)
            fbp.execute(d0, d1, seg_rows, ring, ring_rows=ring_rows, row0=row0, stream=st)
        fb_ev and fb_ev[1].record(st)

    storages = [args.storage] + [x for x in args.compare.split(",") if x and x != args.storage]
    results = {}
    for sname in storages:
        sp = vfrb.SearchPlan(nch, tchan, fch1, chanbw, dms, win_rows, vfrb.STORAGE_NAMES[sname], args.tile)
        acc = vfrb.ACC_F64 if sname == "f32" else vfrb.ACC_F32C
        tune = None
        if args.tile < 0:
            # autotune the dedispersion tile on a real window (not timed)
            front(0, s)
            sp.load(ring, win_rows, ring_rows=ring_rows, row0=0, bandpass=True, stream=s)
            tune = sp.autotune(win_rows, out, seg_rows, iters=2, stream=s)
            log(f"{sname}: autotuned tile {vfrb.tile_names()[sp.tile]}: " +
                " ".join(f"{k}={v:.2f}" for k, v in tune.items()))

        def back(i, st, marks=None):
            sp.load(ring, win_rows, ring_rows=ring_rows, row0=(i * seg_rows) % ring_rows, bandpass=True, stream=st)
            marks and marks[0].record(st)
            sp.dedisperse(win_rows, out, seg_rows, stream=st)
            marks and marks[1].record(st)
            vfrb.call("vfrb_dm_stats", out.ptr, seg_rows, ndm, seg_rows, acc, frb.ptr, ave.ptr, sig.ptr,
                      peak.ptr, peak_t.ptr, st)
            marks and marks[2].record(st)
            vfrb.call("vfrb_candidates", out.ptr, seg_rows, ndm, seg_rows, ave.ptr, sig.ptr, 7.0, cdm.ptr,
                      ct.ptr, csnr.ptr, maxc, ccount.ptr, st)

        # (1) serial steps on one stream: per-kernel times, free of overlap
        ne = len(names) + 1
        ev = [vfrb.Event() for _ in range(ne)]
        stage_ms = {k: [] for k in names}
        nserial = max(2, min(args.steps, 4))
        nf = 2 if frames_mode else 1  # front stages
        for i in range(args.warmup + nserial):
            ev[0].record(s)
            front(i, s, ev[1] if frames_mode else None)
            ev[nf].record(s)
            back(i, s, ev[nf + 1:nf + 4])
            ev[ne - 1].record(s)
            if i >= args.warmup:
                ev[ne - 1].sync()
                for k, nm in enumerate(names):
                    stage_ms[nm].append(ev[k].elapsed_ms(ev[k + 1]))
        stream.sync()
        # (2) the timed run: segments pipelined over two streams.  The front
        # stream filterbanks segment i+1 into the ring while the back stream
        # dedisperses segment i; the front waits until the back has loaded
        # window i (whose rows segment i+1 overwrites), the back waits for the
        # front's rows.  Latency of a step = its filterbank start to its
        # candidates.
        nst = args.warmup + args.steps
        ev_s = [vfrb.Event() for _ in range(nst)]
        ev_f = [vfrb.Event() for _ in range(nst)]
        ev_l = [vfrb.Event() for _ in range(nst)]
        ev_e = [vfrb.Event() for _ in range(nst)]
        # the two heavy kernels of every timed step, on the streams they run on
        ev_fb = [(vfrb.Event(), vfrb.Event()) for _ in range(nst)]
        ev_dd = [vfrb.Event() for _ in range(nst)]
        spare = vfrb.Event()
        t0 = None
        for i in range(nst):
            if i == args.warmup:
                stream.sync()
                stream2.sync()
                vfrb.sync()
                t0 = time.perf_counter()
            if i > 0:
                stream2.wait(ev_l[i - 1])
            ev_s[i].record(s2)
            front(args.warmup + nserial + i, s2, fb_ev=ev_fb[i])
            ev_f[i].record(s2)
            stream.wait(ev_f[i])
            back(args.warmup + nserial + i, s, [ev_l[i], ev_dd[i], spare])
            ev_e[i].record(s)
        stream.sync()
        stream2.sync()
        vfrb.sync()
        wall = time.perf_counter() - t0
        step_ms = [ev_s[i].elapsed_ms(ev_e[i]) for i in range(args.warmup, nst)]
        # kernel durations over the timed region (pipelined: each overlaps the other stream's work)
        timed_kernel_ms = {
            "filterbank": float(np.mean([ev_fb[i][0].elapsed_ms(ev_fb[i][1]) for i in range(args.warmup, nst)])),
            "dedisperse": float(np.mean([ev_l[i].elapsed_ms(ev_dd[i]) for i in range(args.warmup, nst)]))}
        # the last step's dedispersed series and candidates (every step sees the same window)
        out_last = out.download()
        nc_last = min(int(ccount.download()[0]), maxc)
        cand_last = set(zip(cdm.download()[:nc_last].tolist(), ct.download()[:nc_last].tolist()))
        ave_last, sig_last = ave.download(), sig.download()
        ms_step = wall * 1e3 / args.steps
        rtf = args.steps * seg_data_s / wall
        serial_ms = float(np.sum([np.mean(v) for v in stage_ms.values()]))
        ncand = int(ccount.download()[0])
        pk = peak.download()
        best = int(np.argmax(pk))
        lat = np.array(step_ms)
        e2e = None
        if sname == args.storage and hraw is not None:
            e2e = end_to_end(vfrb, args, hraw, raw, front, back, stream, stream2, seg_data_s,
                             args.warmup + nserial + nst)
        results[sname] = dict(rtf=rtf, ms_per_step=ms_step, wall_s=wall, end_to_end=e2e,
                              stage_ms={k: float(np.mean(v)) for k, v in stage_ms.items()},
                              serial_ms_per_step=serial_ms, serial_rtf=seg_data_s / (serial_ms * 1e-3),
                              p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                              candidates=ncand, best_dm=float(dms[best]), best_snr=float(pk[best]),
                              tile=vfrb.tile_names()[sp.tile], autotune_ms=tune,
                              timed_kernel_ms=timed_kernel_ms,
                              snr=(out_last - ave_last[:, None]) / sig_last[:, None], cand=cand_last)
        del out_last
        log(f"{sname}: RTF={rtf:.1f} (serial {seg_data_s / (serial_ms * 1e-3):.1f}) ms/step={ms_step:.2f} "
            f"p99={np.percentile(lat, 99):.2f} stages="
            + " ".join(f"{k}={np.mean(v):.3f}" for k, v in stage_ms.items())
            + f" cand={ncand} best DM={dms[best]:.1f} SNR={pk[best]:.1f}")
        del sp

    main_r = results[args.storage]
    st = main_r["stage_ms"]
    # ---- algorithmic work per launch (per step) for the roofline
    in_bytes = 2.0 * seg_rows * nave * n
    esz = {"f32": 4, "f16": 2, "u8": 1}[args.storage]
    work = {
        # frame headers and the slot table: 2 x 32 B read per pair + 4 B table write
        "ingest": dict(bytes=npairs * (2 * 32.0 + 4.0)),
        "filterbank": dict(bytes=in_bytes + seg_rows * nch * 4.0,
                           flops=2 * 2.5 * n * np.log2(n) * seg_rows * nave),
        "bandpass_pack": dict(bytes=win_rows * nch * (4.0 * (1 if args.storage == "f32" else 2) + esz)),
        "dedisperse": dict(adds=float(ndm) * nch * seg_rows,
                           bytes=win_rows * nch * esz + ndm * seg_rows * 4.0),
        "dm_stats": dict(bytes=3.0 * ndm * seg_rows * 4),
        "candidates": dict(bytes=ndm * seg_rows * 4.0),
    }
    # dominant kernel of the timed (pipelined) steps, timed there with HIP
    # events on the stream it runs on
    tk = main_r["timed_kernel_ms"]
    dom = max(tk, key=lambda k: tk[k])
    dur_s = tk[dom] * 1e-3
    if dom == "dedisperse":
        kind, sl, dmb, w = re.match(r"([TPG])(\d+)_dm(\d+)_w(\d+)", main_r["tile"]).groups()
        sl, dmb, w = int(sl), int(dmb), int(w)
        ts = {"f32": "float", "f16": "__half", "u8": "unsigned char"}[args.storage]
        mode = "TPG".index(kind)  # per-channel, channel-pair, 4-channel-group tile
        kname = f"dedisp_kernel<{ts}, {sl // 64}, {dmb // w}, {w}, {mode}>"
        ach = work[dom]["adds"] / dur_s / 1e12
        roof = dict(bound="valu", achieved=ach, peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=ach / FP32_PEAK_TFLOPS, kernel=kname, kernel_ms=tk[dom],
                    note="no MFMA (not a contraction): FP32 vector adds and LDS reads; 1 add = 1 flop; "
                         "adds = ndm*nch*rows_out per launch; peak = MI355X FP32 vector figure (packed FMA)")
    elif dom == "filterbank":
        # the FFT runs on the vector ALUs: its arithmetic, not its 5.5 GB of
        # HBM traffic per step, is what bounds it
        kname = f"fb_pow2_kernel<{n}, {'true' if frames_mode else 'false'}>"
        ach = work[dom]["flops"] / dur_s / 1e12
        roof = dict(bound="valu", achieved=ach, peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=ach / FP32_PEAK_TFLOPS, kernel=kname, kernel_ms=tk[dom],
                    note="no MFMA: FP32 vector FFT arithmetic; flops = 2.5*N*log2(N) per real N-point FFT, "
                         "two pols, per launch; peak = MI355X FP32 vector figure (packed FMA)")
    else:
        kname = dom
        ach = work[dom]["bytes"] / dur_s / 1e9
        roof = dict(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s", frac=ach / HBM_PEAK_GBS,
                    kernel=kname)
    roof["traffic"], roof["traffic_source"] = profiled_traffic(kname, args.storage, args.config)
    # candidate lists of the other storages against this one's, on the same window
    ref_r = results[args.storage]
    snr_ref = ref_r["snr"]
    near = np.abs(snr_ref / 7.0 - 1.0)
    cand_cmp = {"threshold": 7.0, "reference_storage": args.storage,
                "reference_candidates": len(ref_r["cand"]),
                "reference_min_margin": float(near.min()),
                "reference_min_margin_note": "smallest |SNR/7 - 1| over all (DM, sample) of the reference storage: "
                                             "the relative SNR error a candidate flip needs"}
    for k, v in results.items():
        if k == args.storage:
            continue
        rel = np.abs(v["snr"] - snr_ref) / np.maximum(np.abs(snr_ref), 1.0)
        diff = v["cand"] ^ ref_r["cand"]
        cand_cmp[k] = {"candidates": len(v["cand"]), "identical": not diff, "only_in_one": len(diff),
                       "max_rel_snr_error_vs_reference": float(rel.max()),
                       "max_rel_snr_error_above_6": float(rel[snr_ref > 6.0].max()) if (snr_ref > 6.0).any()
                       else None}
        if diff:
            d = np.array(sorted(diff))
            cand_cmp[k]["differing_reference_snr"] = [float(snr_ref[a, b]) for a, b in d[:8]]
    for v in results.values():
        del v["snr"], v["cand"]
    stages_roof = {}
    for k, v in st.items():
        d = v * 1e-3
        e = {"ms": v}
        if "bytes" in work[k]:
            e["GBps"] = work[k]["bytes"] / d / 1e9
            e["hbm_frac"] = e["GBps"] / HBM_PEAK_GBS
            e["copy_frac"] = e["GBps"] / copy_gbs
        if "adds" in work[k]:
            e["TFLOPs"] = work[k]["adds"] / d / 1e12
        if "flops" in work[k]:
            e["TFLOPs"] = work[k]["flops"] / d / 1e12
        stages_roof[k] = e

    # free the step's device buffers before the extra sections
    del d0, d1, raw, ring, out, cdm, ct, csnr
    extra = {}
    ctx = {"host_payload": host_payload}
    if "f32" in results:
        ctx["headline_fp32_rtf"] = results["f32"]["rtf"]
    for sec in [x for x in args.sections.split(",") if x]:
        fn = {"precision": bench_precision, "batching": bench_batching, "serving": bench_serving,
              "serving_full": bench_serving_full, "stats": bench_stats, "c1": bench_c1, "cli": bench_cli}[sec]
        t = time.perf_counter()
        try:
            extra[sec] = fn(vfrb, args, ctx) if sec == "serving_full" else fn(vfrb, args)
        except Exception as e:  # a failed section is reported, the headline stands
            log(f"section {sec} failed: {e!r}")
            extra[sec] = {"error": repr(e)}
        log(f"section {sec}: {time.perf_counter() - t:.1f} s")

    # ---- CPU baseline: the oracle's FP32 stage mirror on one step
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, n, nave, nch, tchan, dms, off, max_delay, seg_rows, win_rows, host_payload)

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
        "dtype": args.storage,
        "data": "synthetic (seeded VDIF payload: N(128,8^2) noise, 8 tones at 5 sigma, 5 dispersed pulses)",
        "config": {"workload": W["workload"] + f", {args.storage.upper()} storage", "nchan": nch, "ndm": len(dms), "dm_max": W["dm_max"],
                   "fft": n, "nave": nave, "tsamp_s": tchan, "segment_s": W["seg_s"],
                   "segment_rows": seg_rows, "max_delay_rows": max_delay, "window_rows": win_rows,
                   "storage": args.storage, "accumulate": "f32",
                   "input": ("raw VDIF frame pairs resident in HBM, ingested on the GPU (device-resident RTF)"
                             if frames_mode else "de-framed VDIF payload resident in HBM (device-resident RTF)"),
                   "schedule": "segments pipelined on two HIP streams (filterbank of i+1 beside dedispersion of i)",
                   "parallelism": "single GPU"},
        "end_to_end": main_r["end_to_end"],
        "p50_latency_ms": main_r["p50_ms"],
        "p99_latency_ms": main_r["p99_ms"],
        "segment_ms": seg_data_s * 1e3,
        "rtf_by_storage": {k: v["rtf"] for k, v in results.items()},
        "serial_rtf_by_storage": {k: v["serial_rtf"] for k, v in results.items()},
        "p99_ms_by_storage": {k: v["p99_ms"] for k, v in results.items()},
        "stage_ms_by_storage": {k: v["stage_ms"] for k, v in results.items()},
        "candidates_by_storage": {k: v["candidates"] for k, v in results.items()},
        "candidate_lists_vs_reference_storage": cand_cmp,
        "timed_kernel_ms_by_storage": {k: v["timed_kernel_ms"] for k, v in results.items()},
        "tile_by_storage": {k: v["tile"] for k, v in results.items()},
        "autotune_ms_by_storage": {k: v["autotune_ms"] for k, v in results.items()},
        "stages": stages_roof,
        "hbm_copy_GBps": copy_gbs,
        "roofline": roof,
        "cpu_baseline": cpu,
        "device": dev,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


def measured_copy_bandwidth(vfrb, nbytes=2 << 30, reps=5):
    """Achievable HBM bandwidth: bytes read + written per second by a
    device-to-device copy of 2 GiB (the library's copy, hipMemcpyAsync)."""
    a, b = vfrb.DeviceArray(nbytes, np.uint8), vfrb.DeviceArray(nbytes, np.uint8)
    s = vfrb.Stream()
    e0, e1 = vfrb.Event(), vfrb.Event()
    vfrb.call("vfrb_memcpy_d2d", b.ptr, a.ptr, nbytes, s.h)
    e0.record(s.h)
    for _ in range(reps):
        vfrb.call("vfrb_memcpy_d2d", b.ptr, a.ptr, nbytes, s.h)
    e1.record(s.h)
    e1.sync()
    return 2.0 * nbytes * reps / (e0.elapsed_ms(e1) * 1e-3) / 1e9


def end_to_end(vfrb, args, hraw, raw, front, back, stream, stream2, seg_data_s, i0):
    """PCIe-inclusive rate of the same pipeline: every segment's raw frame
    pairs start in pinned host memory and are copied to the device on a
    third stream (one copy per segment, two device buffers), overlapped with
    the ingest + filterbank of the previous segment and the dedispersion of
    the one before.  Never the headline value."""
    copy = vfrb.Stream()
    bufs = [raw, vfrb.DeviceArray(raw.nbytes, np.uint8)]
    nst = args.warmup + args.steps
    ev_c = [vfrb.Event() for _ in range(nst)]   # copy i landed
    ev_u = [vfrb.Event() for _ in range(nst)]   # front i done with its buffer
    ev_l = [vfrb.Event() for _ in range(nst)]   # back i loaded its window
    spare = [vfrb.Event(), vfrb.Event()]
    s, s2 = stream.h, stream2.h
    t0 = None
    for i in range(nst):
        if i == args.warmup:
            stream.sync()
            stream2.sync()
            copy.sync()
            vfrb.sync()
            t0 = time.perf_counter()
        b = bufs[i % 2]
        if i >= 2:
            copy.wait(ev_u[i - 2])
        vfrb.call("vfrb_memcpy_h2d", b.ptr, hraw.ptr, raw.nbytes, copy.h)
        ev_c[i].record(copy.h)
        stream2.wait(ev_c[i])
        if i > 0:
            stream2.wait(ev_l[i - 1])
        front(i0 + i, s2, frames=b)
        ev_u[i].record(s2)
        stream.wait(ev_u[i])
        back(i0 + i, s, [ev_l[i]] + spare)
    stream.sync()
    stream2.sync()
    copy.sync()
    vfrb.sync()
    wall = time.perf_counter() - t0
    rtf = args.steps * seg_data_s / wall
    gbs = args.steps * raw.nbytes / wall / 1e9
    log(f"end-to-end (frames from pinned host memory): RTF={rtf:.1f} data-s/s, {gbs:.1f} GB/s host-to-device")
    del bufs[1]
    return {"rtf": rtf, "ms_per_step": wall * 1e3 / args.steps, "h2d_GBps": gbs,
            "bytes_per_step": raw.nbytes,
            "input": "raw VDIF frame pairs in pinned host memory, one host-to-device copy per segment on a "
                     "third stream into two device buffers, overlapped with ingest + filterbank and dedispersion"}


def profiled_traffic(kernel, storage, config):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE and WRITE_SIZE scaled by a
    calibration copy), or None when that summary does not cover it."""
    path = os.path.join(ROOT, "profiles", "pmc_4096ch_2048dm_20s.md")
    if config != "c2" or not os.path.exists(path):
        return None, None
    for line in open(path):
        f = line.rstrip("\n").split("\t")
        if len(f) == 4 and f[0] == "traffic_bytes_per_launch" and f[1] == storage and f[2] == kernel:
            return float(f[3]), "profiles/pmc_4096ch_2048dm_20s.md"
    return None, None


def c1_shape(vfrb):
    """1024 channels (2048-point FFT), 64 spectra averaged: 1.024 ms samples."""
    n, nave = 2048, 64
    p = vfrb.params_derive(n, nave, 1, 1.0, fullband=True)
    return n, nave, p


def bench_precision(vfrb, args):
    """Precision variant of the 1 s, 1024 ch x 256 DM workload (one pulse at
    DM 12.3, FWHM 2 ms, SNR 20, seed 0): the GPU chain in f32, f16 and u8
    storage against the oracle's FP64 chain on the same payload.  Reports the
    relative error of each DM trial's peak SNR and whether the candidate
    (DM, sample) lists at 7 sigma are identical."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    n, nave, p = c1_shape(vfrb)
    nch, tchan = p.nchanout, p.tchan
    rows = int(1.0 / tchan)
    dms = (np.arange(256) * 0.1).astype(np.float32)
    amp = vfrb.amp_for_snr(20.0, nave, nch)
    spec = vfrb.synth_spec(n, seed=0, pulses=[(12.3, 0.35, 0.002, amp)])
    p0 = vfrb.synth_payload(spec, 0, 0, rows * nave)
    p1 = vfrb.synth_payload(spec, 1, 0, rows * nave)
    # FP64 oracle chain
    fb64 = orc.filterbank_f64(p0, p1, rows, n, nave, p.nskip, nch)
    off, max_delay, margin = vfrb.delay_table(dms, nch, p.fch1, p.chanbw, tchan)
    nt = rows - max_delay
    out64 = orc.dedisperse_f64(orc.bandpass_f64(fb64), off, nt)
    ave64, sig64, peak64, pt64 = orc.dm_stats_f64(out64)
    snr64 = (out64 - ave64[:, None]) / sig64[:, None]
    thr = 7.0
    ref_set = set(zip(*np.nonzero(snr64 > thr)))
    near = float(np.min(np.abs(snr64 - thr)))
    # GPU chain
    d0, d1 = vfrb.DeviceArray.from_host(p0), vfrb.DeviceArray.from_host(p1)
    fb = vfrb.DeviceArray((rows, nch), np.float32)
    vfrb.FilterbankPlan(n, nave, p.nskip, nch).execute(d0, d1, rows, fb)
    res = {"workload": "1.0 s, 1024 ch x 256 DM (0-25.5), pulse DM 12.3 SNR 20, seed 0",
           "threshold": thr, "fp64_candidates": len(ref_set),
           "fp64_min_distance_to_threshold": near, "delay_table_min_margin": margin,
           "fp64_best": {"dm": float(dms[int(np.argmax(peak64))]), "snr": float(np.max(peak64))}}
    for sname, acc in (("f32", vfrb.ACC_F64), ("f16", vfrb.ACC_F32C), ("u8", vfrb.ACC_F32C)):
        sp = vfrb.SearchPlan(nch, tchan, p.fch1, p.chanbw, dms, rows, vfrb.STORAGE_NAMES[sname])
        sp.load(fb, rows, bandpass=True)
        out = vfrb.DeviceArray((len(dms), nt), np.float32)
        sp.dedisperse(rows, out, nt)
        st = vfrb.dm_stats(out, nt, len(dms), nt, acc=acc)
        cnt, cdm, ct, _ = vfrb.candidates(out, nt, len(dms), nt, st.ave, st.sigma, thr, max_cand=1 << 16)
        got = set(zip(cdm.tolist(), ct.tolist()))
        rel = np.abs(st.peak.astype(np.float64) - peak64) / peak64
        res[sname] = {"accumulate": "f64" if acc == vfrb.ACC_F64 else "f32 compensated",
                      "peak_snr_rel_err_max": float(rel.max()),
                      "peak_snr_rel_err_p50": float(np.median(rel)),
                      "peak_snr_rel_err_at_best": float(rel[int(np.argmax(peak64))]),
                      "peak_t_equal_fraction": float(np.mean(st.peak_t == pt64)),
                      "candidates": cnt, "candidates_identical": got == ref_set,
                      "missing": len(ref_set - got), "extra": len(got - ref_set)}
        log(f"precision {sname}: max rel SNR err {rel.max():.2e} (median {np.median(rel):.2e}), "
            f"candidates {cnt} identical={got == ref_set}")
    return res


def bench_c1(vfrb, args):
    """The 1024 ch x 256 DM workload (2048-point FFT, NAVE 64: 1.024 ms rows;
    DMs 0-25.5 step 0.1; one pulse at DM 12.3, SNR 20, seed 0), 1 s segments
    of 976 new rows plus the 308-row overlap.  GPU: the chain on one stream
    (filterbank of the payload, bandpass + pack, dedispersion, stats,
    candidates), f32 and f16 storage, payload in HBM, timed over 20
    segments.  CPU: the oracle's FP32 stage mirror on the same payload bytes."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    n, nave, p = c1_shape(vfrb)
    nch, tchan = p.nchanout, p.tchan
    rows = int(1.0 / tchan)
    dms = (np.arange(256) * 0.1).astype(np.float32)
    off, max_delay, _ = vfrb.delay_table(dms, nch, p.fch1, p.chanbw, tchan)
    win = rows + max_delay
    amp = vfrb.amp_for_snr(20.0, nave, nch)
    spec = vfrb.synth_spec(n, seed=0, pulses=[(12.3, 0.35, 0.002, amp)])
    p0 = vfrb.synth_payload(spec, 0, 0, rows * nave)
    p1 = vfrb.synth_payload(spec, 1, 0, rows * nave)
    d0, d1 = vfrb.DeviceArray.from_host(p0), vfrb.DeviceArray.from_host(p1)
    ring = vfrb.DeviceArray((win, nch), np.float32)
    ring.zero()
    fbp = vfrb.FilterbankPlan(n, nave, p.nskip, nch)
    out = vfrb.DeviceArray((len(dms), rows), np.float32)
    maxc = 1 << 16
    cdm, ct = vfrb.DeviceArray(maxc, np.int32), vfrb.DeviceArray(maxc, np.int32)
    csnr, cnt = vfrb.DeviceArray(maxc, np.float32), vfrb.DeviceArray(1, np.int32)
    ave, sig, peak = (vfrb.DeviceArray(len(dms), np.float32) for _ in range(3))
    frb, pt = vfrb.DeviceArray(len(dms), np.int32), vfrb.DeviceArray(len(dms), np.int32)
    s = vfrb.Stream()
    ev0, ev1 = vfrb.Event(), vfrb.Event()
    res = {"workload": "1.0 s segments, 1024 ch x 256 DM (0-25.5), pulse DM 12.3 SNR 20, seed 0",
           "segment_rows": rows, "max_delay_rows": max_delay}
    for sname, acc in (("f32", vfrb.ACC_F64), ("f16", vfrb.ACC_F32C)):
        sp = vfrb.SearchPlan(nch, tchan, p.fch1, p.chanbw, dms, win, vfrb.STORAGE_NAMES[sname])

        def step(i):
            fbp.execute(d0, d1, rows, ring, ring_rows=win, row0=(max_delay + i * rows) % win, stream=s.h)
            sp.load(ring, win, ring_rows=win, row0=(i * rows) % win, bandpass=True, stream=s.h)
            sp.dedisperse(win, out, rows, stream=s.h)
            vfrb.call("vfrb_dm_stats", out.ptr, rows, len(dms), rows, acc, frb.ptr, ave.ptr, sig.ptr, peak.ptr,
                      pt.ptr, s.h)
            vfrb.call("vfrb_candidates", out.ptr, rows, len(dms), rows, ave.ptr, sig.ptr, 7.0, cdm.ptr, ct.ptr,
                      csnr.ptr, maxc, cnt.ptr, s.h)
        for i in range(3):
            step(i)
        # a fixed 1 s window: the autotuner picks the tile on the live window
        # (small DM blocks fill the GPU better than the default at this size)
        tune = sp.autotune(win, out, rows, iters=3, stream=s.h)
        nsteps = 20
        ev0.record(s.h)
        for i in range(nsteps):
            step(3 + i)
        ev1.record(s.h)
        ev1.sync()
        ms = ev0.elapsed_ms(ev1) / nsteps
        pk = peak.download()
        # the same chain captured once (one segment's ring positions) and
        # replayed: what launching the seven kernels one by one costs
        g = vfrb.Graph(s)
        g.begin()
        step(3)
        g.end()
        g.launch()
        ev0.record(s.h)
        for _ in range(nsteps):
            g.launch()
        ev1.record(s.h)
        ev1.sync()
        gms = ev0.elapsed_ms(ev1) / nsteps
        del g
        res[sname] = {"ms_per_segment": ms, "rtf": rows * tchan / (ms * 1e-3), "tile": vfrb.tile_names()[sp.tile],
                      "autotune_ms": {k: round(v, 4) for k, v in tune.items() if v > 0},
                      "best_dm": float(dms[int(np.argmax(pk))]), "best_snr": float(np.max(pk)),
                      "graph_ms_per_segment": gms, "graph_note": "one segment's chain captured in a hipGraph and "
                      "replayed (fixed ring positions): the launch cost of the eager chain is ms - graph_ms"}
        log(f"c1 {sname}: {ms:.3f} ms per 1 s segment, RTF {rows * tchan / (ms * 1e-3):.0f}, best DM "
            f"{dms[int(np.argmax(pk))]:.1f} SNR {np.max(pk):.1f}; graph replay {gms:.3f} ms")
        del sp
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        fb = orc.filterbank_f32(p0, p1, rows, n, nave, p.nskip, nch)
        window = np.concatenate([fb[rows - max_delay:], fb])
        fbcm = np.ascontiguousarray(orc.bandpass_f32(window).T)
        dd = orc.dedisperse_cm_f32(fbcm, off, rows)
        orc.dm_stats(dd)
        total = time.perf_counter() - t0
        res["cpu"] = {"rtf": rows * tchan / total, "seconds": total, "threads": orc.num_threads(),
                      "kind": "port", "sample": "one 1 s segment: FP32 stage-mirror filterbank, bandpass over the "
                                                "1284-row window, transpose, dedispersion, per-DM stats"}
        log(f"c1 cpu ({orc.num_threads()} threads): {total:.2f} s per 1 s segment, RTF {rows * tchan / total:.2f}")
    return res


def bench_cli(vfrb, args):
    """The drop-in program end to end: vdifFILE2dedisp_PIPELINE -d (12500-point
    FFT, NAVE 4, 32-row bundles, 6250 channels) on an 8 s VDIF file of
    uniform-noise frame pairs in a temporary directory, 101 DMs 0-100, 4 s windows.  Wall
    time of the process, start-up and plan creation included; input from the
    file through the host frame sequencer and host-to-device copies, as the
    reference reads it."""
    import subprocess
    import tempfile

    exe = os.path.join(ROOT, "vlite-frb_amd", "_build", "vdifFILE2dedisp_PIPELINE")
    if not os.path.exists(exe):
        return {"error": "vdifFILE2dedisp_PIPELINE not built"}
    seconds = 8.0
    npairs = int(seconds * 25600)
    rng = np.random.default_rng(11)
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        path = os.path.join(d, "in.vdif")
        with open(path, "wb") as f:
            step = 25600
            for k0 in range(0, npairs, step):
                k = np.arange(k0, min(npairs, k0 + step))
                words = np.zeros((k.size, 8), np.uint32)
                words[:, 0] = 1000 + k // 25600
                words[:, 1] = (k % 25600) | (30 << 24)
                words[:, 2] = 629
                blk = np.empty((k.size, 2, 5032), np.uint8)
                for th in (0, 1):
                    words[:, 3] = 0x5856 | (7 << 26) | (th << 16)
                    blk[:, th, :32] = words.view(np.uint8).reshape(k.size, 32)
                    blk[:, th, 32:] = rng.integers(96, 160, (k.size, 5000), dtype=np.uint8)
                blk.tofile(f)
        with open(os.path.join(d, "DM101.txt"), "w") as f:
            for i in range(101):
                f.write(f"{float(i):f}\n")
        runs = {}
        for mode, extra in (("fast_input", []), ("host_sequencer", ["--host-sequencer"])):
            t0 = time.perf_counter()
            r = subprocess.run([exe, "-d", "-i", path, "-o", "cli", "--window", "4"] + extra, cwd=d,
                               capture_output=True, text=True, timeout=600)
            wall = time.perf_counter() - t0
            if r.returncode != 0:
                return {"error": f"{mode}: exit {r.returncode}: {r.stderr[-500:]}"}
            log_lines = open(os.path.join(d, "cli.log")).read().splitlines()
            total = [float(x.split("=")[1].split()[0]) for x in log_lines if "Total run time" in x]
            runs[mode] = {"wall_s": wall, "in_program_s": total[0] if total else None,
                          "log_tail": [x for x in log_lines if "Processed" in x or "Total run time" in x
                                       or "fast input path" in x or "set-up" in x]}
    data_s = npairs * 5000 / 128e6
    res = {"program": "vdifFILE2dedisp_PIPELINE -d -i in.vdif --window 4", "data_s": data_s,
           "file_GB": npairs * 2 * 5032 / 1e9,
           "wall_s": runs["fast_input"]["wall_s"], "rtf": data_s / runs["fast_input"]["wall_s"],
           "note": "wall time of the process, start-up included; fast_input is the default path, "
                   "host_sequencer the reference's frame-by-frame loop (--host-sequencer)",
           "runs": {k: dict(v, rtf=data_s / v["wall_s"]) for k, v in runs.items()}}
    for k, v in runs.items():
        log(f"cli {k}: {data_s:.2f} data-s in {v['wall_s']:.3f} s wall (in program {v['in_program_s']}) -> "
            f"RTF {data_s / v['wall_s']:.1f} (file {res['file_GB']:.2f} GB)")
    return res


def bench_batching(vfrb, args):
    """16 antenna windows of uneven length (segments 2-20 s plus each stream's
    max-delay overlap) with their own DM ranges (max DM 200-1000, 512-2048
    trials), 1024 channels at 1.024 ms, one dispersed pulse each.  The
    dedispersion of all 16 windows: one launch per window vs one persistent
    batched launch (eager, and replayed from a captured hipGraph), the tile
    picked by the batch autotuner.  Inputs are filterbank windows already
    loaded (bandpassed and packed) in HBM."""
    n, nave, p = c1_shape(vfrb)
    nch, tchan = p.nchanout, p.tchan
    rng = np.random.default_rng(4)
    streams = []
    for i in range(16):
        seg_s = rng.uniform(2.0, 20.0)
        dmax = rng.uniform(200.0, 1000.0)
        ndm = int(rng.integers(512, 2049))
        dms = np.linspace(0.0, dmax, ndm).astype(np.float32)
        off, maxd, _ = vfrb.delay_table(dms, nch, p.fch1, p.chanbw, tchan)
        seg_rows = int(seg_s / tchan)
        rows = seg_rows + maxd
        fb = rng.standard_normal((rows, nch), dtype=np.float32) * 22.0 + 256.0
        d = int(ndm * 0.6)
        t0 = seg_rows // 3
        for c in range(nch):  # a 2-row pulse at DM index d, SNR ~20 after dedispersion
            fb[t0 + off[d, c]: t0 + off[d, c] + 2, c] += 13.75
        streams.append(dict(dms=dms, seg_rows=seg_rows, rows=rows, fb=vfrb.DeviceArray.from_host(fb),
                            pulse=(d, t0)))
        del fb
    data_s = sum(x["seg_rows"] for x in streams) * tchan
    s = vfrb.Stream()
    ev0, ev1 = vfrb.Event(), vfrb.Event()

    def timed(fn, reps=3):
        fn()
        ev0.record(s.h)
        for _ in range(reps):
            fn()
        ev1.record(s.h)
        ev1.sync()
        return ev0.elapsed_ms(ev1) / reps

    res = {"workload": "16 windows, segments U(2,20) s, max DM U(200,1000), 512-2048 DMs, 1024 ch",
           "data_s": data_s, "windows": [{"segment_rows": x["seg_rows"], "ndm": len(x["dms"]),
                                          "dm_max": float(x["dms"][-1])} for x in streams]}
    for sname in ("f32", "f16"):
        jobs = []
        for x in streams:
            sp = vfrb.SearchPlan(nch, tchan, p.fch1, p.chanbw, x["dms"], x["rows"], vfrb.STORAGE_NAMES[sname])
            sp.load(x["fb"], x["rows"], bandpass=True, stream=s.h)
            out = vfrb.DeviceArray((len(x["dms"]), x["seg_rows"]), np.float32)
            jobs.append((sp, x["rows"], out, x["seg_rows"]))
        b = vfrb.Batch(vfrb.STORAGE_NAMES[sname], max_jobs=16)
        b.set(jobs, stream=s.h)
        tune = b.autotune(iters=2, stream=s.h)
        tile = vfrb.tile_names()[b.tile]
        seq = timed(lambda: [sp.dedisperse(r, o, ld, stream=s.h) for sp, r, o, ld in jobs])
        eager = timed(lambda: b.dedisperse(stream=s.h))
        graph = timed(lambda: b.graph_launch(s.h))
        # the pulse of every window is found at its DM in the batched output
        found = 0
        for (sp, r, o, ld), x in zip(jobs, streams):
            d, t0 = x["pulse"]
            found += int(np.argmax(o.download()[d]) in (t0, t0 + 1))
        res[sname] = {"tile": tile, "grid": b.grid, "sequential_ms": seq, "batched_ms": eager,
                      "graph_ms": graph, "speedup_graph_vs_sequential": seq / graph,
                      "dedisp_rtf_sequential": data_s / (seq * 1e-3), "dedisp_rtf_graph": data_s / (graph * 1e-3),
                      "pulses_found": found, "autotune_ms": tune}
        log(f"batching {sname}: tile {tile} sequential {seq:.2f} ms, batched {eager:.2f} ms, graph "
            f"{graph:.2f} ms ({seq / graph:.2f}x), pulses found {found}/16")
        del b, jobs
    return res


def _device_fb(vfrb, rows, nch, rng, block_rows=4096):
    """f32 [rows][nch] noise filterbank in HBM (channel means 50-150, sigma
    10): one random block uploaded and repeated by device copies."""
    blk = (rng.standard_normal((block_rows, nch), dtype=np.float32) * 10.0
           + rng.uniform(50.0, 150.0, nch).astype(np.float32))
    d = vfrb.DeviceArray((rows, nch), np.float32)
    b = vfrb.DeviceArray.from_host(blk)
    row_b = nch * 4
    for r0 in range(0, rows, block_rows):
        nr = min(block_rows, rows - r0)
        vfrb.call("vfrb_memcpy_d2d", d.offset(r0 * row_b), b.ptr, nr * row_b, None)
    vfrb.call("vfrb_device_synchronize")
    return d


def bench_stats(vfrb, args):
    """The STATS program's window work and the dedisperseSTATS variant at
    scale.  (1) One 60 s window of 1 s statistic blocks: subtractbpCAL +
    calcstats as one fused pass (vfrb_bandpass_chan_stats) against the two
    launches, at the program's defaults (12500-point FFT, NAVE 4) and at
    4096 channels (8192-point FFT, NAVE 16); with the window's filterbank
    (timed on 2 s of payload, scaled to the window), the program's
    device-resident rate.  (2) dedisperseSTATS (mean and
    spread over channels of every (DM, t)) of a 4096 ch x 2048 DM f32 window
    of 20 s + max delay.  Inputs already in HBM."""
    rng = np.random.default_rng(7)
    s = vfrb.Stream()
    ev0, ev1 = vfrb.Event(), vfrb.Event()

    def timed(fn, reps=3):
        fn()
        ev0.record(s.h)
        for _ in range(reps):
            fn()
        ev1.record(s.h)
        ev1.sync()
        return ev0.elapsed_ms(ev1) / reps

    res = {}
    for name, (n, nave, nb, fullband) in {"stats_default_12500pt": (12500, 4, 32, False),
                                          "stats_4096ch": (8192, 16, 16, True)}.items():
        p = vfrb.params_derive(n, nave, nb, 60.0, fullband=fullband)
        rows, nch = p.maxnt, p.nchanout
        nblk, ncol = 60, int(1.0 / p.tchan)
        fb = _device_fb(vfrb, rows, nch, rng)
        outs = tuple(vfrb.DeviceArray((nblk, nch), np.float32) for _ in range(3))
        fused = timed(lambda: vfrb.bandpass_chan_stats(fb, rows, nch, ncol, nblk, stream=s.h, outs=outs))
        ref = [o.download() for o in outs]
        # two launches on a scratch copy (the in-place bandpass changes it)
        tmp = vfrb.DeviceArray((rows, nch), np.float32)

        def two():
            vfrb.call("vfrb_memcpy_d2d", tmp.ptr, fb.ptr, fb.nbytes, s.h)
            vfrb.bandpass_inplace(tmp, rows, nch, stream=s.h)
            vfrb.call("vfrb_chan_stats", tmp.ptr, rows, nch, ncol, nblk, outs[0].ptr, outs[1].ptr,
                      outs[2].ptr, s.h)
        copy_ms = timed(lambda: vfrb.call("vfrb_memcpy_d2d", tmp.ptr, fb.ptr, fb.nbytes, s.h))
        two_ms = timed(two) - copy_ms
        same = all(np.array_equal(a, o.download(), equal_nan=True) for a, o in zip(ref, outs))
        gb = rows * nch * 4 / 1e9  # compulsory: one read of the window
        del fb, tmp
        # the program's filterbank, timed on 2 s of payload resident in HBM
        # and scaled to the window (rows are independent)
        frows = int(2.0 / p.tchan)
        pay = [vfrb.DeviceArray.from_host(rng.integers(96, 160, frows * nave * n, dtype=np.uint8)) for _ in (0, 1)]
        fbp = vfrb.FilterbankPlan(n, nave, p.nskip, nch)
        fbo = vfrb.DeviceArray((frows, nch), np.float32)
        fb_ms = timed(lambda: fbp.execute(pay[0], pay[1], frows, fbo, stream=s.h)) * rows / frows
        del pay, fbo, fbp
        res[name] = {"fft": n, "nave": nave, "nchan": nch, "window_rows": rows, "stat_block_rows": ncol,
                     "window_GB": gb, "fused_ms": fused, "two_launch_ms": two_ms, "speedup": two_ms / fused,
                     "fused_GBps": gb / (fused * 1e-3), "fused_hbm_frac": gb / (fused * 1e-3) / HBM_PEAK_GBS,
                     "stage_rtf": 60.0 / (fused * 1e-3), "identical_to_two_launch": bool(same),
                     "filterbank_ms_per_window": fb_ms,
                     "program_rtf": rows * p.tchan / ((fb_ms + fused) * 1e-3)}
        log(f"stats {name}: {nch} ch x {rows} rows, fused {fused:.2f} ms ({gb / (fused * 1e-3):.0f} GB/s), "
            f"two launches {two_ms:.2f} ms, identical {same}; filterbank {fb_ms:.1f} ms per window -> "
            f"STATS program {rows * p.tchan / ((fb_ms + fused) * 1e-3):.0f} data-s/s (device-resident)")
        del outs

    # dedisperseSTATS at 4096 ch x 2048 DMs
    W = workload("c2")
    p = vfrb.params_derive(W["twonchan"], W["nave"], 16, 1.0, fullband=True)
    dms = np.linspace(0.0, W["dm_max"], W["ndm"]).astype(np.float32)
    nch, tchan = p.nchanout, p.tchan
    seg_rows = int(W["seg_s"] / tchan)
    sp = vfrb.SearchPlan(nch, tchan, p.fch1, p.chanbw, dms, seg_rows + 20000, vfrb.STORE_F32)
    rows = seg_rows + sp.max_delay
    fb = _device_fb(vfrb, rows, nch, rng)
    sp.load(fb, rows, bandpass=True, stream=s.h)
    nt = rows - sp.max_delay
    out = vfrb.DeviceArray((len(dms), nt), np.float32)
    sig = vfrb.DeviceArray((len(dms), nt), np.float32)
    ms = timed(lambda: sp.dedisperse_stats(rows, out, sig, nt, stream=s.h))
    adds = 2.0 * len(dms) * nch * nt  # a sum and a sum of squares per (DM, channel, sample)
    res["dedisperse_stats_4096ch_2048dm"] = {
        "window_rows": rows, "nt": nt, "ms": ms, "segment_s": nt * tchan, "stage_rtf": nt * tchan / (ms * 1e-3),
        "TFLOPs": adds / (ms * 1e-3) / 1e12, "fp32_frac": adds / (ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS}
    log(f"dedisperse_stats 4096ch x 2048dm: {ms:.2f} ms per {nt * tchan:.1f} s window "
        f"({nt * tchan / (ms * 1e-3):.0f} data-s/s)")
    del sp, fb, out, sig
    return res


def bench_serving(vfrb, args):
    """Request loop at the 1024 ch x 256 DM shape: 16 streams submit segments
    of U(0.5, 10) s with Poisson arrivals at aggregate offered loads of 0.5,
    1 and 2 x real time for all 16 streams (8, 16, 32 data-s/s), replayed in
    real time against the FP16 continuous-batching server and against the
    plain FP32 path (one segment at a time) on the same trace; then every
    segment of a trace submitted at once (closed loop) for the capacity.
    Payload: segments are slices of a pinned 10.5 s host pool built from one
    synthesised second (pulse at DM 12.3), so arrival-to-candidate latency
    includes the host-to-device copy."""
    n, nave, p = c1_shape(vfrb)
    nch, tchan = p.nchanout, p.tchan
    dms = (np.arange(256) * 0.1).astype(np.float32)
    rate = 128e6  # bytes per pol per data-second
    sec = int(rate)
    spec = vfrb.synth_spec(n, seed=0, pulses=[(12.3, 0.35, 0.002, vfrb.amp_for_snr(20.0, nave, nch))])
    pools = []
    pool_s = 10.5
    for pol in (0, 1):
        one = vfrb.synth_payload(spec, pol, 0, sec // n)
        pb = vfrb.PinnedBytes(int(pool_s * sec))
        reps = -(-pb.array.size // one.size)
        for k in range(reps):
            a = k * one.size
            m = min(one.size, pb.array.size - a)
            pb.array[a: a + m] = one[:m]
        pools.append(pb)
    dpools = [vfrb.DeviceArray.from_host(pb.array) for pb in pools]
    rng = np.random.default_rng(5)
    dur = args.serve_seconds

    def trace(load):
        lam = load * 16.0 / 5.25  # arrivals per second (mean length 5.25 s)
        t, ev = 0.0, []
        while True:
            t += rng.exponential(1.0 / lam)
            if t > dur:
                return ev
            length = rng.uniform(0.5, 10.0)
            nbytes = int(length * rate)
            off = int(rng.integers(0, pools[0].array.size - nbytes))
            ev.append((t, int(rng.integers(0, 16)), off, nbytes))

    def make(batched, storage):
        srv = vfrb.Server(n, nave, nch, p.nskip, p.fch1, p.chanbw, tchan, storage=storage,
                          acc=vfrb.ACC_F32C if storage != vfrb.STORE_F32 else vfrb.ACC_F64,
                          threshold=7.0, max_cand=4096, max_streams=16,
                          max_segment_bytes=int(10.0 * rate) + 1, arena_bytes=48 << 30,
                          max_batch=16, batched=batched)
        ids = [srv.open_stream(dms) for _ in range(16)]
        return srv, ids

    def replay(srv, ids, ev, realtime=True, device=True):
        t0 = time.perf_counter()
        got = []
        for k, (t, sidx, off, nb) in enumerate(ev):
            if realtime:
                while time.perf_counter() - t0 < t:
                    got += srv.poll(wait_ms=1)
            if device:
                srv.submit_device(ids[sidx], dpools[0], dpools[1], off, nb, k, outputs=False)
            else:
                srv.submit(ids[sidx], pools[0].array[off: off + nb], pools[1].array[off: off + nb], k,
                           outputs=False)
        srv.drain()
        wall = time.perf_counter() - t0
        got += srv.poll(wait_ms=0, max_results=4096)
        lat = np.array([r["done_s"] - r["arrival_s"] for r in got])
        # throughput counts the segments that completed without error
        seglen = np.array([ev[r["tag"]][3] / rate if r["status"] == 0 else 0.0 for r in got])
        bs = np.array([r["batch_size"] for r in got])
        st = srv.stats()
        return {"segments": len(got), "data_s": float(seglen.sum()), "wall_s": wall,
                "throughput_data_s_per_s": float(seglen.sum() / wall),
                "p50_latency_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_latency_ms": float(np.percentile(lat, 99) * 1e3),
                "max_latency_ms": float(lat.max() * 1e3),
                "p99_latency_over_segment": float(np.percentile(lat / seglen, 99)),
                "mean_batch": float(bs.mean()), "iterations": st["iterations"],
                "device_busy_s": st["device_busy_s"], "errors": int(sum(r["status"] != 0 for r in got))}

    res = {"workload": "16 streams, 1024 ch x 256 DM, segments U(0.5,10) s, Poisson arrivals",
           "trace_wall_s": dur,
           "input": {"device": "payload resident in HBM (NIC-to-GPU path; primary)",
                     "host": "payload in pinned host memory, copied per segment (PCIe-inclusive)"}}
    modes = (("fp16_batched", True, vfrb.STORE_F16), ("fp32_plain", False, vfrb.STORE_F32))
    for inp in ("device", "host"):
        for load in (0.5, 1.0, 2.0):
            ev = trace(load)
            for name, batched, storage in modes:
                srv, ids = make(batched, storage)
                r = replay(srv, ids, ev, device=inp == "device")
                srv.close()
                res.setdefault(inp, {}).setdefault(f"load_{load:g}x", {})[name] = r
                log(f"serving {inp} load {load}x {name}: {r['segments']} segs, p50 "
                    f"{r['p50_latency_ms']:.1f} ms, p99 {r['p99_latency_ms']:.1f} ms, thr "
                    f"{r['throughput_data_s_per_s']:.1f} data-s/s, mean batch {r['mean_batch']:.2f}"
                    + (f", {r['errors']} SEGMENTS FAILED" if r["errors"] else ""))
        # capacity: a 2x trace submitted at once
        ev = trace(2.0)
        for name, batched, storage in modes:
            srv, ids = make(batched, storage)
            r = replay(srv, ids, ev, realtime=False, device=inp == "device")
            srv.close()
            res[inp].setdefault("closed_loop", {})[name] = r
            log(f"serving {inp} closed loop {name}: {r['throughput_data_s_per_s']:.1f} data-s/s, "
                f"p99 {r['p99_latency_ms']:.1f} ms, mean batch {r['mean_batch']:.2f}"
                + (f", {r['errors']} SEGMENTS FAILED" if r["errors"] else ""))
        c = res[inp]["closed_loop"]
        res[inp]["capacity_ratio_fp16_batched_vs_fp32_plain"] = (
            c["fp16_batched"]["throughput_data_s_per_s"] / c["fp32_plain"]["throughput_data_s_per_s"])
    return res


def bench_serving_full(vfrb, args, ctx):
    """The serving comparison at the full shape: 4096 ch x 2048 DM (0-1000),
    8192-point FFT, NAVE 16, 20 s segments, 8 antenna streams.  The FP16
    continuous-batching server against the plain FP32 path (one segment at a
    time) on the same traces: closed loop (every segment submitted at once:
    capacity) and Poisson arrivals at 50 % and 90 % of the plain path's
    capacity.  Input: the headline workload's payload resident in HBM, read in
    place by the filterbank (device-resident)."""
    host_payload = ctx["host_payload"]
    n, nave = 8192, 16
    p = vfrb.params_derive(n, nave, 1, 1.0, fullband=True)
    nch, tchan = p.nchanout, p.tchan
    dms = np.linspace(0.0, 1000.0, 2048).astype(np.float32)
    rb = n * nave
    seg_rows = int(20.0 / tchan)
    nbytes = seg_rows * rb
    assert host_payload[0].size >= nbytes
    pools = [vfrb.DeviceArray.from_host(host_payload[pol][:nbytes]) for pol in (0, 1)]
    nstreams, per_stream = 8, 3
    seg_s = seg_rows * tchan
    # the tile the autotuner keeps on 20 s windows at this shape, for both servers
    tile = vfrb.tile_names().index("W504_dm64_w16")

    def make(batched, storage):
        srv = vfrb.Server(n, nave, nch, p.nskip, p.fch1, p.chanbw, tchan, storage=storage,
                          acc=vfrb.ACC_F32C if storage != vfrb.STORE_F32 else vfrb.ACC_F64,
                          threshold=7.0, max_cand=1 << 16, max_streams=nstreams, max_segment_bytes=nbytes,
                          arena_bytes=8 << 30, max_batch=4, batched=batched, tile=tile)
        ids = [srv.open_stream(dms) for _ in range(nstreams)]
        return srv, ids

    def run(srv, ids, times):
        """times: arrival offsets (s) of the segments, stream-interleaved; None = all at once."""
        nseg = nstreams * per_stream
        t0 = time.perf_counter()
        got = []
        for k in range(nseg):
            if times is not None:
                while time.perf_counter() - t0 < times[k]:
                    got += srv.poll(wait_ms=1)
            srv.submit_device(ids[k % nstreams], pools[0], pools[1], 0, nbytes, k, outputs=False)
        srv.drain()
        wall = time.perf_counter() - t0
        got += srv.poll(wait_ms=0, max_results=4096)
        lat = np.array([r["done_s"] - r["arrival_s"] for r in got])
        bs = np.array([r["batch_size"] for r in got])
        st = srv.stats()
        nok = sum(r["status"] == 0 for r in got)  # throughput counts error-free segments
        return {"segments": len(got), "data_s": nok * seg_s, "wall_s": wall,
                "throughput_data_s_per_s": nok * seg_s / wall,
                "p50_latency_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_latency_ms": float(np.percentile(lat, 99) * 1e3),
                "p99_latency_over_segment": float(np.percentile(lat, 99) / seg_s),
                "mean_batch": float(bs.mean()), "iterations": st["iterations"],
                "device_busy_s": st["device_busy_s"], "errors": int(sum(r["status"] != 0 for r in got))}

    modes = (("fp16_batched", True, vfrb.STORE_F16), ("fp32_plain", False, vfrb.STORE_F32))
    res = {"workload": f"{nstreams} streams x {per_stream} segments of {seg_s:.2f} s, 4096 ch x 2048 DM "
                       "(0-1000), device-resident payload read in place, tile W504_dm64_w16",
           "closed_loop": {}}
    for name, batched, storage in modes:
        srv, ids = make(batched, storage)
        run(srv, ids, None)  # warm-up: plans, graphs, first-touch
        srv.close()
        srv, ids = make(batched, storage)
        r = run(srv, ids, None)
        srv.close()
        res["closed_loop"][name] = r
        log(f"serving 4096x2048 closed loop {name}: {r['throughput_data_s_per_s']:.1f} data-s/s, "
            f"p99 {r['p99_latency_ms']:.1f} ms, mean batch {r['mean_batch']:.2f}"
            + (f", {r['errors']} SEGMENTS FAILED" if r["errors"] else ""))
    cap = res["closed_loop"]["fp32_plain"]["throughput_data_s_per_s"]
    res["capacity_ratio_fp16_batched_vs_fp32_plain"] = (
        res["closed_loop"]["fp16_batched"]["throughput_data_s_per_s"] / cap)
    rng = np.random.default_rng(8)
    for frac in (0.5, 0.9):
        lam = frac * cap / seg_s  # segments per second
        times = np.cumsum(rng.exponential(1.0 / lam, nstreams * per_stream))
        for name, batched, storage in modes:
            srv, ids = make(batched, storage)
            r = run(srv, ids, times)
            srv.close()
            r["offered_data_s_per_s"] = frac * cap
            res.setdefault(f"poisson_{int(frac * 100)}pct_of_fp32_capacity", {})[name] = r
            log(f"serving 4096x2048 at {frac:.0%} of FP32 capacity {name}: p50 {r['p50_latency_ms']:.1f} ms, "
                f"p99 {r['p99_latency_ms']:.1f} ms, mean batch {r['mean_batch']:.2f}"
                + (f", {r['errors']} SEGMENTS FAILED" if r["errors"] else ""))
    if "headline_fp32_rtf" in ctx:
        res["ratio_fp16_batched_server_vs_headline_fp32_pipeline"] = (
            res["closed_loop"]["fp16_batched"]["throughput_data_s_per_s"] / ctx["headline_fp32_rtf"])
    del pools
    return res


def cpu_baseline(args, n, nave, nch, tchan, dms, off, max_delay, seg_rows, win_rows, host_payload):
    """The oracle's FP32 stage mirror timed on this host on one step of the
    same workload: filterbank of the segment's payload, bandpass over the
    (max_delay + segment)-row window, transposition to channel-major,
    dedispersion of the segment's output samples, per-DM statistics.  The
    carried overlap rows are stood in for by the segment's last rows (same
    size and cost)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    if args.cpu_threads > 0:
        orc.set_threads(args.cpu_threads)
    threads = orc.num_threads()
    rows = seg_rows
    p0, p1 = host_payload[0], host_payload[1]
    t0 = time.perf_counter()
    fb = orc.filterbank_f32(p0, p1, rows, n, nave, 1, nch)
    t_fb = time.perf_counter() - t0
    window = np.concatenate([fb[rows - max_delay:], fb])
    del fb
    t0 = time.perf_counter()
    bp = orc.bandpass_f32(window)
    t_bp = time.perf_counter() - t0
    t0 = time.perf_counter()
    fbcm = np.ascontiguousarray(bp.T)
    t_tr = time.perf_counter() - t0
    del bp, window
    t0 = time.perf_counter()
    dd = orc.dedisperse_cm_f32(fbcm, off, rows)
    t_dd = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.dm_stats(dd)
    t_st = time.perf_counter() - t0
    total = t_fb + t_bp + t_tr + t_dd + t_st
    log(f"cpu baseline ({threads} threads): fb={t_fb:.2f}s bp={t_bp:.2f}s transpose={t_tr:.2f}s "
        f"dedisp={t_dd:.2f}s stats={t_st:.2f}s -> RTF {rows * tchan / total:.3f}")
    return {"value": rows * tchan / total, "unit": "data-s/s", "cores": threads, "kind": "port",
            "sample": f"one step of the same workload ({rows} new rows = {rows * tchan:.2f} data-s): FP32 "
                      f"stage-mirror filterbank of the segment payload, bandpass over the {win_rows}-row "
                      f"window, transpose, dedispersion of {rows} samples x {len(dms)} DMs x {nch} ch, "
                      f"per-DM stats; {total:.1f} s of CPU work",
            "stage_s": {"filterbank": t_fb, "bandpass": t_bp, "transpose": t_tr, "dedisperse": t_dd,
                        "dm_stats": t_st}}


if __name__ == "__main__":
    main()
