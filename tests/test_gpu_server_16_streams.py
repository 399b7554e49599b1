"""GPU: the serving workload at its own shape - 16 antenna streams feeding
segments of U(0.5, 10) s into one continuously batching server at 1024
channels x 256 DMs (2048-point FFT, NAVE 64: 1.024 ms rows), FP16 window
storage with compensated FP32 statistics, device-resident input.

Every segment's result is checked against the oracle replay of its window
(the server's window rules: the stream's last max_delay + new rows), and each
stream run alone in its own server must give the same results as in the mix,
bit for bit.  The candidate lists are compared with the FP32 oracle's: every
(DM, sample) whose oracle SNR clears 7 by more than the FP16 margin is found,
and nothing below 7 by more than that margin is reported.

The payload is a 2-s synthetic block (noise N(128, 8^2), two dispersed
pulses) repeated to 10 s at whole filterbank rows, so the oracle filterbank of
the block, tiled, is the filterbank of every stream; each stream starts at its
own row of it and cuts its own segment lengths."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, NAVE = 2048, 64
NSTREAMS = 16
DMS = (np.arange(256) * 0.1).astype(np.float32)
BLOCK_ROWS = 1953          # 2.0 s of 1.024 ms rows
REPEAT = 5                 # 10.0 s of payload
MARGIN = 2e-3              # relative SNR error allowed to FP16 storage near the threshold


def _payload(vfrb, p):
    rb = N * NAVE
    amp = vfrb.amp_for_snr(20, NAVE, p.nchanout)
    spec = vfrb.synth_spec(N, seed=31, pulses=[(12.3, 0.6, 0.002, amp), (20.1, 1.5, 0.004, amp)])
    blk = [vfrb.synth_payload(spec, pol, 0, BLOCK_ROWS * NAVE)[: BLOCK_ROWS * rb] for pol in (0, 1)]
    return blk


def _plan_streams(rng, rb):
    total = BLOCK_ROWS * REPEAT * rb
    out = []
    for _ in range(NSTREAMS):
        o = int(rng.integers(0, BLOCK_ROWS)) * rb
        cuts = [0]
        while True:
            n = int(rng.uniform(0.5, 10.0) * 128e6)
            if o + cuts[-1] + n > total:
                if len(cuts) == 1:
                    cuts.append(total - o)
                break
            cuts.append(cuts[-1] + n)
        out.append((o, np.array(cuts, np.int64)))
    return out


def _server(vfrb, p, arena):
    return vfrb.Server(N, NAVE, p.nchanout, p.nskip, p.fch1, p.chanbw, p.tchan, storage=vfrb.STORE_F16,
                       acc=vfrb.ACC_F32C, threshold=7.0, max_cand=1 << 16, max_streams=NSTREAMS,
                       max_segment_bytes=int(10.0 * 128e6) + N * NAVE, arena_bytes=arena, max_batch=16,
                       batched=True)


def _collect(srv, n):
    srv.drain()
    res = {}
    while True:
        got = srv.poll(wait_ms=0)
        if not got:
            break
        for r in got:
            res[r["tag"]] = r
    assert len(res) == n
    return res


@pytest.mark.timeout(900)
def test_server_16_streams_match_oracle_and_alone(gpu, orc):
    vfrb = gpu
    p = vfrb.params_derive(N, NAVE, 1, 1.0, fullband=True)
    rb = N * NAVE
    blk = _payload(vfrb, p)
    dev = [vfrb.DeviceArray.from_host(np.tile(b, REPEAT)) for b in blk]
    streams = _plan_streams(np.random.default_rng(55), rb)
    nseg = sum(len(c) - 1 for _, c in streams)
    data_s = sum(int(c[-1]) for _, c in streams) / 128e6
    print(f"[case] 16-stream server: {nseg} segments, {data_s:.1f} data-s, FP16", flush=True)

    def submit(srv, sids, which):
        for k in range(max(len(c) - 1 for _, c in streams)):
            for i in which:
                o, c = streams[i]
                if k + 1 < len(c):
                    srv.submit_device(sids[i], dev[0], dev[1], o + int(c[k]), int(c[k + 1] - c[k]), 1000 * i + k)

    srv = _server(vfrb, p, 24 << 30)
    sids = [srv.open_stream(DMS) for _ in range(NSTREAMS)]
    maxd = srv.max_delay(sids[0])
    submit(srv, sids, range(NSTREAMS))
    mix = _collect(srv, nseg)
    stats = srv.stats()
    srv.close()
    assert stats["completed"] == nseg
    assert max(r["batch_size"] for r in mix.values()) > 1, "no iteration batched two segments"
    print(f"[case] 16-stream server: mix done, mean batch {nseg / stats['iterations']:.2f}", flush=True)

    # each stream alone
    for i in range(NSTREAMS):
        s1 = _server(vfrb, p, 3 << 30)
        sid = s1.open_stream(DMS)
        submit(s1, {i: sid}, [i])
        alone = _collect(s1, len(streams[i][1]) - 1)
        s1.close()
        for tag, a in alone.items():
            m = mix[tag]
            for key in ("status", "rows", "sample0", "nt", "ncand", "best_dm", "best_t", "best_snr"):
                assert a[key] == m[key], (tag, key)
            if a["nt"]:
                for key in ("peak", "peak_t", "ave", "sigma", "cand_dm", "cand_t", "cand_snr"):
                    assert np.array_equal(a[key], m[key]), (tag, key)
    print("[case] 16-stream server: alone == mix for every stream", flush=True)

    # the oracle replay of every window
    fb_blk = orc.filterbank_f32(blk[0], blk[1], BLOCK_ROWS, N, NAVE, p.nskip, p.nchanout)
    fb_all = np.tile(fb_blk, (REPEAT, 1))
    off, maxd_o = orc.delay_table(DMS, p.nchanout, p.fch1, p.chanbw, p.tchan)
    assert maxd == maxd_o
    searched = found = 0
    for i, (o, c) in enumerate(streams):
        r0 = o // rb
        prev = 0
        for k in range(len(c) - 1):
            r = mix[1000 * i + k]
            tot = int(c[k + 1] // rb)
            R = tot - prev
            prev = tot
            W = min(tot, maxd + R)
            nt = W - maxd
            assert r["status"] == 0 and r["rows"] == R, (i, k)
            if nt < 2:
                assert r["nt"] == 0
                continue
            searched += 1
            assert r["nt"] == nt and r["sample0"] == tot - W, (i, k)
            out = orc.dedisperse_f32(orc.bandpass_f32(fb_all[r0 + tot - W: r0 + tot]), off, nt)
            _, ave, sig, peak, pt = orc.dm_stats(out)
            assert np.max(np.abs(r["ave"] - ave) / sig) < 1e-3, (i, k)
            assert np.max(np.abs(r["sigma"] / sig - 1)) < 1e-3, (i, k)
            assert np.max(np.abs(r["peak"] - peak) / np.maximum(peak, 1.0)) < MARGIN, (i, k)
            snr = (out - ave[:, None]) / sig[:, None]
            must = set(zip(*np.nonzero(snr > 7.0 * (1 + MARGIN))))
            may = set(zip(*np.nonzero(snr > 7.0 * (1 - MARGIN))))
            got = set(zip(r["cand_dm"].tolist(), (r["cand_t"] - (tot - W)).tolist()))
            assert r["ncand"] == len(got)
            assert must <= got <= may, (i, k, len(must), len(got), len(may))
            found += len(got)
            if r["best_dm"] >= 0 and peak.max() > 10:
                bt = r["best_t"] - (tot - W)
                assert snr[r["best_dm"], bt] > peak.max() * (1 - MARGIN), (i, k)
    print(f"[case] 16-stream server: {searched} windows vs oracle, {found} candidates", flush=True)
    assert searched >= NSTREAMS and found > 0
