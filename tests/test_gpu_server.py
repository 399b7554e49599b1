"""GPU: the segment server (vfrb_server_*) against the oracle replay of its
window rules, and the same stream alone vs in a mix of streams."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, NAVE = 2048, 8
DMS = (np.arange(32) * 0.1).astype(np.float32)
# uneven segment lengths in rows (fractions leave partial rows for the carry)
SEG_ROWS = [350.5, 200.25, 611.75, 90.1, 480.0, 260.6]


def _shape(vfrb):
    return vfrb.params_derive(N, NAVE, 1, 1.0, fullband=True)


def _stream_bytes(vfrb, p, seed, pulse_t0, snr=30, fwhm=0.002):
    nbytes = [int(r * N * NAVE) for r in SEG_ROWS]
    total = sum(nbytes)
    nblk = -(-total // N)
    amp = vfrb.amp_for_snr(snr, NAVE, p.nchanout)
    spec = vfrb.synth_spec(N, seed=seed, pulses=[(2.0, pulse_t0, fwhm, amp)])
    p0 = vfrb.synth_payload(spec, 0, 0, nblk)[:total]
    p1 = vfrb.synth_payload(spec, 1, 0, nblk)[:total]
    cuts = np.concatenate([[0], np.cumsum(nbytes)])
    return p0, p1, cuts


def _server(vfrb, p, storage, batched, **kw):
    return vfrb.Server(N, NAVE, p.nchanout, p.nskip, p.fch1, p.chanbw, p.tchan, storage=storage,
                       acc=kw.pop("acc", vfrb.ACC_F64), threshold=7.0, max_cand=8192,
                       max_streams=kw.pop("max_streams", 8),
                       max_segment_bytes=kw.pop("max_segment_bytes", 16 << 20),
                       arena_bytes=kw.pop("arena_bytes", 1 << 28), max_batch=8, batched=batched, **kw)


def _run(srv, streams, device=False):
    """streams: list of (sid, p0, p1, cuts, tag_base); submits segment k of
    every stream before segment k+1 of any, then collects all results.
    device=True uploads each stream's payload once and submits device slices."""
    nseg = len(SEG_ROWS)
    if device:
        import vfrb
        dev = {sid: (vfrb.DeviceArray.from_host(p0), vfrb.DeviceArray.from_host(p1))
               for sid, p0, p1, _, _ in streams}
    for k in range(nseg):
        for sid, p0, p1, cuts, base in streams:
            a, b = cuts[k], cuts[k + 1]
            if device:
                srv.submit_device(sid, dev[sid][0], dev[sid][1], int(a), int(b - a), base + k)
            else:
                srv.submit(sid, p0[a:b], p1[a:b], base + k)
    srv.drain()
    res = {}
    while True:
        got = srv.poll(wait_ms=0)
        if not got:
            break
        for r in got:
            res[r["tag"]] = r
    assert len(res) == nseg * len(streams)
    return res


def test_server_windows_match_oracle_replay(gpu, orc):
    vfrb = gpu
    p = _shape(vfrb)
    print("[case] server f32 plain vs oracle replay", flush=True)
    # a narrow, strong pulse, so that the last window's best DM and time are checked
    p0, p1, cuts = _stream_bytes(vfrb, p, seed=3, pulse_t0=0.19, snr=120, fwhm=0.0003)
    srv = _server(vfrb, p, vfrb.STORE_F32, batched=False)
    sid = srv.open_stream(DMS)
    maxd = srv.max_delay(sid)
    res = _run(srv, [(sid, p0, p1, cuts, 0)])
    srv.close()
    rb = N * NAVE
    total_rows = int(cuts[-1] // rb)
    fb = orc.filterbank_f32(p0, p1, total_rows, N, NAVE, p.nskip, p.nchanout)
    off, maxd_o = orc.delay_table(DMS, p.nchanout, p.fch1, p.chanbw, p.tchan)
    assert maxd == maxd_o
    prev = 0
    searched = pulsed = 0
    for k in range(len(SEG_ROWS)):
        r = res[k]
        tot = int(cuts[k + 1] // rb)
        R = tot - prev
        prev = tot
        W = min(tot, maxd + R)
        nt = W - maxd
        assert r["status"] == 0 and r["rows"] == R
        if nt < 2:
            assert r["nt"] == 0 and r["best_dm"] == -1
            continue
        searched += 1
        assert r["nt"] == nt and r["sample0"] == tot - W
        out = orc.dedisperse_f32(orc.bandpass_f32(fb[tot - W: tot]), off, nt)
        _, ave, sig, peak, pt = orc.dm_stats(out)
        assert np.max(np.abs(r["peak"] - peak) / np.maximum(peak, 1.0)) < 2e-3
        assert np.max(np.abs(r["ave"] - ave) / sig) < 1e-3
        assert np.max(np.abs(r["sigma"] / sig - 1)) < 1e-3
        b = int(np.argmax(peak))
        if peak[b] > 10:  # the pulse is inside this window
            assert r["best_dm"] == b and r["best_t"] == pt[b] + tot - W
            pulsed += 1
    assert searched >= 4 and pulsed == 1


@pytest.mark.parametrize("storage", ["f32", "f16", "u8"])
def test_server_result_is_the_same_alone_and_in_the_mix(gpu, storage):
    vfrb = gpu
    p = _shape(vfrb)
    a = _stream_bytes(vfrb, p, seed=5, pulse_t0=0.12)
    others = [_stream_bytes(vfrb, p, seed=s, pulse_t0=0.05 * s) for s in (6, 7, 8)]
    print(f"[case] server {storage}: stream alone (batched)", flush=True)
    s1 = _server(vfrb, p, vfrb.STORAGE_NAMES[storage], batched=True)
    sid = s1.open_stream(DMS)
    alone = _run(s1, [(sid, *a, 0)])
    s1.close()
    print(f"[case] server {storage}: stream alone (plain)", flush=True)
    s2 = _server(vfrb, p, vfrb.STORAGE_NAMES[storage], batched=False)
    sid = s2.open_stream(DMS)
    plain = _run(s2, [(sid, *a, 0)])
    s2.close()
    print(f"[case] server {storage}: stream alone (batched, device-resident payload)", flush=True)
    s4 = _server(vfrb, p, vfrb.STORAGE_NAMES[storage], batched=True)
    sid = s4.open_stream(DMS)
    dev = _run(s4, [(sid, *a, 0)], device=True)
    s4.close()
    print(f"[case] server {storage}: stream in a mix of four", flush=True)
    s3 = _server(vfrb, p, vfrb.STORAGE_NAMES[storage], batched=True)
    ids = [s3.open_stream(DMS) for _ in range(4)]
    # stream A is the second one opened; the others get their own DM lists
    streams = [(ids[0], *others[0], 100), (ids[1], *a, 0), (ids[2], *others[1], 200),
               (ids[3], *others[2], 300)]
    mix = _run(s3, streams)
    stats = s3.stats()
    s3.close()
    assert stats["completed"] == 4 * len(SEG_ROWS)
    assert max(r["batch_size"] for r in mix.values()) > 1, "no iteration batched two segments"
    keys = ("status", "rows", "sample0", "nt", "ncand", "best_dm", "best_t", "best_snr")
    arrays = ("peak", "peak_t", "ave", "sigma", "cand_dm", "cand_t", "cand_snr")
    for k in range(len(SEG_ROWS)):
        for other in (plain[k], mix[k], dev[k]):
            for key in keys:
                assert alone[k][key] == other[key], (k, key)
            if alone[k]["nt"]:
                for key in arrays:
                    assert np.array_equal(alone[k][key], other[key]), (k, key)
    assert any(alone[k]["ncand"] > 0 for k in alone), "the pulse gave no candidate"
    for r in mix.values():
        assert r["arrival_s"] <= r["start_s"] <= r["done_s"]


def test_server_admission_and_errors(gpu):
    vfrb = gpu
    p = _shape(vfrb)
    rb = N * NAVE
    print("[case] server admission limits", flush=True)
    # an arena that holds one 300-row segment (both pols) at a time
    arena = 2 * (300 * rb + rb + 4096)
    srv = _server(vfrb, p, vfrb.STORE_F16, batched=True, arena_bytes=arena, max_streams=2)
    s0, s1 = srv.open_stream(DMS), srv.open_stream(DMS)
    with pytest.raises(vfrb.VfrbError) as e:
        srv.open_stream(DMS)
    assert e.value.status == vfrb.ERR_CAPACITY
    seg = np.full(300 * rb, 128, np.uint8)
    with pytest.raises(vfrb.VfrbError) as e:
        srv.submit(s0, np.zeros(17 << 20, np.uint8), np.zeros(17 << 20, np.uint8), 99)
    assert e.value.status == vfrb.ERR_CAPACITY
    with pytest.raises(vfrb.VfrbError):
        srv.submit(7, seg, seg, 98)
    for k in range(4):
        srv.submit(s0 if k % 2 == 0 else s1, seg, seg, k)
    srv.drain()
    got = []
    while True:
        r = srv.poll(wait_ms=0)
        if not r:
            break
        got += r
    st = srv.stats()
    srv.close()
    assert sorted(r["tag"] for r in got) == [0, 1, 2, 3]
    assert all(r["status"] == 0 and r["batch_size"] == 1 for r in got)
    assert st["arena_peak_bytes"] <= arena and st["iterations"] == 4


# Per-stream DM lists at this shape (2048-point FFT, NAVE 8: 0.128 ms rows)
# that need different tiles: DM step 0.25 fits the pair tile but not the
# group tile, step 0.4 only the per-channel tile, DMS the default group tile.
TILE_LISTS = [("P512_dm64_w16", (np.arange(64) * 0.25).astype(np.float32)),
              ("T512_dm64_w16", (np.arange(64) * 0.4).astype(np.float32)),
              ("G256_dm64_w16", DMS)]
LONG_ROWS = [900.5, 1100.25, 700.75, 1300.0, 650.1, 1250.6]


def _long_stream(vfrb, p, seed):
    nbytes = [int(r * N * NAVE) for r in LONG_ROWS]
    total = sum(nbytes)
    nblk = -(-total // N)
    amp = vfrb.amp_for_snr(30, NAVE, p.nchanout)
    spec = vfrb.synth_spec(N, seed=seed, pulses=[(3.0, 0.45, 0.002, amp)])
    p0 = vfrb.synth_payload(spec, 0, 0, nblk)[:total]
    p1 = vfrb.synth_payload(spec, 1, 0, nblk)[:total]
    return p0, p1, np.concatenate([[0], np.cumsum(nbytes)])


def _run_long(srv, streams):
    for k in range(len(LONG_ROWS)):
        for sid, p0, p1, cuts, base in streams:
            a, b = cuts[k], cuts[k + 1]
            srv.submit(sid, p0[a:b], p1[a:b], base + k)
    srv.drain()
    res = {}
    while True:
        got = srv.poll(wait_ms=0)
        if not got:
            break
        for r in got:
            res[r["tag"]] = r
    assert len(res) == len(LONG_ROWS) * len(streams)
    return res


def test_server_per_stream_dm_lists_keep_their_tiles(gpu, orc):
    """Streams whose DM lists need different dedispersion tiles share one
    2-lane batched server.  Each iteration launches one persistent batch per
    tile in use, so every segment runs on its own plan's tile: each stream's
    results equal the stream alone bit for bit, and the oracle replay."""
    vfrb = gpu
    p = _shape(vfrb)
    names = vfrb.tile_names()
    for want, dms in TILE_LISTS:
        assert names[vfrb.SearchPlan(p.nchanout, p.tchan, p.fch1, p.chanbw, dms, 4000).tile] == want
    data = [_long_stream(vfrb, p, seed=40 + i) for i in range(len(TILE_LISTS))]
    alone = []
    for (want, dms), (p0, p1, cuts) in zip(TILE_LISTS, data):
        print(f"[case] server per-stream lists: {want} alone", flush=True)
        srv = _server(vfrb, p, vfrb.STORE_F32, batched=True, arena_bytes=1 << 30, max_segment_bytes=32 << 20)
        sid = srv.open_stream(dms)
        alone.append(_run_long(srv, [(sid, p0, p1, cuts, 0)]))
        srv.close()
    print("[case] server per-stream lists: all three in one server", flush=True)
    srv = _server(vfrb, p, vfrb.STORE_F32, batched=True, arena_bytes=1 << 30, max_segment_bytes=32 << 20)
    ids = [srv.open_stream(dms) for _, dms in TILE_LISTS]
    mix = _run_long(srv, [(sid, *d, 100 * i) for i, (sid, d) in enumerate(zip(ids, data))])
    stats = srv.stats()
    srv.close()
    assert stats["completed"] == 3 * len(LONG_ROWS)
    assert max(r["batch_size"] for r in mix.values()) > 1, "no iteration batched two segments"
    keys = ("status", "rows", "sample0", "nt", "ncand", "best_dm", "best_t", "best_snr")
    arrays = ("peak", "peak_t", "ave", "sigma", "cand_dm", "cand_t", "cand_snr")
    searched = 0
    for i, ((want, dms), (p0, p1, cuts)) in enumerate(zip(TILE_LISTS, data)):
        for k in range(len(LONG_ROWS)):
            a, m = alone[i][k], mix[100 * i + k]
            assert a["status"] == 0, (want, k, a["status"])
            for key in keys:
                assert a[key] == m[key], (want, k, key)
            if a["nt"]:
                searched += 1
                for key in arrays:
                    assert np.array_equal(a[key], m[key]), (want, k, key)
        # the oracle replay of the stream's last window
        rb = N * NAVE
        tot = int(cuts[-1] // rb)
        off, maxd = orc.delay_table(dms, p.nchanout, p.fch1, p.chanbw, p.tchan)
        R = tot - int(cuts[-2] // rb)
        W = min(tot, maxd + R)
        r = alone[i][len(LONG_ROWS) - 1]
        assert r["nt"] == W - maxd and r["sample0"] == tot - W
        fb = orc.filterbank_f32(p0, p1, tot, N, NAVE, p.nskip, p.nchanout)
        out = orc.dedisperse_f32(orc.bandpass_f32(fb[tot - W: tot]), off, W - maxd)
        _, ave, sig, peak, pt = orc.dm_stats(out)
        assert np.max(np.abs(r["ave"] - ave) / sig) < 1e-3
        assert np.max(np.abs(r["sigma"] / sig - 1)) < 1e-3
        assert np.max(np.abs(r["peak"] - peak) / np.maximum(peak, 1.0)) < 2e-3
    assert searched >= 9


def test_server_device_input_in_place_equals_copy(gpu):
    """Device-resident segments are filterbanked in place when their rows start
    4-byte aligned (a carried partial row is completed in the stream's carry
    buffer first); other segments are copied into the arena.  Segments that
    start a stream with less than a row, complete a carried row exactly, or
    leave odd byte counts mix both paths; results equal the host-input
    (copy) path bit for bit."""
    vfrb = gpu
    p = _shape(vfrb)
    rb = N * NAVE
    # carries: 8192, 0, 1000, 6, 9, 26, 126 (a segment short of the carried
    # row: copied), 106, 106
    nbytes = [rb // 2, rb // 2, 300 * rb + 1000, 250 * rb - 1000 + 6, 2 * rb + 3, 400 * rb + 17, 100, rb - 20,
              350 * rb]
    total = sum(nbytes)
    amp = vfrb.amp_for_snr(30, NAVE, p.nchanout)
    spec = vfrb.synth_spec(N, seed=9, pulses=[(2.0, 0.1, 0.002, amp)])
    nblk = -(-total // N)
    p0 = vfrb.synth_payload(spec, 0, 0, nblk)[:total]
    p1 = vfrb.synth_payload(spec, 1, 0, nblk)[:total]
    cuts = np.concatenate([[0], np.cumsum(nbytes)])
    res = []
    # host input; device input in place (rows 16- and 4-byte aligned); device
    # input one byte off alignment (every segment through the arena copy)
    for device, pad in ((False, 0), (True, 0), (True, 4), (True, 1)):
        print(f"[case] server device input in place: device={device} pad={pad}", flush=True)
        srv = _server(vfrb, p, vfrb.STORE_F16, batched=True)
        sid = srv.open_stream(DMS)
        if device:
            z = np.zeros(pad, np.uint8)
            d0 = vfrb.DeviceArray.from_host(np.concatenate([z, p0]))
            d1 = vfrb.DeviceArray.from_host(np.concatenate([z, p1]))
        for k in range(len(nbytes)):
            a, b = int(cuts[k]), int(cuts[k + 1])
            if device:
                srv.submit_device(sid, d0, d1, pad + a, b - a, k)
            else:
                srv.submit(sid, p0[a:b], p1[a:b], k)
        srv.drain()
        got = {}
        while True:
            r = srv.poll(wait_ms=0)
            if not r:
                break
            got.update({x["tag"]: x for x in r})
        srv.close()
        res.append(got)
    host = res[0]
    assert sum(host[k]["nt"] > 0 for k in host) >= 3
    for dev in res[1:]:
        assert sorted(host) == sorted(dev) == list(range(len(nbytes)))
        for k in host:
            for key in ("status", "rows", "sample0", "nt", "ncand", "best_dm", "best_t", "best_snr"):
                assert host[k][key] == dev[k][key], (k, key)
            if host[k]["nt"]:
                for key in ("peak", "peak_t", "ave", "sigma", "cand_dm", "cand_t", "cand_snr"):
                    assert np.array_equal(host[k][key], dev[k][key]), (k, key)


def _collect(srv, n):
    srv.drain()
    res = {}
    while len(res) < n:
        got = srv.poll(wait_ms=0)
        assert got, f"{len(res)} of {n} results"
        for r in got:
            res[r["tag"]] = r
    return res


def test_server_stream_state_save_and_restore(gpu):
    """A stream's state saved after two segments (a partial row carried) and
    restored into a stream of a new server (batched, another segment limit,
    so another ring length) continues as the saved stream does: the next
    four segments give the same windows, per-DM results and candidates bit
    for bit.  A state restores only into a stream with the same DM list and
    band; a truncated state and a short buffer are refused."""
    vfrb = gpu
    p = _shape(vfrb)
    # a short window's own spread caps a pulse's SNR near sqrt(nt / width):
    # a narrow, strong pulse clears the 7-sigma threshold in the last window
    p0, p1, cuts = _stream_bytes(vfrb, p, seed=5, pulse_t0=0.19, snr=120, fwhm=0.0003)
    nseg, k_save = len(SEG_ROWS), 2
    rb = N * NAVE
    assert cuts[k_save] % rb != 0, "a partial row is carried at the save point"
    print("[case] server stream state: uninterrupted stream", flush=True)
    a = _server(vfrb, p, vfrb.STORE_F32, batched=False)
    sa = a.open_stream(DMS)
    for k in range(k_save):
        a.submit(sa, p0[cuts[k]:cuts[k + 1]], p1[cuts[k]:cuts[k + 1]], k)
    state = a.save_stream(sa)
    maxd = a.max_delay(sa)
    assert len(state) > maxd * p.nchanout * 4
    for k in range(k_save, nseg):
        a.submit(sa, p0[cuts[k]:cuts[k + 1]], p1[cuts[k]:cuts[k + 1]], k)
    ref = _collect(a, nseg)
    print("[case] server stream state: restored into a new server", flush=True)
    b = _server(vfrb, p, vfrb.STORE_F32, batched=True, max_segment_bytes=12 << 20)
    sb = b.open_stream(DMS)
    other = b.open_stream(DMS[:-1])
    with pytest.raises(vfrb.VfrbError):
        b.restore_stream(other, state)
    with pytest.raises(vfrb.VfrbError):
        b.restore_stream(sb, state[:-4])
    with pytest.raises(vfrb.VfrbError):
        vfrb.call("vfrb_server_save_stream", b.h, sb, vfrb.C.create_string_buffer(16), 16,
                  vfrb.C.byref(vfrb.i64(0)))
    b.restore_stream(sb, state)
    for k in range(k_save, nseg):
        b.submit(sb, p0[cuts[k]:cuts[k + 1]], p1[cuts[k]:cuts[k + 1]], k)
    got = _collect(b, nseg - k_save)
    n_cand = 0
    for k in range(k_save, nseg):
        x, y = got[k], ref[k]
        for f in ("status", "rows", "sample0", "nt", "ncand", "best_dm", "best_t", "best_snr"):
            assert x[f] == y[f], (k, f, x[f], y[f])
        for f in ("peak", "peak_t", "ave", "sigma", "cand_dm", "cand_t", "cand_snr"):
            assert np.array_equal(x[f], y[f]), (k, f)
        n_cand += x["ncand"]
    assert n_cand > 0, ("the pulse after the save point gives candidates", [ref[k]["best_snr"] for k in ref])
    # the state saved now from the restored stream equals the uninterrupted one's
    assert b.save_stream(sb) == a.save_stream(sa)
    a.close()
    b.close()
