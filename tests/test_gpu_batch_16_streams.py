"""GPU: the batching workload at its own size - 16 antenna streams of uneven
length in one persistent launch: 1024 channels of 1.024 ms rows, segments
U(2, 20) s, each stream with its own DM list (maximum DM U(200, 1000),
512-2048 trials), in f32 and f16 storage, eagerly and through the captured
hipGraph replayed on the work list.

Unit impulses are laid along the delay curves of a few DMs of every stream's
otherwise zero window.  Each output sample is then (channels whose delayed
sample hits an impulse) / 1024, a small integer over a power of two, exact in
f32 and f16 whatever the summation order or tile: every stream's whole
output must equal the scatter of its delay table bit for bit, in the batch,
in every graph replay, and in per-plan launches."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NCH = 1024
TCHAN = 64 * 2048 / 128e6
CHANBW = 64.0 / NCH
FCH1 = 320.0 + NCH * CHANBW - 0.5 * CHANBW
NSTREAMS = 16


def _streams(orc, seed=2024):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(NSTREAMS):
        seg_s = rng.uniform(2.0, 20.0)
        dm_max = rng.uniform(200.0, 1000.0)
        ndm = int(rng.integers(512, 2049))
        dms = np.linspace(0.0, dm_max, ndm).astype(np.float32)
        off, maxd = orc.delay_table(dms, NCH, FCH1, CHANBW, TCHAN)
        nt = int(seg_s / TCHAN)
        pulses = [(0, 0), (ndm - 1, nt - 1)] + [(int(rng.integers(0, ndm)), int(rng.integers(0, nt)))
                                                 for _ in range(3)]
        out.append(dict(dms=dms, off=off, maxd=maxd, nt=nt, rows=nt + maxd, pulses=pulses))
    return out


def _window(s):
    fb = np.zeros((s["rows"], NCH), np.float32)
    cols = np.arange(NCH)
    for d, t0 in s["pulses"]:
        fb[t0 + s["off"][d], cols] += 1.0
    return fb


def _expected(s):
    if "exp" in s:
        return s["exp"]
    off, nt = s["off"], s["nt"]
    ndm = off.shape[0]
    cnt = np.zeros(ndm * nt, np.int64)
    dd = np.repeat(np.arange(ndm, dtype=np.int64), NCH)
    for d, t0 in s["pulses"]:
        t = (t0 + off[d][None, :] - off).ravel().astype(np.int64)
        ok = (t >= 0) & (t < nt)
        cnt += np.bincount(dd[ok] * nt + t[ok], minlength=ndm * nt)
    s["exp"] = (cnt.reshape(ndm, nt) / NCH).astype(np.float32)
    return s["exp"]


def _check(outs, streams, what):
    for i, (o, s) in enumerate(zip(outs, streams)):
        got = o.download()
        exp = _expected(s)
        bad = np.argwhere(got != exp)
        assert bad.size == 0, (what, i, bad[:5])


@pytest.mark.timeout(600)
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_batch_16_uneven_streams_impulse_response(gpu, orc, storage):
    vfrb = gpu
    streams = _streams(orc)
    plans, outs, keep = [], [], []
    for s in streams:
        plan = vfrb.SearchPlan(NCH, TCHAN, FCH1, CHANBW, s["dms"], s["rows"], vfrb.STORAGE_NAMES[storage])
        assert plan.max_delay == s["maxd"]
        d = vfrb.DeviceArray.from_host(_window(s))
        plan.load(d, s["rows"], bandpass=False)
        del d
        plans.append(plan)
        outs.append(vfrb.DeviceArray((len(s["dms"]), s["nt"]), np.float32))
    names = vfrb.tile_names()
    total_s = sum(s["nt"] for s in streams) * TCHAN
    print(f"[case] 16-stream batch {storage}: {total_s:.1f} data-s, DMs "
          f"{min(len(s['dms']) for s in streams)}-{max(len(s['dms']) for s in streams)}, per-plan tiles "
          f"{sorted(set(names[p.tile] for p in plans))}", flush=True)
    # per-plan launches, each on its own plan's tile
    for p, o, s in zip(plans, outs, streams):
        o.zero()
        p.dedisperse(s["rows"], o, s["nt"])
    _check(outs, streams, "per-plan")
    # one persistent launch over all 16 windows
    batch = vfrb.Batch(vfrb.STORAGE_NAMES[storage], max_jobs=NSTREAMS, max_tiles=1 << 20)
    stream = vfrb.Stream()
    jobs = [(p, s["rows"], o, s["nt"]) for p, o, s in zip(plans, outs, streams)]
    ntiles = batch.set(jobs, stream=stream)
    print(f"[case] 16-stream batch {storage}: tile {names[batch.tile]}, {ntiles} work items, "
          f"grid {batch.grid}", flush=True)
    for o in outs:
        o.zero()
    batch.dedisperse(stream=stream)
    stream.sync()
    _check(outs, streams, "batch")
    # graph replays, with eager work in between
    for k in range(2):
        for o in outs:
            o.zero()
        batch.graph_launch(stream)
        stream.sync()
        _check(outs, streams, f"graph replay {k}")
        plans[k].dedisperse(streams[k]["rows"], outs[k], streams[k]["nt"])
    # a new job set (the first 8 streams, 1 s shorter) through the same graph
    jobs2 = [(p, r - int(1.0 / TCHAN), o, nt - int(1.0 / TCHAN)) for p, r, o, nt in jobs[:8]]
    batch.set(jobs2, stream=stream)
    for o in outs:
        o.zero()
    batch.graph_launch(stream)
    stream.sync()
    for (p, r, o, ld), s in zip(jobs2, streams[:8]):
        got = o.download()
        nt2 = r - p.max_delay
        exp = _expected(s)[:, :nt2]
        # rows past the shortened window were never written
        got2 = got.reshape(-1)[: exp.shape[0] * ld].reshape(exp.shape[0], ld)[:, :nt2]
        assert np.array_equal(got2, exp), "graph replay on the shorter set"
