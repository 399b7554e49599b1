"""GPU: device-side VDIF frame ingest (vfrb_ingest_*) against the host frame
sequencer, whose byte stream is pinned to the restated reference sequencer
(PIPELINE:888-1171, TIMESERIES:893-937) in the host tests; and the
filterbank read through the slot table against the filterbank of the
de-framed bytes."""
import numpy as np
import pytest

import vdif_oracle as vo

pytestmark = pytest.mark.gpu
FPS = 25600


def _stream(rng, npairs, drop=(), start_frame=0, swap_threads_at=None, fn_mismatch_at=None):
    p0 = rng.integers(0, 256, (npairs, 5000), dtype=np.uint8)
    p1 = rng.integers(0, 256, (npairs, 5000), dtype=np.uint8)
    frames = []
    kept = 0
    for k in range(npairs):
        if k in drop:
            continue
        a = start_frame + k
        fn, sec = a % FPS, 1000 + a // FPS
        th = (1, 0) if k == swap_threads_at else (0, 1)
        fn1 = (fn + 1) % FPS if k == fn_mismatch_at else fn
        frames.append((sec, fn, th[0], p0[k].tobytes()))
        frames.append((sec, fn1, th[1], p1[k].tobytes()))
        kept += 1
    return vo.build_stream(frames), kept


def _host_streams(vfrb, data, seconds, nbytes):
    seq = vfrb.Sequencer(data=data, seconds_check=seconds)
    assert seq.init() == 0
    st, a, b = seq.fill(nbytes)
    assert st == 0
    return a, b


def _device_streams(vfrb, data, npairs, seconds, batches=1):
    raw = vfrb.DeviceArray.from_host(np.frombuffer(data, np.uint8))
    ing = vfrb.Ingest(npairs, seconds_check=seconds)
    out0, out1 = [], []
    edges = np.linspace(0, npairs, batches + 1).astype(int)
    for a, b in zip(edges[:-1], edges[1:]):
        table = vfrb.DeviceArray(4 * (b - a) * 2 + 64 * FPS, np.int32)
        nslots = ing.frames(raw, b - a, table, offset=int(a) * 2 * 5032)
        for pol, acc in ((0, out0), (1, out1)):
            o = vfrb.DeviceArray(nslots * 5000, np.uint8)
            vfrb.Ingest.gather(raw, table, nslots, pol, o, offset=int(a) * 2 * 5032)
            acc.append(o.download())
    return np.concatenate(out0), np.concatenate(out1)


@pytest.mark.parametrize("case", ["clean", "drops", "second_wrap", "seconds_rule", "batches"])
def test_ingest_matches_host_sequencer(gpu, case):
    vfrb = gpu
    rng = np.random.default_rng(["clean", "drops", "second_wrap", "seconds_rule", "batches"].index(case))
    seconds = case == "seconds_rule"
    if case == "clean":
        data, kept = _stream(rng, 300)
    elif case in ("drops", "batches"):
        data, kept = _stream(rng, 400, drop={5, 6, 7, 50, 121, 122, 300})
    else:  # wrap of the frame counter and the second, with a gap across it
        data, kept = _stream(rng, 2700, drop={2558, 2559, 2560, 2561}, start_frame=FPS - 2560)
    print(f"[case] ingest {case}: {kept} pairs", flush=True)
    d0, d1 = _device_streams(vfrb, data, kept, seconds, batches=3 if case == "batches" else 1)
    nbytes = d0.size - 5000 - 17   # the host fill reads one pair ahead of what it copies
    h0, h1 = _host_streams(vfrb, data, seconds, nbytes)
    assert np.array_equal(d0[:nbytes], h0) and np.array_equal(d1[:nbytes], h1)
    zero_slots = int(np.sum(np.all(d0.reshape(-1, 5000) == 0, axis=1)))
    assert zero_slots >= (0 if case == "clean" else 4)


def test_ingest_rejects_malformed_pairs(gpu):
    vfrb = gpu
    rng = np.random.default_rng(3)
    for kw in ({"swap_threads_at": 10}, {"fn_mismatch_at": 20}):
        data, kept = _stream(rng, 40, **kw)
        raw = vfrb.DeviceArray.from_host(np.frombuffer(data, np.uint8))
        ing = vfrb.Ingest(kept)
        table = vfrb.DeviceArray(4096, np.int32)
        with pytest.raises(vfrb.VfrbError) as e:
            ing.frames(raw, kept, table)
        assert e.value.status == vfrb.ERR_FORMAT


def test_filterbank_from_frames_equals_deframed(gpu):
    vfrb = gpu
    rng = np.random.default_rng(9)
    data, kept = _stream(rng, 500, drop={33, 34, 200})
    raw = vfrb.DeviceArray.from_host(np.frombuffer(data, np.uint8))
    ing = vfrb.Ingest(kept)
    table = vfrb.DeviceArray(4096, np.int32)
    nslots = ing.frames(raw, kept, table)
    n, nave = 2048, 4
    p = vfrb.params_derive(n, nave, 1, 1.0, fullband=True)
    fbp = vfrb.FilterbankPlan(n, nave, p.nskip, p.nchanout)
    rowb = n * nave
    pols = []
    for pol in (0, 1):
        o = vfrb.DeviceArray(nslots * 5000, np.uint8)
        vfrb.Ingest.gather(raw, table, nslots, pol, o)
        pols.append(o)
    for first_row in (0, 3):
        byte0 = first_row * rowb
        nrows = (nslots * 5000 - byte0) // rowb
        print(f"[case] filterbank from frames, first row {first_row}, {nrows} rows", flush=True)
        a = vfrb.DeviceArray((nrows, p.nchanout), np.float32)
        b = vfrb.DeviceArray((nrows, p.nchanout), np.float32)
        fbp.execute_frames(raw, table, nslots, byte0, nrows, a)
        fbp.execute(pols[0], pols[1], nrows, b, pol_offset=byte0)
        assert np.array_equal(a.download(), b.download())
    with pytest.raises(vfrb.VfrbError):   # rows past the table
        fbp.execute_frames(raw, table, nslots, 0, nslots * 5000 // rowb + 1, a)


@pytest.mark.parametrize("n", [2048, 8192, 16384])
def test_filterbank_from_frames_slots_far_apart(gpu, n):
    """Consecutive payload slots read from frame pairs more than 4 GB apart
    in one buffer (and missing slots): the filterbank reads each slot where
    the table puts it, equal to the filterbank of the gathered bytes."""
    vfrb = gpu
    rng = np.random.default_rng(n)
    npairs, far = 460000, 459900          # 4.63 GB of frame pairs; pair `far` lies past 4 GiB
    nslots = 96
    raw = vfrb.DeviceArray(npairs * 2 * 5032, np.uint8)
    for first in (0, far):
        blk = rng.integers(0, 256, 64 * 2 * 5032, dtype=np.uint8)
        vfrb.call("vfrb_memcpy_h2d", raw.offset(first * 2 * 5032), blk.ctypes.data_as(vfrb.C.c_void_p),
                  blk.nbytes, None)
    t = np.where(np.arange(nslots) % 2 == 0, np.arange(nslots) // 2, far + np.arange(nslots) // 2)
    t[[5, 6, 40]] = -1
    table = vfrb.DeviceArray.from_host(t.astype(np.int32))
    nave = 2
    fbp = vfrb.FilterbankPlan(n, nave, 1, n // 2)
    pols = []
    for pol in (0, 1):
        o = vfrb.DeviceArray(nslots * 5000, np.uint8)
        vfrb.Ingest.gather(raw, table, nslots, pol, o)
        pols.append(o)
    rowb = n * nave
    for byte0 in (0, 16 * 195):
        nrows = (nslots * 5000 - byte0) // rowb
        print(f"[case] {n}-point filterbank, slots far apart, byte0 {byte0}, {nrows} rows", flush=True)
        a = vfrb.DeviceArray((nrows, n // 2), np.float32)
        b = vfrb.DeviceArray((nrows, n // 2), np.float32)
        fbp.execute_frames(raw, table, nslots, byte0, nrows, a)
        fbp.execute(pols[0], pols[1], nrows, b, pol_offset=byte0)
        assert np.array_equal(a.download(), b.download())


def test_ingest_failed_call_keeps_the_parse_state(gpu):
    """A batch that fails its capacity check leaves the carried parse state
    (expected frame) untouched: retrying it with a larger table, then the
    next batch, gives the same payload as two good calls."""
    vfrb = gpu
    rng = np.random.default_rng(77)
    npairs = 600
    data, nkept = _stream(rng, npairs, drop=(5, 6, 7, 301, 302, 450))
    raw = vfrb.DeviceArray.from_host(np.frombuffer(data, np.uint8))
    print("[case] ingest capacity failure then retry", flush=True)
    half = 300  # frame pairs in the first batch
    ing = vfrb.Ingest(nkept)
    small = vfrb.DeviceArray(64, np.int32)
    with pytest.raises(vfrb.VfrbError) as e:
        ing.frames(raw, half, small)
    assert e.value.status == vfrb.ERR_CAPACITY
    outs = []
    off = 0
    for n in (half, nkept - half):
        table = vfrb.DeviceArray(4 * n + 64, np.int32)
        nslots = ing.frames(raw, n, table, offset=off)
        o = vfrb.DeviceArray(nslots * 5000, np.uint8)
        vfrb.Ingest.gather(raw, table, nslots, 0, o, offset=off)
        outs.append(o.download())
        off += n * 2 * 5032
    got = np.concatenate(outs)
    want, _ = _device_streams(vfrb, data, nkept, False)
    assert got.shape == want.shape and np.array_equal(got, want)
    nbytes = want.size - 5000 - 17  # the host fill reads one pair ahead of what it copies
    host0, _ = _host_streams(vfrb, data, False, nbytes)
    assert np.array_equal(got[:nbytes], host0)
