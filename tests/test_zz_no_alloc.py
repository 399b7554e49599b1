"""GPU: the no-allocation ingest entry point (vfrb_ingest_frames_noalloc),
called through the C ABI, against vfrb_ingest_frames on the same frames:
two calls in a row (the parse state carried by the caller), the second with
a smaller input, then its limits and its error path."""
import numpy as np
import pytest

from test_gpu_ingest import _stream

pytestmark = pytest.mark.gpu


def test_noalloc_ingest_equals_ingest_frames(gpu):
    vfrb = gpu
    rng = np.random.default_rng(88)
    data, kept = _stream(rng, 1500, drop={4, 5, 6, 640, 641, 1200})
    raw = vfrb.DeviceArray.from_host(np.frombuffer(data, np.uint8))
    ing = vfrb.Ingest(kept)
    state = [-1, 0]
    off = 0
    for n in (900, kept - 900):  # the second call is smaller
        print(f"[case] no-alloc ingest: {n} pairs at byte {off}", flush=True)
        t_ref = vfrb.DeviceArray(2 * n + 64, np.int32)
        t_new = vfrb.DeviceArray(2 * n + 64, np.int32)
        t_ref.upload(np.full(2 * n + 64, 7, np.int32))
        t_new.upload(np.full(2 * n + 64, 7, np.int32))
        ns_ref = ing.frames(raw, n, t_ref, offset=off)
        ns_new = vfrb.ingest_frames_noalloc(raw, n, state, t_new, offset=off)
        assert ns_new == ns_ref
        assert np.array_equal(t_new.download(), t_ref.download())
        a = vfrb.DeviceArray(ns_ref * 5000, np.uint8)
        vfrb.Ingest.gather(raw, t_new, ns_new, 1, a, offset=off)
        b = vfrb.DeviceArray(ns_ref * 5000, np.uint8)
        vfrb.Ingest.gather(raw, t_ref, ns_ref, 1, b, offset=off)
        assert np.array_equal(a.download(), b.download())
        off += n * 2 * 5032
    # the state after the last pair (frame 1499 of second 1000, the stream starting at frame 0)
    assert state == [1500, 1000]


def test_noalloc_ingest_limits_and_errors(gpu):
    vfrb = gpu
    rng = np.random.default_rng(89)
    data, kept = _stream(rng, 200, drop={50, 51})
    raw = vfrb.DeviceArray.from_host(np.frombuffer(data, np.uint8))
    table = vfrb.DeviceArray(4096, np.int32)
    state = [-1, 0]
    print("[case] no-alloc ingest: rejected sizes", flush=True)
    with pytest.raises(vfrb.VfrbError) as e:   # past the scratch: rejected before any launch
        vfrb.ingest_frames_noalloc(raw, (1 << 20) + 1, state, table)
    assert e.value.status == vfrb.ERR_CAPACITY and "no-allocation" in str(e.value)
    small = vfrb.DeviceArray(16, np.int32)
    with pytest.raises(vfrb.VfrbError) as e:   # slots past the caller's table
        vfrb.ingest_frames_noalloc(raw, kept, state, small)
    assert e.value.status == vfrb.ERR_CAPACITY and state == [-1, 0]
    print("[case] no-alloc ingest: retry after the capacity error", flush=True)
    n = vfrb.ingest_frames_noalloc(raw, kept, state, table)
    assert n == 200 and state == [200, 1000]
    bad, kbad = _stream(rng, 30, swap_threads_at=7)
    rbad = vfrb.DeviceArray.from_host(np.frombuffer(bad, np.uint8))
    st2 = [-1, 0]
    with pytest.raises(vfrb.VfrbError) as e:
        vfrb.ingest_frames_noalloc(rbad, kbad, st2, table)
    assert e.value.status == vfrb.ERR_FORMAT and st2 == [-1, 0]
