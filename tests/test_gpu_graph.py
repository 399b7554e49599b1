"""hipGraph capture of the search chain (vfrb_graph_*): the 1024-channel,
256-DM chain of one 1 s segment (filterbank into the overlap ring, bandpass +
pack, dedispersion, per-DM statistics, candidates) captured once, after
two segments have filled the ring, and replayed; a replay rewrites the same
ring rows with the same payload, so each one repeats the eager segment.  Every replay, with eager work and a synchronisation between
replays and its outputs cleared beforehand, equals the eager chain bit for
bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", ["f32", "f16", "u8"])
def test_captured_chain_replays_equal_eager(gpu, storage):
    vfrb = gpu
    n, nave = 2048, 64
    p = vfrb.params_derive(n, nave, 16, 1.0, fullband=True)
    nch, tchan = p.nchanout, p.tchan
    rows = int(1.0 / tchan)
    dms = (np.arange(256) * 0.1).astype(np.float32)
    _, max_delay, _ = vfrb.delay_table(dms, nch, p.fch1, p.chanbw, tchan)
    win = rows + max_delay
    spec = vfrb.synth_spec(n, seed=4, pulses=[(12.3, 0.35, 0.002, vfrb.amp_for_snr(20.0, nave, nch))])
    d0 = vfrb.DeviceArray.from_host(vfrb.synth_payload(spec, 0, 0, rows * nave))
    d1 = vfrb.DeviceArray.from_host(vfrb.synth_payload(spec, 1, 0, rows * nave))
    ring = vfrb.DeviceArray((win, nch), np.float32)
    ring.zero()
    fbp = vfrb.FilterbankPlan(n, nave, p.nskip, nch)
    sp = vfrb.SearchPlan(nch, tchan, p.fch1, p.chanbw, dms, win, vfrb.STORAGE_NAMES[storage])
    ndm = len(dms)
    out = vfrb.DeviceArray((ndm, rows), np.float32)
    maxc = 1 << 14
    cdm, ct = vfrb.DeviceArray(maxc, np.int32), vfrb.DeviceArray(maxc, np.int32)
    csnr, cnt = vfrb.DeviceArray(maxc, np.float32), vfrb.DeviceArray(1, np.int32)
    ave, sig, peak = (vfrb.DeviceArray(ndm, np.float32) for _ in range(3))
    frb, pt = vfrb.DeviceArray(ndm, np.int32), vfrb.DeviceArray(ndm, np.int32)
    acc = vfrb.ACC_F64 if storage == "f32" else vfrb.ACC_F32C
    s = vfrb.Stream()

    def chain(i):
        # segment i's rows follow the previous segments' overlap in the ring
        fbp.execute(d0, d1, rows, ring, ring_rows=win, row0=(max_delay + i * rows) % win, stream=s.h)
        sp.load(ring, win, ring_rows=win, row0=(i * rows) % win, bandpass=True, stream=s.h)
        sp.dedisperse(win, out, rows, stream=s.h)
        vfrb.call("vfrb_dm_stats", out.ptr, rows, ndm, rows, acc, frb.ptr, ave.ptr, sig.ptr, peak.ptr,
                  pt.ptr, s.h)
        vfrb.call("vfrb_candidates", out.ptr, rows, ndm, rows, ave.ptr, sig.ptr, 7.0, cdm.ptr, ct.ptr,
                  csnr.ptr, maxc, cnt.ptr, s.h)

    def results():
        s.sync()
        k = int(cnt.download()[0])
        order = np.lexsort((ct.download()[:k], cdm.download()[:k]))
        return (out.download(), ave.download(), sig.download(), peak.download(), pt.download(),
                cdm.download()[:k][order], ct.download()[:k][order])

    def clear():
        for a in (out, ave, sig, peak, pt, cnt, cdm, ct):
            a.zero(stream=s.h)
        s.sync()

    for i in range(3):  # the ring holds whole segments from the second on
        chain(i)
    eager = results()
    print(f"[case] eager chain: {eager[6].size} candidates above 7 sigma, peak SNR {eager[3].max():.1f}", flush=True)
    assert eager[6].size > 0, "the pulse gives candidates"
    g = vfrb.Graph(s)
    g.begin()
    chain(2)
    g.end()
    other = vfrb.DeviceArray((ndm, rows), np.float32)
    for rep in range(3):
        print(f"[case] graph replay {rep} ({storage})", flush=True)
        clear()
        g.launch()
        got = results()
        for a, b in zip(got, eager):
            assert np.array_equal(a, b)
        # eager work of the program between replays
        sp.dedisperse(win, other, rows, stream=s.h)
        s.sync()
        assert np.array_equal(other.download(), eager[0])
