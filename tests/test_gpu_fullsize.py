"""The headline shape at full size: 4096 channels (8192-point FFT, NAVE 16,
1.024 ms rows), 2048 DMs 0-1000, windows of 19 531 new rows + the 12 090-row
maximum delay.  The oracle does not finish at this size, so these tests check
properties that hold exactly whatever the size:

- impulse response: unit impulses laid along the delay curves of a few DMs
  in an otherwise zero window.  Every output sample is then (number of
  channels whose delayed sample hits an impulse) / 4096, a small integer over
  a power of two, exact in f32 and f16 whatever the summation order, so the
  whole 2048 x 19 531 output must equal a scatter of the delay table bit for
  bit, for the default group tile, the pair tile and the per-channel tile,
  and through the persistent batch kernel;
- DM 0 is the channel mean of the window row;
- the delay table is the oracle's bit for bit, max_delay 12 090;
- dedisperseSTATS of non-overlapping unit impulses: every (DM, t) sees k
  ones, so ave = k/4096 exactly and sig = sqrt((k - k^2/N) / (N(N-1))) to
  the last bit of the float division and square root;
- the per-DM statistics of a full output match the oracle's arithmetic;
- the filterbank of a whole 20 s segment read as raw VDIF frame pairs (5.15 GB,
  past 2^32 bytes, with dropped pairs the device ingest zero-fills) equals the
  filterbank of the same payload de-framed on the host, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NCH, NDM = 4096, 2048


def _shape(vfrb):
    p = vfrb.params_derive(8192, 16, 16, 1.0, fullband=True)
    dms = np.linspace(0.0, 1000.0, NDM).astype(np.float32)
    return p, dms


@pytest.fixture(scope="module")
def full(gpu, orc):
    vfrb = gpu
    p, dms = _shape(vfrb)
    off, maxd = orc.delay_table(dms, NCH, p.fch1, p.chanbw, p.tchan)
    seg = int(20.0 / p.tchan)
    return p, dms, off, maxd, seg, seg + maxd


def test_delay_table_full_size(gpu, orc, full):
    vfrb = gpu
    p, dms, off, maxd, seg, rows = full
    assert maxd == 12090 and seg == 19531
    plan = vfrb.SearchPlan(NCH, p.tchan, p.fch1, p.chanbw, dms, rows)
    assert plan.max_delay == maxd
    assert np.array_equal(plan.offsets(), off)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_impulse_response_full_size(gpu, full, storage):
    vfrb = gpu
    p, dms, off, maxd, seg, rows = full
    nt = rows - maxd
    # impulses along the curves of DMs d at output samples t0 (edges included)
    pulses = [(0, 0), (NDM - 1, nt - 1), (700, 4321), (1500, 12345), (1501, 12345), (333, 19000)]
    fb = np.zeros((rows, NCH), np.float32)
    cols = np.arange(NCH)
    for d, t0 in pulses:
        fb[t0 + off[d], cols] += 1.0
    # expected: out[d'][t] = #{(pulse, c): t0 + off[d][c] - off[d'][c] == t} / NCH
    cnt = np.zeros(NDM * nt, np.int64)
    dd = np.repeat(np.arange(NDM, dtype=np.int64), NCH)
    for d, t0 in pulses:
        t = (t0 + off[d][None, :] - off).ravel().astype(np.int64)
        ok = (t >= 0) & (t < nt)
        cnt += np.bincount(dd[ok] * nt + t[ok], minlength=NDM * nt)
    exp = (cnt.reshape(NDM, nt) / NCH).astype(np.float32)
    del cnt, dd
    plan = vfrb.SearchPlan(NCH, p.tchan, p.fch1, p.chanbw, dms, rows, vfrb.STORAGE_NAMES[storage])
    d_fb = vfrb.DeviceArray.from_host(fb)
    del fb
    plan.load(d_fb, rows, bandpass=False)
    out = vfrb.DeviceArray((NDM, nt), np.float32)
    names = vfrb.tile_names()
    default = plan.tile
    for tile in (default, 16, 20, 12, 8, 3):
        print(f"[case] impulse response {storage} tile {names[tile]}", flush=True)
        plan.set_tile(tile)
        out.zero()
        plan.dedisperse(rows, out, nt)
        got = out.download()
        bad = np.argwhere(got != exp)
        assert bad.size == 0, (names[tile], storage, bad[:5], got[tuple(bad[0])], exp[tuple(bad[0])])
    print(f"[case] impulse response {storage} batched", flush=True)
    b = vfrb.Batch(vfrb.STORAGE_NAMES[storage], max_jobs=1)
    out.zero()
    b.set([(plan, rows, out, nt)])
    b.dedisperse()
    assert np.array_equal(out.download(), exp), "batched"
    for d, t0 in pulses:
        assert exp[d, t0] >= 1.0


def test_dm0_is_the_channel_mean_full_size(gpu, full):
    vfrb = gpu
    p, dms, off, maxd, seg, rows = full
    nt = rows - maxd
    rng = np.random.default_rng(77)
    fb = rng.standard_normal((rows, NCH), dtype=np.float32) * 9.0 + 40.0
    plan = vfrb.SearchPlan(NCH, p.tchan, p.fch1, p.chanbw, dms, rows)
    d_fb = vfrb.DeviceArray.from_host(fb)
    plan.load(d_fb, rows, bandpass=False)
    out = vfrb.DeviceArray((NDM, nt), np.float32)
    plan.dedisperse(rows, out, nt)
    row0 = out.download()[0]
    ref = fb[:nt].astype(np.float64).mean(axis=1)
    # f32 sums of 4096 terms of size ~40: within a few ulps of the mean
    assert np.max(np.abs(row0 - ref)) < 2e-5 * 40.0


def test_dm_stats_full_size(gpu, orc, full):
    vfrb = gpu
    p, dms, off, maxd, seg, rows = full
    rng = np.random.default_rng(78)
    x = rng.normal(0.0, 0.02, (NDM, seg)).astype(np.float32)
    x[1234, 9876] += 0.3
    d = vfrb.DeviceArray.from_host(x)
    st = vfrb.dm_stats(d, seg, NDM, seg, vfrb.ACC_F64)
    frb, ave, sig, peak, pt = orc.dm_stats(x)
    # FP64 sums of the same floats: equal but for a double sum on a float
    # rounding boundary
    assert np.max(np.abs(st.ave - ave) / np.maximum(np.abs(ave), 1e-6)) < 1e-5
    assert np.max(np.abs(st.sigma - sig) / sig) < 2e-7
    assert np.array_equal(st.frb, frb) and np.array_equal(st.peak_t, pt)
    assert st.peak_t[1234] == 9876 and np.allclose(st.peak, peak, rtol=1e-6)


def test_filterbank_from_frames_full_segment(gpu, full):
    vfrb = gpu
    p, dms, off, maxd, seg, rows = full
    n, nave = 8192, 16
    need = seg * nave * n                   # payload bytes per pol for the segment
    npairs = -(-need // 5000)
    drop = np.zeros(npairs, bool)
    drop[[1000, 1001, 1002, 333333, npairs - 7]] = True
    rng = np.random.default_rng(79)
    pay = rng.integers(0, 256, (2, npairs, 5000), dtype=np.uint8)
    pay[:, drop] = 0                        # what the ingest inserts for a missing pair
    kept = np.flatnonzero(~drop)
    frames = np.empty((kept.size, 2, 5032), np.uint8)
    words = np.zeros((kept.size, 8), np.uint32)
    words[:, 0] = 1000 + kept // 25600
    words[:, 1] = (kept % 25600) | (30 << 24)
    words[:, 2] = 629
    for th in (0, 1):
        words[:, 3] = 0x5856 | (7 << 26) | (th << 16)
        frames[:, th, :32] = words.view(np.uint8).reshape(kept.size, 32)
        frames[:, th, 32:] = pay[th, kept]
    print(f"[case] framed filterbank, {kept.size} pairs ({frames.nbytes / 1e9:.2f} GB), {seg} rows", flush=True)
    raw = vfrb.DeviceArray.from_host(frames.reshape(-1))
    del frames
    ing = vfrb.Ingest(kept.size)
    table = vfrb.DeviceArray(npairs + 64, np.int32)
    nslots = ing.frames(raw, kept.size, table)
    assert nslots == npairs
    fbp = vfrb.FilterbankPlan(n, nave, p.nskip, p.nchanout)
    a = vfrb.DeviceArray((seg, NCH), np.float32)
    fbp.execute_frames(raw, table, nslots, 0, seg, a)
    got = a.download()
    del raw, a
    d0 = vfrb.DeviceArray.from_host(pay[0].reshape(-1))
    d1 = vfrb.DeviceArray.from_host(pay[1].reshape(-1))
    del pay
    b = vfrb.DeviceArray((seg, NCH), np.float32)
    fbp.execute(d0, d1, seg, b)
    assert np.array_equal(got, b.download())
    assert np.all(np.isfinite(got)) and got.min() >= 0.0


def test_dedisperse_stats_impulses_full_size(gpu, full):
    vfrb = gpu
    p, dms, off, maxd, seg, rows = full
    nt = rows - maxd
    pulses = [(0, 0), (NDM - 1, nt - 1), (700, 4321), (1500, 12345), (333, 17500), (1200, 800)]
    fb = np.zeros((rows, NCH), np.float32)
    cols = np.arange(NCH)
    for d, t0 in pulses:
        fb[t0 + off[d], cols] += 1.0
    assert fb.max() == 1.0  # no two impulses share a sample (checked on the delay table)
    cnt = np.zeros(NDM * nt, np.int64)
    dd = np.repeat(np.arange(NDM, dtype=np.int64), NCH)
    for d, t0 in pulses:
        t = (t0 + off[d][None, :] - off).ravel().astype(np.int64)
        ok = (t >= 0) & (t < nt)
        cnt += np.bincount(dd[ok] * nt + t[ok], minlength=NDM * nt)
    k = cnt.reshape(NDM, nt).astype(np.float64)
    del cnt, dd
    exp_ave = (k / NCH).astype(np.float32)
    num = (k - k * k / NCH).astype(np.float32)
    exp_sig = np.sqrt(num / np.float32(NCH * (NCH - 1)))
    print("[case] dedisperseSTATS impulses, full size", flush=True)
    plan = vfrb.SearchPlan(NCH, p.tchan, p.fch1, p.chanbw, dms, rows)
    d_fb = vfrb.DeviceArray.from_host(fb)
    del fb
    plan.load(d_fb, rows, bandpass=False)
    out, sig = vfrb.DeviceArray((NDM, nt), np.float32), vfrb.DeviceArray((NDM, nt), np.float32)
    plan.dedisperse_stats(rows, out, sig, nt)
    assert np.array_equal(out.download(), exp_ave)
    g = sig.download()
    assert np.all(np.abs(g - exp_sig) <= np.spacing(np.maximum(exp_sig, np.float32(1e-30))))
    assert np.count_nonzero(exp_sig) > 0
