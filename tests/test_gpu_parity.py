"""GPU parity: every HIP kernel against the CPU oracle on the same seeded
inputs, through the C ABI (libvfrb.so)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noise_bytes(rng, n):
    return np.clip(np.rint(rng.normal(128, 8, n)), 0, 255).astype(np.uint8)


def _rel_err(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b)) / np.mean(np.abs(b)))


# ---------------------------------------------------------- front end ----
@pytest.mark.parametrize("n,nave,nrows,fullband", [
    (2048, 64, 6, True),     # 1024-channel shape
    (8192, 16, 5, True),     # 4096-channel shape
    (4096, 8, 7, False),     # MUOS cut
    (1024, 5, 9, True),      # nave not a multiple of the FFTs in flight
    (256, 3, 4, True),
    (512, 4, 3, True),       # one radix-4 middle pass
    (16384, 2, 3, True),
    (12500, 2, 3, False),    # reference default length: fixed-length kernel, radices 5^5 * 4
    (6250, 3, 4, True),      # fixed-length kernels of its halvings: 5^5 * 2
    (2500, 4, 3, True),      # 5^4 * 4
    (1250, 2, 5, True),      # 5^4 * 2
    (240, 7, 5, True),       # run-time radices 8*2*3*5
    (3000, 2, 3, True),      # run-time radices 8*5*5*3
])
def test_filterbank_vs_oracle(gpu, orc, n, nave, nrows, fullband):
    vfrb = gpu
    p = vfrb.params_derive(n, nave, 1, 1.0, fullband=fullband)
    print(f"[case] filterbank n={n} nave={nave} rows={nrows} nskip={p.nskip} nch={p.nchanout}", flush=True)
    rng = np.random.default_rng(n + nave)
    tot = nrows * nave * n
    p0, p1 = _noise_bytes(rng, tot), _noise_bytes(rng, tot)
    d0, d1 = vfrb.DeviceArray.from_host(p0), vfrb.DeviceArray.from_host(p1)
    fb = vfrb.DeviceArray((nrows, p.nchanout), np.float32)
    plan = vfrb.FilterbankPlan(n, nave, p.nskip, p.nchanout)
    plan.execute(d0, d1, nrows, fb)
    got = fb.download()
    ref = orc.filterbank_f64(p0, p1, nrows, n, nave, p.nskip, p.nchanout)
    # FP32 transform vs FP64: relative to the mean power of the row
    assert _rel_err(got, ref) < 2e-4, _rel_err(got, ref)


def test_filterbank_ring_wrap(gpu, orc):
    vfrb = gpu
    n, nave, nrows, ring = 512, 4, 6, 8
    rng = np.random.default_rng(9)
    p0, p1 = _noise_bytes(rng, nrows * nave * n), _noise_bytes(rng, nrows * nave * n)
    d0, d1 = vfrb.DeviceArray.from_host(p0), vfrb.DeviceArray.from_host(p1)
    fb = vfrb.DeviceArray((ring, n // 2), np.float32)
    fb.zero()
    plan = vfrb.FilterbankPlan(n, nave, 1, n // 2)
    plan.execute(d0, d1, nrows, fb, ring_rows=ring, row0=5)
    got = fb.download()
    ref = orc.filterbank_f64(p0, p1, nrows, n, nave, 1, n // 2)
    order = [(5 + r) % ring for r in range(nrows)]
    assert _rel_err(got[order], ref) < 2e-4
    untouched = sorted(set(range(ring)) - set(order))
    assert np.all(got[untouched] == 0)


def test_filterbank_tone_closed_form(gpu):
    """Period-4 tone on pol 0: all power in bin N/4, |X| = a*N/2."""
    vfrb = gpu
    n, nave, nrows, a = 8192, 16, 2, 40
    pat = np.tile(np.array([128 + a, 128, 128 - a, 128], np.uint8), nrows * nave * n // 4)
    flat = np.full(nrows * nave * n, 128, np.uint8)
    d0, d1 = vfrb.DeviceArray.from_host(pat), vfrb.DeviceArray.from_host(flat)
    fb = vfrb.DeviceArray((nrows, n // 2), np.float32)
    vfrb.FilterbankPlan(n, nave, 1, n // 2).execute(d0, d1, nrows, fb)
    got = fb.download()
    expect = (a * n / 2.0) ** 2 / (n // 2 + 1)
    assert np.allclose(got[:, n // 4 - 1], expect, rtol=1e-5)
    others = np.delete(got, n // 4 - 1, axis=1)
    assert np.max(np.abs(others)) < 1e-6 * expect


# ----------------------------------------------------------- bandpass ----
def test_bandpass_inplace_vs_oracle(gpu, orc):
    vfrb = gpu
    rng = np.random.default_rng(1)
    rows, nch = 976, 1024
    fb = rng.normal(256, 22, (rows, nch)).astype(np.float32)
    fb[:, 77] += 5000.0                # RFI channel
    fb[300:310, 5] += 800.0            # burst
    fb[:, 9] = 3.0                     # constant channel: left untouched
    d = vfrb.DeviceArray.from_host(fb)
    vfrb.bandpass_inplace(d, rows, nch)
    got = d.download()
    ref, margin = orc.bandpass_f32(fb, with_margin=True)
    ok = np.repeat(margin > 1e-4, 128, axis=0)[:rows]    # blocks with no value at a clip edge
    assert ok.mean() > 0.99
    assert np.max(np.abs(got - ref)[ok]) < 2e-3
    assert np.all(got[:, 9] == 3.0)


def test_chan_stats_vs_oracle(gpu, orc):
    vfrb = gpu
    rng = np.random.default_rng(2)
    rows, nch, ncol, nblk = 2000, 300, 400, 5
    fb = rng.normal(0, 10, (rows, nch)).astype(np.float32)
    fb[:, 3] = 1.0
    d = vfrb.DeviceArray.from_host(fb)
    ave, sig, out = vfrb.chan_stats(d, rows, nch, ncol, nblk)
    rave, rsig, rout = orc.chan_stats(fb, ncol, nblk)
    assert np.allclose(ave, rave, atol=1e-4) and np.allclose(sig, rsig, rtol=1e-4, atol=1e-4)
    assert np.array_equal(out, rout)


@pytest.mark.parametrize("rows,nch,ncol,nblk", [
    (2000, 300, 400, 5),     # stat blocks straddle bandpass blocks
    (3000, 130, 977, 3),     # rows past the last stat block, ragged last bandpass block
    (1000, 64, 128, 8),      # aligned blocks, short last one
    (40, 70, 1, 40),         # one-row stat blocks: 0/0 variance, as the reference
    (70000, 64, 33000, 2),   # long stat blocks (260 bandpass blocks each)
])
def test_bandpass_chan_stats_fused(gpu, orc, rows, nch, ncol, nblk):
    """The one-pass STATS window (bandpass + channel statistics) equals the
    two launches bit for bit, leaves the filterbank unchanged, and agrees
    with the oracle's subtractbpCAL + calcstats."""
    vfrb = gpu
    rng = np.random.default_rng(rows + nch)
    fb = rng.normal(100, 10, (rows, nch)).astype(np.float32)
    fb[:, 3] = 1.0
    fb[rows // 3, 5] += 500.0
    fb[rows // 2:rows // 2 + 9, 7] += 90.0
    d = vfrb.DeviceArray.from_host(fb)
    fused = vfrb.bandpass_chan_stats(d, rows, nch, ncol, nblk)
    assert np.array_equal(d.download(), fb)
    vfrb.bandpass_inplace(d, rows, nch)
    two = vfrb.chan_stats(d, rows, nch, ncol, nblk)
    for a, b in zip(fused, two):
        assert np.array_equal(a, b, equal_nan=True)
    if rows <= 3000 or ncol > 30000:
        rave, rsig, rout = orc.chan_stats(orc.bandpass_f32(fb), ncol, nblk)
        ok = np.isfinite(rsig)
        assert np.array_equal(ok, np.isfinite(fused[1]))
        assert np.allclose(fused[0], rave, atol=1e-4)
        assert np.allclose(fused[1][ok], rsig[ok], rtol=1e-4, atol=1e-4)
        assert np.array_equal(fused[2][ok], rout[ok])


def test_chan_stats_all_above_hi(gpu, orc):
    """The outlier fraction is count/npts in integer division: 1 exactly when
    every sample exceeds HI = ave + 5*sig.  A channel of equal samples whose
    float mean rounds below them (977 x 0.1f or 1.1f in the reference's
    sequential order; 0.3f rounds above) with sig 0 has fraction 1.  The
    kernels decide it from the smallest sample, which must agree with the
    count against the HI of their own ave and sig; a NaN sample is never
    above HI."""
    vfrb = gpu
    rows, nch, ncol, nblk = 977, 64, 977, 1
    fb = np.random.default_rng(5).normal(0, 1, (rows, nch)).astype(np.float32)
    for c, v in ((9, 0.1), (10, 0.1), (11, 1.1), (12, 0.3)):
        fb[:, c] = np.float32(v)
    fb[100, 10] = np.nan
    rave, rsig, rout = orc.chan_stats(fb, ncol, nblk)
    assert rout[0, 9] == 1.0 and rout[0, 11] == 1.0 and rout[0, 12] == 0.0 and rout[0, 10] == 0.0
    d = vfrb.DeviceArray.from_host(fb)
    for ave, sig, out in (vfrb.chan_stats(d, rows, nch, ncol, nblk),):
        for c in (9, 11, 12):
            hi = np.float32(np.float64(ave[0, c]) + 5.0 * np.float64(sig[0, c]))
            assert out[0, c] == float(np.all(fb[:, c] > hi)), (c, ave[0, c], sig[0, c])
        assert out[0, 10] == 0.0
        normal = np.ones(nch, bool)
        normal[9:13] = False
        assert np.array_equal(out[0, normal], rout[0, normal])
        assert np.allclose(ave[0, normal], rave[0, normal], atol=1e-5)
        assert np.allclose(sig[0, normal], rsig[0, normal], rtol=1e-4)


# ------------------------------------------------------- dedispersion ----
def _c1_shape():
    nch, tchan = 1024, 64 * 2048 / 128e6
    chanbw = 64.0 / nch
    fch1 = 320.0 + nch * chanbw - 0.5 * chanbw
    dms = (np.arange(256) * 0.1).astype(np.float32)
    return nch, tchan, fch1, chanbw, dms


def test_delay_table_device_matches_oracle(gpu, orc):
    vfrb = gpu
    nch, tchan, fch1, chanbw, dms = _c1_shape()
    plan = vfrb.SearchPlan(nch, tchan, fch1, chanbw, dms, 976)
    off, maxd = orc.delay_table(dms, nch, fch1, chanbw, tchan)
    assert np.array_equal(plan.offsets(), off) and plan.max_delay == maxd == 308


@pytest.mark.parametrize("tile", range(26))
def test_dedisperse_f32_all_tiles(gpu, orc, tile):
    vfrb = gpu
    nch, tchan, fch1, chanbw, dms = _c1_shape()
    rows = 976
    rng = np.random.default_rng(tile)
    fb = rng.normal(0, 22, (rows, nch)).astype(np.float32)
    print(f"[case] dedisperse f32 tile={tile} ({vfrb.tile_names()[tile]})", flush=True)
    plan = vfrb.SearchPlan(nch, tchan, fch1, chanbw, dms, rows, vfrb.STORE_F32, tile)
    nt = rows - plan.max_delay
    d = vfrb.DeviceArray.from_host(fb)
    out = vfrb.DeviceArray((len(dms), nt), np.float32)
    plan.execute(rows, d, out, nt)
    got = out.download()
    ref = orc.dedisperse_f32(fb, plan.offsets(), nt)
    # same terms, different summation tree: f32 rounding only
    assert np.max(np.abs(got - ref)) < 1e-4 * 22


@pytest.mark.parametrize("storage", ["f32", "f16", "u8"])
def test_pair_tiles_match_channel_tiles(gpu, storage):
    """The pair- and group-presum tiles add the same terms as the per-channel
    tiles, each channel pair (group of 4) first: equal to f32 rounding on the
    same stored window.  The
    u8 accumulator holds sum(scale*byte), of the size of the channel offsets
    subtracted afterwards (about 60x the output here), so its summation-order
    rounding is that much larger relative to the output."""
    vfrb = gpu
    nch, tchan, fch1, chanbw, dms = _c1_shape()
    rows = 976
    rng = np.random.default_rng(31)
    fb = rng.normal(0, 22, (rows, nch)).astype(np.float32)
    plan = vfrb.SearchPlan(nch, tchan, fch1, chanbw, dms, rows, vfrb.STORAGE_NAMES[storage], 3)
    nt = rows - plan.max_delay
    d = vfrb.DeviceArray.from_host(fb)
    out = vfrb.DeviceArray((len(dms), nt), np.float32)
    plan.load(d, rows, bandpass=True)
    plan.dedisperse(rows, out, nt)
    ref = out.download()
    names = vfrb.tile_names()
    for i in (i for i, n in enumerate(names) if n[0] in "PGW"):
        print(f"[case] pair tile {names[i]} {storage}", flush=True)
        plan.set_tile(i)
        out.zero()
        plan.dedisperse(rows, out, nt)
        got = out.download()
        tol = (1e-3 if storage == "u8" else 1e-5) * np.max(np.abs(ref))
        assert np.max(np.abs(got - ref)) <= tol, (names[i], storage, np.max(np.abs(got - ref)))


@pytest.mark.parametrize("nch", [1021, 1022, 1023])
@pytest.mark.parametrize("storage", ["f32", "f16", "u8"])
def test_presum_tiles_ragged_channel_count(gpu, orc, storage, nch):
    """A channel count that is not a multiple of 4 (2) leaves the last group
    (pair) with missing channels, which the staging reads as zeros.  Every
    pair and group tile equals the per-channel tile on the same stored window
    (to f32 rounding; u8 as above), and in f32 the oracle's channel mean."""
    vfrb = gpu
    _, tchan, _, _, dms = _c1_shape()
    chanbw = 64.0 / 1024
    fch1 = 320.0 + nch * chanbw - 0.5 * chanbw
    rows = 976
    rng = np.random.default_rng(nch)
    fb = rng.normal(0, 22, (rows, nch)).astype(np.float32)
    plan = vfrb.SearchPlan(nch, tchan, fch1, chanbw, dms, rows, vfrb.STORAGE_NAMES[storage], 3)
    nt = rows - plan.max_delay
    d = vfrb.DeviceArray.from_host(fb)
    out = vfrb.DeviceArray((len(dms), nt), np.float32)
    plan.load(d, rows, bandpass=False)
    plan.dedisperse(rows, out, nt)
    ref = out.download()
    if storage == "f32":
        oref = orc.dedisperse_f32(fb, plan.offsets(), nt)
        assert np.max(np.abs(ref - oref)) < 1e-4 * 22
    names = vfrb.tile_names()
    for i in (i for i, n in enumerate(names) if n[0] in "PGW"):
        print(f"[case] ragged {nch} ch, tile {names[i]} {storage}", flush=True)
        plan.set_tile(i)
        out.zero()
        plan.dedisperse(rows, out, nt)
        got = out.download()
        tol = (1e-3 if storage == "u8" else 1e-5) * np.max(np.abs(ref))
        assert np.max(np.abs(got - ref)) <= tol, (names[i], storage, nch, np.max(np.abs(got - ref)))


@pytest.mark.parametrize("storage", ["f32", "f16", "u8"])
def test_wide_group_tiles_equal_group_tiles(gpu, storage):
    """The wide group tiles ("W") stage the same presum rows as the group
    tiles ("G") and add them in the same order, reading four samples per lane
    at a time: equal bit for bit, on a window whose DM starts fall on every
    phase of a 4-sample quad, at a length that leaves a partial last tile."""
    vfrb = gpu
    nch, tchan, fch1, chanbw, dms = _c1_shape()
    rows = 976
    rng = np.random.default_rng(41)
    fb = rng.normal(0, 22, (rows, nch)).astype(np.float32)
    names = vfrb.tile_names()
    plan = vfrb.SearchPlan(nch, tchan, fch1, chanbw, dms, rows, vfrb.STORAGE_NAMES[storage],
                           names.index("G512_dm64_w16"))
    nt = rows - plan.max_delay
    phases = np.unique(plan.offsets()[:, ::4] % 4)
    assert phases.size == 4, phases
    d = vfrb.DeviceArray.from_host(fb)
    out = vfrb.DeviceArray((len(dms), nt), np.float32)
    plan.load(d, rows, bandpass=True)
    plan.dedisperse(rows, out, nt)
    ref = out.download()
    for wide in (n for n in names if n[0] == "W"):
        print(f"[case] wide tile {wide} {storage}", flush=True)
        plan.set_tile(names.index(wide))
        out.zero()
        plan.dedisperse(rows, out, nt)
        assert np.array_equal(out.download(), ref), wide


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_presum_tiles_irregular_dm_list(gpu, orc, storage):
    """200 DMs (a partial last DM block for every tile) at irregular spacing
    (a sorted random list, two DMs repeated), a window whose output length is
    not a multiple of any tile: every pair, group and wide tile that fits the
    list equals the per-channel tile to f32 rounding, the wide tiles equal the
    group tiles bit for bit, and in f32 the oracle's channel mean."""
    vfrb = gpu
    nch, tchan, fch1, chanbw, _ = _c1_shape()
    rng = np.random.default_rng(77)
    dms = np.sort(np.concatenate([rng.uniform(0.0, 25.0, 198), [3.0, 3.0]])).astype(np.float32)
    rows = 901
    fb = rng.normal(0, 22, (rows, nch)).astype(np.float32)
    plan = vfrb.SearchPlan(nch, tchan, fch1, chanbw, dms, rows, vfrb.STORAGE_NAMES[storage], 3)
    nt = rows - plan.max_delay
    assert all(nt % t for t in (252, 256, 504, 512))
    d = vfrb.DeviceArray.from_host(fb)
    out = vfrb.DeviceArray((len(dms), nt), np.float32)
    plan.load(d, rows, bandpass=False)
    plan.dedisperse(rows, out, nt)
    ref = out.download()
    if storage == "f32":
        oref = orc.dedisperse_f32(fb, plan.offsets(), nt)
        assert np.max(np.abs(ref - oref)) < 1e-4 * 22
    names = vfrb.tile_names()
    got = {}
    for i in (i for i, n in enumerate(names) if n[0] in "PGW"):
        try:
            plan.set_tile(i)
        except vfrb.VfrbError:
            continue  # the list is too coarse for this tile
        print(f"[case] irregular DM list, tile {names[i]} {storage}", flush=True)
        out.zero()
        plan.dedisperse(rows, out, nt)
        got[names[i]] = out.download()
        assert np.max(np.abs(got[names[i]] - ref)) <= 1e-5 * np.max(np.abs(ref)), names[i]
    for w, g in (("W504_dm64_w16", "G512_dm64_w16"), ("W252_dm64_w16", "G256_dm64_w16")):
        assert w in got and g in got
        assert np.array_equal(got[w], got[g]), (w, g)


def test_autotune_picks_a_fitting_tile_and_keeps_results(gpu, orc):
    vfrb = gpu
    nch, tchan, fch1, chanbw, dms = _c1_shape()
    rows = 976
    rng = np.random.default_rng(21)
    fb = rng.normal(0, 22, (rows, nch)).astype(np.float32)
    for st in ("f32", "f16"):
        plan = vfrb.SearchPlan(nch, tchan, fch1, chanbw, dms, rows, vfrb.STORAGE_NAMES[st])
        nt = rows - plan.max_delay
        d = vfrb.DeviceArray.from_host(fb)
        out = vfrb.DeviceArray((len(dms), nt), np.float32)
        plan.load(d, rows, bandpass=False)
        plan.dedisperse(rows, out, nt)
        before = out.download()
        ms = plan.autotune(rows, out, nt, iters=1)
        assert plan.tile >= 0 and ms[vfrb.tile_names()[plan.tile]] == min(v for v in ms.values() if v > 0)
        plan.dedisperse(rows, out, nt)
        after = out.download()
        assert np.max(np.abs(after - before)) < 1e-4 * 22


def test_load_bandpass_then_dedisperse_matches_oracle_chain(gpu, orc):
    vfrb = gpu
    nch, tchan, fch1, chanbw, dms = _c1_shape()
    rows = 976
    rng = np.random.default_rng(4)
    fb = rng.normal(256, 22, (rows, nch)).astype(np.float32)
    plan = vfrb.SearchPlan(nch, tchan, fch1, chanbw, dms, rows)
    nt = rows - plan.max_delay
    d = vfrb.DeviceArray.from_host(fb)
    plan.load(d, rows, bandpass=True)
    out = vfrb.DeviceArray((len(dms), nt), np.float32)
    plan.dedisperse(rows, out, nt)
    got = out.download()
    bp = orc.bandpass_f32(fb)
    ref = orc.dedisperse_f32(bp, plan.offsets(), nt)
    assert np.max(np.abs(got - ref)) < 2e-3


def test_load_from_ring(gpu, orc):
    vfrb = gpu
    nch, tchan, fch1, chanbw, dms = _c1_shape()
    rows, ring, row0 = 700, 1000, 800
    rng = np.random.default_rng(6)
    fb = rng.normal(256, 22, (ring, nch)).astype(np.float32)
    plan = vfrb.SearchPlan(nch, tchan, fch1, chanbw, dms, rows)
    nt = rows - plan.max_delay
    d = vfrb.DeviceArray.from_host(fb)
    plan.load(d, rows, ring_rows=ring, row0=row0, bandpass=True)
    out = vfrb.DeviceArray((len(dms), nt), np.float32)
    plan.dedisperse(rows, out, nt)
    window = fb[[(row0 + r) % ring for r in range(rows)]]
    ref = orc.dedisperse_f32(orc.bandpass_f32(window), plan.offsets(), nt)
    assert np.max(np.abs(out.download() - ref)) < 2e-3


# Tolerances, as max |err| / std(dedispersed series) over 171k samples:
#  f16: rounding 2^-11 relative per element, averaged over 1024 channels.
#  u8 : per-channel step = range/255 ~ 9 sigma_c/255, uniform rounding error
#       sigma_c*9/(255*sqrt(12)) = 1.0% of the series std after the channel
#       mean; the maximum over 1.7e5 Gaussian samples is ~4.6 of those std.
# The RMS error follows the same model without the extreme-value factor (the
# channel mean divides error and noise alike, so element ratios carry over):
#  f16: rounding is uniform over one ulp, at most 2^-10 |x|, so its RMS is at
#       most 2^-10/sqrt(12) = 2.8e-4 of the elements' spread; the bound 4e-4
#       leaves room for the f32 bandpass's own rounding (measured 2.9e-4);
#  u8 : step/sqrt(12) with step = range/255; the range of 976 Gaussian samples
#       is about 6.4 sigma, i.e. 7.2e-3 (measured 7.3e-3); the bound 1.3e-2
#       holds for ranges up to 11 sigma.
@pytest.mark.parametrize("storage,tol,rms_tol", [("f16", 2e-3, 4e-4), ("u8", 6e-2, 1.3e-2)])
def test_reduced_precision_storage(gpu, orc, storage, tol, rms_tol):
    vfrb = gpu
    nch, tchan, fch1, chanbw, dms = _c1_shape()
    rows = 976
    rng = np.random.default_rng(8)
    fb = rng.normal(256, 22, (rows, nch)).astype(np.float32)
    fb[:, 40] += 3.0e6            # strong RFI channel: exercises the scale logic
    plan = vfrb.SearchPlan(nch, tchan, fch1, chanbw, dms, rows, vfrb.STORAGE_NAMES[storage])
    nt = rows - plan.max_delay
    d = vfrb.DeviceArray.from_host(fb)
    plan.load(d, rows, bandpass=True)
    out = vfrb.DeviceArray((len(dms), nt), np.float32)
    plan.dedisperse(rows, out, nt)
    got = out.download()
    ref = orc.dedisperse_f64(orc.bandpass_f64(fb.astype(np.float64)), plan.offsets(), nt)
    sd = ref.std()
    # error relative to the dedispersed noise level: its maximum and its RMS
    assert np.max(np.abs(got - ref)) / sd < tol, np.max(np.abs(got - ref)) / sd
    rms = float(np.sqrt(np.mean((got - ref) ** 2))) / sd
    print(f"[case] {storage}: max |err|/sd {np.max(np.abs(got - ref)) / sd:.3e}, rms/sd {rms:.3e}", flush=True)
    assert rms < rms_tol, rms


@pytest.mark.parametrize("storage", ["f32", "f16", "u8"])
def test_storage_exact_on_representable_window(gpu, storage):
    """Each storage mode is exact where its format holds the window exactly.
    f32/f16: integers in [-1000, 1000] (exact in f16, power-of-two scale 1).
    u8: channel c holds lo_c + 2^-(c mod 4) * k, k in 0..255 with both ends
    present, lo_c an integer: the range gives scale 2^-(c mod 4) and offset
    lo_c exactly, so every byte is k.  All partial sums are multiples of 1/8
    below 2^21, exact in f32, and the channel mean divides by 1024: every
    output equals the exact dedispersed mean bit for bit, in every tile
    family (what the reduced-precision tolerances above do not pin)."""
    vfrb = gpu
    nch, tchan, fch1, chanbw, dms = _c1_shape()
    rows = 976
    rng = np.random.default_rng(12)
    if storage == "u8":
        k = rng.integers(0, 256, (rows, nch))
        k[3] = 0
        k[7] = 255
        j = np.arange(nch) % 4
        lo = rng.integers(-300, 301, nch)
        units = lo[None, :] * 8 + k * (8 >> j)[None, :]  # window in units of 1/8
    else:
        units = rng.integers(-1000, 1001, (rows, nch)) * 8
    fb = (units / 8.0).astype(np.float32)
    assert np.array_equal(fb.astype(np.float64) * 8, units)
    names = vfrb.tile_names()
    plan = vfrb.SearchPlan(nch, tchan, fch1, chanbw, dms, rows, vfrb.STORAGE_NAMES[storage])
    nt = rows - plan.max_delay
    off = plan.offsets().astype(np.int64)
    cols = np.arange(nch)
    t = np.arange(nt)
    exp = np.empty((len(dms), nt), np.float32)
    for d in range(len(dms)):
        s = units[t[:, None] + off[d][None, :], cols[None, :]].sum(axis=1)
        exp[d] = (s / 8.0 / nch).astype(np.float32)
        assert np.array_equal(exp[d].astype(np.float64) * 8 * nch, s)  # exact in f32
    d_fb = vfrb.DeviceArray.from_host(fb)
    plan.load(d_fb, rows, bandpass=False)
    out = vfrb.DeviceArray((len(dms), nt), np.float32)
    for tile in ("T512_dm64_w16", "P512_dm64_w16", "G256_dm64_w16", "W504_dm64_w16", "W252_dm64_w16"):
        print(f"[case] exact window {storage} tile {tile}", flush=True)
        plan.set_tile(names.index(tile))
        out.zero()
        plan.dedisperse(rows, out, nt)
        got = out.download()
        bad = np.argwhere(got != exp)
        assert bad.size == 0, (storage, tile, bad[:5], got[tuple(bad[0])], exp[tuple(bad[0])])


def test_dedisperse_stats_variant(gpu, orc):
    vfrb = gpu
    nch, tchan, fch1, chanbw = 3840, 4 * 12500 / 128e6, 64.0 / 6250, None
    chanbw = 64.0 / 6250
    fch1 = 320.0 + nch * chanbw - 0.5 * chanbw
    dms = np.array([0.0, 119.3], np.float32)
    off, maxd = orc.delay_table(dms, nch, fch1, chanbw, tchan)
    rows = maxd + 500
    rng = np.random.default_rng(3)
    fb = rng.normal(0, 5, (rows, nch)).astype(np.float32)
    plan = vfrb.SearchPlan(nch, tchan, fch1, chanbw, dms, rows)
    nt = rows - plan.max_delay
    d = vfrb.DeviceArray.from_host(fb)
    plan.load(d, rows, bandpass=False)
    out, sig = vfrb.DeviceArray((2, nt), np.float32), vfrb.DeviceArray((2, nt), np.float32)
    plan.dedisperse_stats(rows, out, sig, nt)
    rout, rsig = orc.dedisperse_stats_f32(fb, off, nt)
    assert np.allclose(out.download(), rout, atol=1e-4)
    assert np.allclose(sig.download(), rsig, rtol=1e-3)


@pytest.mark.parametrize("mean,rtol", [(0.0, 2e-5), (40.0, 2e-5), (-300.0, 2e-5)])
def test_dedisperse_stats_tiles(gpu, orc, mean, rtol):
    """The tiled dedisperseSTATS over several DM blocks and time tiles, a
    window longer than one tile row and a ragged last tile: ave within float
    summation order of the oracle; sigma from one-pass sums of x - K, with K
    the window's level (0 for a centred window), so they do not cancel: the
    same accuracy at mean 40 and -300 (not bandpassed) as at mean 0 (the
    bandpassed window the programs feed it)."""
    vfrb = gpu
    nch, tchan = 1000, 4 * 12500 / 128e6
    chanbw = 64.0 / 6250
    fch1 = 320.0 + nch * chanbw - 0.5 * chanbw
    dms = np.linspace(0.0, 90.0, 77).astype(np.float32)
    off, maxd = orc.delay_table(dms, nch, fch1, chanbw, tchan)
    rows = maxd + 1333
    rng = np.random.default_rng(31)
    fb = (rng.normal(0, 5, (rows, nch)) + mean).astype(np.float32)
    fb[maxd // 2, 17] += 300.0  # one outlier seen by many (dm, t)
    plan = vfrb.SearchPlan(nch, tchan, fch1, chanbw, dms, rows)
    nt = rows - plan.max_delay
    d = vfrb.DeviceArray.from_host(fb)
    plan.load(d, rows, bandpass=False)
    ld = nt + 5
    out, sig = vfrb.DeviceArray((len(dms), ld), np.float32), vfrb.DeviceArray((len(dms), ld), np.float32)
    plan.dedisperse_stats(rows, out, sig, ld)
    rout, rsig = orc.dedisperse_stats_f32(fb, off, nt)
    go, gs = out.download()[:, :nt], sig.download()[:, :nt]
    assert np.max(np.abs(go - rout)) <= 1e-5 * max(1.0, abs(mean)) + 1e-6
    assert np.max(np.abs(gs - rsig) / rsig) < rtol, np.max(np.abs(gs - rsig) / rsig)
    # the plain dedispersion of the same window is the mean series
    o2 = vfrb.DeviceArray((len(dms), nt), np.float32)
    plan.dedisperse(rows, o2, nt)
    assert np.max(np.abs(o2.download() - go)) <= 1e-5 * max(1.0, abs(mean)) + 1e-6


# --------------------------------------------------------- statistics ----
def test_dm_stats_f64_matches_reference_arithmetic(gpu, orc):
    vfrb = gpu
    rng = np.random.default_rng(10)
    ndm, nt = 64, 19531
    x = rng.normal(3.0, 0.7, (ndm, nt)).astype(np.float32)
    x[5, 1234] += 9.0
    x[7] = 2.0  # constant series: sigma 0
    d = vfrb.DeviceArray.from_host(x)
    st = vfrb.dm_stats(d, nt, ndm, nt, vfrb.ACC_F64)
    frb, ave, sig, peak, pt = orc.dm_stats(x)
    ok = np.arange(ndm) != 7
    # FP64 sums of the same floats: the float results agree to the last bit
    # except where the double sum sits on a float rounding boundary.
    assert np.max(np.abs(st.ave - ave)[ok] / np.abs(ave[ok])) < 2e-7
    assert np.max(np.abs(st.sigma - sig)[ok] / sig[ok]) < 2e-7
    assert np.array_equal(st.frb, frb) and np.array_equal(st.peak_t, pt)
    assert np.allclose(st.peak, peak, rtol=1e-6)
    assert st.peak_t[5] == 1234
    stc = vfrb.dm_stats(d, nt, ndm, nt, vfrb.ACC_F32C)
    assert np.max(np.abs(stc.sigma - sig)[ok] / sig[ok]) < 1e-5
    assert np.array_equal(stc.peak_t[ok], pt[ok])


def test_candidates(gpu):
    vfrb = gpu
    rng = np.random.default_rng(12)
    ndm, nt = 32, 5000
    x = rng.normal(0, 1, (ndm, nt)).astype(np.float32)
    x[3, 100] = 12.0
    x[20, 4000] = 9.0
    d = vfrb.DeviceArray.from_host(x)
    st = vfrb.dm_stats(d, nt, ndm, nt)
    n, cdm, ct, cs = vfrb.candidates(d, nt, ndm, nt, st.ave, st.sigma, 7.0)
    snr = (x - st.ave[:, None]) / st.sigma[:, None]
    want = set(zip(*np.nonzero(snr > 7.0)))
    assert n == len(want) and set(zip(cdm.tolist(), ct.tolist())) == want


def test_default_tile_falls_back_for_coarse_dm_lists(gpu, orc):
    """A DM list coarse enough that adjacent channels' relative delay varies
    by more than 4 rows across a 64-DM block (DM step 5 at 1024 ch) cannot
    use the channel-pair tile: a default plan, and a default batch, fall
    back to a tile that fits, with the same results."""
    vfrb = gpu
    nch, tchan, fch1, chanbw, _ = _c1_shape()
    dms = (np.arange(64) * 5.0).astype(np.float32)
    rows = 4200
    rng = np.random.default_rng(41)
    fb = rng.normal(0, 22, (rows, nch)).astype(np.float32)
    print("[case] default tile fallback", flush=True)
    plan = vfrb.SearchPlan(nch, tchan, fch1, chanbw, dms, rows)
    names = vfrb.tile_names()
    assert names[plan.tile] in ("G256_dm64_w16", "T512_dm64_w16")
    keep = plan.tile
    with pytest.raises(vfrb.VfrbError):
        plan.set_tile(names.index("P512_dm64_w16"))
    assert plan.tile == keep
    nt = rows - plan.max_delay
    d = vfrb.DeviceArray.from_host(fb)
    plan.load(d, rows, bandpass=False)
    ref = orc.dedisperse_f32(fb, plan.offsets(), nt)
    out = vfrb.DeviceArray((len(dms), nt), np.float32)
    plan.dedisperse(rows, out, nt)
    assert np.max(np.abs(out.download() - ref)) < 1e-4 * 22
    out.zero()
    batch = vfrb.Batch(vfrb.STORE_F32)
    batch.set([(plan, rows, out, nt)])
    assert names[batch.tile] in ("G256_dm64_w16", "T512_dm64_w16")
    batch.dedisperse()
    vfrb.sync()
    assert np.max(np.abs(out.download() - ref)) < 1e-4 * 22
