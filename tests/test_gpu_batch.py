"""GPU: batched persistent dedispersion (vfrb_batch_*) against the same plans
run one at a time, the oracle, and hipGraph replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NCH = 1024
TCHAN = 64 * 2048 / 128e6
CHANBW = 64.0 / NCH
FCH1 = 320.0 + NCH * CHANBW - 0.5 * CHANBW

# uneven windows and DM ranges: (rows, ndm, max DM)
JOBS = [(976, 256, 25.5), (520, 64, 6.3), (800, 200, 19.9), (450, 32, 3.1), (700, 128, 12.7)]


def _make_jobs(vfrb, storage, seed):
    rng = np.random.default_rng(seed)
    jobs, fbs = [], []
    for rows, ndm, dmax in JOBS:
        dms = np.linspace(0.0, dmax, ndm).astype(np.float32)
        plan = vfrb.SearchPlan(NCH, TCHAN, FCH1, CHANBW, dms, rows, vfrb.STORAGE_NAMES[storage])
        fb = rng.normal(256, 22, (rows, NCH)).astype(np.float32)
        d = vfrb.DeviceArray.from_host(fb)
        plan.load(d, rows, bandpass=True)
        nt = rows - plan.max_delay
        out = vfrb.DeviceArray((ndm, nt + 7), np.float32)  # ld_out > nt
        jobs.append((plan, rows, out, nt + 7))
        fbs.append((fb, d))
    return jobs, fbs


def _singles(vfrb, jobs):
    res = []
    for plan, rows, out, ld in jobs:
        o = vfrb.DeviceArray(out.shape, np.float32)
        o.zero()
        plan.dedisperse(rows, o, ld)
        res.append(o.download()[:, : rows - plan.max_delay])
    return res


def _batch_results(jobs):
    return [out.download()[:, : rows - plan.max_delay] for plan, rows, out, _ in jobs]


@pytest.mark.parametrize("tile", [None, "W504_dm64_w16", "W252_dm64_w16"])
@pytest.mark.parametrize("storage", ["f32", "f16", "u8"])
def test_batch_equals_single_plans(gpu, orc, storage, tile):
    """The default batch (group tile) and batches fixed to the wide group
    tiles: bit-identical to each plan run alone on the batch's tile."""
    vfrb = gpu
    print(f"[case] batch vs single plans storage={storage} tile={tile}", flush=True)
    jobs, fbs = _make_jobs(vfrb, storage, seed=11)
    t = -1 if tile is None else vfrb.tile_names().index(tile)
    batch = vfrb.Batch(vfrb.STORAGE_NAMES[storage], tile=t)
    ntiles = batch.set(jobs)
    assert ntiles > 0 and batch.grid > 0
    for _, _, out, _ in jobs:
        out.zero()
    batch.dedisperse()
    vfrb.sync()
    got = _batch_results(jobs)
    # same tile configuration and summation order: bit-identical to one plan alone
    for g, s in zip(got, _singles(vfrb, jobs)):
        assert np.array_equal(g, s)
    if storage == "f32":
        for (plan, rows, _, _), (fb, _), g in zip(jobs, fbs, got):
            ref = orc.dedisperse_f32(orc.bandpass_f32(fb), plan.offsets(), rows - plan.max_delay)
            assert np.max(np.abs(g - ref)) < 2e-3


@pytest.mark.parametrize("storage", ["f32", "f16", "u8"])
def test_batch_graph_replays_match_eager(gpu, storage):
    vfrb = gpu
    print(f"[case] batch graph replay {storage}", flush=True)
    jobs, fbs = _make_jobs(vfrb, storage, seed=12)
    batch = vfrb.Batch(vfrb.STORAGE_NAMES[storage])
    s = vfrb.Stream()
    batch.set(jobs, stream=s)
    batch.dedisperse(stream=s)
    s.sync()
    eager = _batch_results(jobs)
    for k in range(3):
        for _, _, out, _ in jobs:
            out.zero()
        batch.graph_launch(s)
        s.sync()
        for g, e in zip(_batch_results(jobs), eager):
            assert np.array_equal(g, e), f"replay {k}"
        _singles(vfrb, jobs)  # eager work between replays
    # new job set (fewer, shorter windows) through the same captured graph
    jobs2 = [(p, r - 100, o, ld) for p, r, o, ld in jobs[:3]]
    batch.set(jobs2, stream=s)
    for _, _, out, _ in jobs2:
        out.zero()
    batch.graph_launch(s)
    s.sync()
    for g, e in zip(_batch_results(jobs2), _singles(vfrb, jobs2)):
        assert np.array_equal(g, e)


def test_batch_autotune_keeps_results(gpu):
    vfrb = gpu
    print("[case] batch autotune", flush=True)
    jobs, _ = _make_jobs(vfrb, "f32", seed=13)
    batch = vfrb.Batch(vfrb.STORE_F32)
    batch.set(jobs)
    ms = batch.autotune(iters=1)
    names = vfrb.tile_names()
    assert batch.tile >= 0 and ms[names[batch.tile]] == min(v for v in ms.values() if v > 0)
    for _, _, out, _ in jobs:
        out.zero()
    batch.dedisperse()
    vfrb.sync()
    for plan, *_ in jobs:
        assert plan.tile == batch.tile
    for g, s in zip(_batch_results(jobs), _singles(vfrb, jobs)):
        assert np.array_equal(g, s)


def test_batch_rejects_bad_jobs(gpu):
    vfrb = gpu
    jobs, _ = _make_jobs(vfrb, "f32", seed=14)
    small = vfrb.Batch(vfrb.STORE_F32, max_jobs=2, max_tiles=8)
    with pytest.raises(vfrb.VfrbError) as e:
        small.set(jobs)
    assert e.value.status == vfrb.ERR_CAPACITY
    with pytest.raises(vfrb.VfrbError) as e:
        small.set(jobs[:2])  # two jobs fit, their work items do not
    assert e.value.status == vfrb.ERR_CAPACITY
    other = vfrb.Batch(vfrb.STORE_F16)
    with pytest.raises(vfrb.VfrbError):
        other.set(jobs[:1])  # storage mismatch
    b = vfrb.Batch(vfrb.STORE_F32)
    plan, rows, out, ld = jobs[0]
    with pytest.raises(vfrb.VfrbError):
        b.set([(plan, rows + 1, out, ld)])  # more rows than loaded
    with pytest.raises(vfrb.VfrbError):
        b.dedisperse()  # no jobs set after the failure


# DM lists at 1024 ch x 1.024 ms rows whose 64-DM blocks fit only some tiles:
# DM step 2 is too coarse for the group tile (a group's presum rows exceed
# the LDS budget) but fits the pair tile; DM step 4 spans more than 4 rows of
# relative delay per channel pair, so only the per-channel tile fits; DM step
# 0.1 fits the default group tile.
FALLBACK_LISTS = [("P512_dm64_w16", 2.0, 64, 2200), ("T512_dm64_w16", 4.0, 64, 3600),
                  ("G256_dm64_w16", 0.1, 256, 976)]


def test_default_batch_walks_the_fallback_chain_on_every_set(gpu, orc):
    """A default-tiled batch restarts from its preferred tile on every set():
    a coarse list that moved it to the pair tile, then one that needs the
    per-channel tile, do not keep a later fine list off the group tile."""
    vfrb = gpu
    names = vfrb.tile_names()
    rng = np.random.default_rng(21)
    batch = vfrb.Batch(vfrb.STORE_F32)
    assert names[batch.tile] == "G256_dm64_w16"
    keep = []
    for want, step, ndm, rows in FALLBACK_LISTS + FALLBACK_LISTS[:1] + FALLBACK_LISTS[2:]:
        print(f"[case] batch fallback set: {want} (DM step {step})", flush=True)
        dms = (np.arange(ndm) * step).astype(np.float32)
        plan = vfrb.SearchPlan(NCH, TCHAN, FCH1, CHANBW, dms, rows)
        assert names[plan.tile] == want
        fb = rng.normal(0, 22, (rows, NCH)).astype(np.float32)
        d = vfrb.DeviceArray.from_host(fb)
        plan.load(d, rows, bandpass=False)
        nt = rows - plan.max_delay
        out = vfrb.DeviceArray((ndm, nt), np.float32)
        out.zero()
        batch.set([(plan, rows, out, nt)])
        assert names[batch.tile] == want
        batch.dedisperse()
        vfrb.sync()
        got = out.download()
        single = vfrb.DeviceArray((ndm, nt), np.float32)
        plan.dedisperse(rows, single, nt)
        assert np.array_equal(got, single.download())  # same tile: same summation order
        ref = orc.dedisperse_f32(fb, plan.offsets(), nt)
        assert np.max(np.abs(got - ref)) < 1e-4 * 22
        keep.append((plan, rows, fb, d))
    # a mixed set goes to the first tile that fits every job (the fine plan's
    # tables are rebuilt for it), then the fine list alone returns to the group tile
    print("[case] batch fallback: mixed set", flush=True)
    mixed = [keep[2], keep[0]]
    outs = [vfrb.DeviceArray((p.ndm, r - p.max_delay), np.float32) for p, r, _, _ in mixed]
    batch.set([(p, r, o, r - p.max_delay) for (p, r, _, _), o in zip(mixed, outs)])
    assert names[batch.tile] == "P512_dm64_w16"
    batch.dedisperse()
    vfrb.sync()
    for (p, r, fb, _), o in zip(mixed, outs):
        ref = orc.dedisperse_f32(fb, p.offsets(), r - p.max_delay)
        assert np.max(np.abs(o.download() - ref)) < 1e-4 * 22
    p, r, _, _ = keep[2]
    batch.set([(p, r, outs[0], r - p.max_delay)])
    assert names[batch.tile] == "G256_dm64_w16" and names[p.tile] == "G256_dm64_w16"
