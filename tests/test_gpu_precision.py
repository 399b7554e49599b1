"""Precision of the reduced-storage search against the oracle's FP64 chain on
the 1 s, 1024-channel x 256-DM workload (one pulse at DM 12.3, FWHM 2 ms,
SNR 20, seed 0): per-DM peak SNR and the candidate (DM, sample) lists at
7 sigma."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THRESHOLD = 7.0


@pytest.fixture(scope="module")
def c3(gpu, orc):
    vfrb = gpu
    n, nave = 2048, 64
    p = vfrb.params_derive(n, nave, 1, 1.0, fullband=True)
    rows = int(1.0 / p.tchan)
    dms = (np.arange(256) * 0.1).astype(np.float32)
    amp = vfrb.amp_for_snr(20.0, nave, p.nchanout)
    spec = vfrb.synth_spec(n, seed=0, pulses=[(12.3, 0.35, 0.002, amp)])
    p0 = vfrb.synth_payload(spec, 0, 0, rows * nave)
    p1 = vfrb.synth_payload(spec, 1, 0, rows * nave)
    off, max_delay, _ = vfrb.delay_table(dms, p.nchanout, p.fch1, p.chanbw, p.tchan)
    nt = rows - max_delay
    out64 = orc.dedisperse_f64(orc.bandpass_f64(orc.filterbank_f64(p0, p1, rows, n, nave, p.nskip, p.nchanout)),
                               off, nt)
    ave64, sig64, peak64, pt64 = orc.dm_stats_f64(out64)
    snr64 = (out64 - ave64[:, None]) / sig64[:, None]
    ref = set(zip(*[a.tolist() for a in np.nonzero(snr64 > THRESHOLD)]))
    d0, d1 = vfrb.DeviceArray.from_host(p0), vfrb.DeviceArray.from_host(p1)
    fb = vfrb.DeviceArray((rows, p.nchanout), np.float32)
    vfrb.FilterbankPlan(n, nave, p.nskip, p.nchanout).execute(d0, d1, rows, fb)
    return dict(p=p, rows=rows, dms=dms, nt=nt, fb=fb, peak64=peak64, pt64=pt64, snr64=snr64, ref=ref)


@pytest.mark.parametrize("storage,acc,tol", [("f32", "ACC_F64", 1e-5), ("f16", "ACC_F32C", 1e-3),
                                             ("u8", "ACC_F32C", 2e-2)])
def test_snr_error_and_candidates_vs_fp64(gpu, c3, storage, acc, tol):
    vfrb = gpu
    print(f"[case] precision {storage}", flush=True)
    p, rows, dms, nt = c3["p"], c3["rows"], c3["dms"], c3["nt"]
    sp = vfrb.SearchPlan(p.nchanout, p.tchan, p.fch1, p.chanbw, dms, rows, vfrb.STORAGE_NAMES[storage])
    sp.load(c3["fb"], rows, bandpass=True)
    out = vfrb.DeviceArray((len(dms), nt), np.float32)
    sp.dedisperse(rows, out, nt)
    st = vfrb.dm_stats(out, nt, len(dms), nt, acc=getattr(vfrb, acc))
    # SNR error vs FP64 for every DM trial (tolerance: f16 is the target's
    # 1e-3; u8's 1/255-of-range steps are ~1% of a channel's sigma)
    rel = np.abs(st.peak.astype(np.float64) - c3["peak64"]) / c3["peak64"]
    assert rel.max() < tol, (storage, rel.max(), int(np.argmax(rel)))
    best = int(np.argmax(c3["peak64"]))
    assert dms[best] == pytest.approx(12.3, abs=0.11)
    assert int(np.argmax(st.peak)) == best and st.peak_t[best] == c3["pt64"][best]
    cnt, cdm, ct, _ = vfrb.candidates(out, nt, len(dms), nt, st.ave, st.sigma, THRESHOLD)
    got = set(zip(cdm.tolist(), ct.tolist()))
    # identical lists; the FP64 SNRs nearest the threshold say how much error a
    # flip would need
    margin = float(np.min(np.abs(c3["snr64"] - THRESHOLD)))
    assert cnt == len(got) and got == c3["ref"], (storage, len(got ^ c3["ref"]), margin)
