"""Pins the CPU oracle (oracle/vfrb_oracle.c) against closed forms, brute
force and independent library routines, so that the GPU parity tests compare
against something known to be right."""
from decimal import Decimal, getcontext

import numpy as np
import pytest


# ---------------------------------------------------------------- FFT ----
@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12, 15, 16, 20, 30, 64, 100, 125, 128, 250, 256])
def test_fft_matches_bruteforce_dft(orc, n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    X = orc.fft(x)
    D = orc.dft(x)
    assert np.max(np.abs(X - D)) <= 1e-12 * max(1.0, np.max(np.abs(D)))
    # numpy's pocketfft as an independent library check
    assert np.allclose(X, np.fft.fft(x), rtol=1e-12, atol=1e-12 * n)


@pytest.mark.parametrize("n", [16, 60, 2048])
def test_dft_closed_forms(orc, n):
    j0, k0 = 3, 5
    imp = np.zeros(n, complex)
    imp[j0] = 1.0
    k = np.arange(n)
    X = orc.fft(imp)
    assert np.allclose(X, np.exp(-2j * np.pi * j0 * k / n), atol=1e-12)
    tone = np.exp(2j * np.pi * k0 * k / n)
    T = orc.fft(tone)
    expect = np.zeros(n, complex)
    expect[k0] = n
    assert np.allclose(T, expect, atol=1e-9)
    c = orc.fft(np.full(n, 2.5 + 0j))
    assert abs(c[0] - 2.5 * n) < 1e-9 and np.max(np.abs(c[1:])) < 1e-9
    if n <= 64:   # the brute-force DFT itself obeys the same closed forms
        assert np.allclose(orc.dft(imp), np.exp(-2j * np.pi * j0 * k / n), atol=1e-12)
    rng = np.random.default_rng(1)
    x = rng.normal(size=n) + 0j
    assert abs(np.sum(np.abs(x) ** 2) - np.sum(np.abs(orc.fft(x)) ** 2) / n) < 1e-9 * n


def test_fft_large_matches_numpy(orc):
    rng = np.random.default_rng(7)
    for n in (8192, 12500):
        x = rng.integers(0, 256, n).astype(float)
        assert np.allclose(orc.fft(x), np.fft.fft(x), rtol=1e-10, atol=1e-7)


# ---------------------------------------------------------- filterbank ----
def _fb_numpy(p0, p1, nrows, n, nave, nskip, nch):
    """Independent restatement with numpy.fft.rfft (PIPELINE:1292-1322)."""
    x0 = p0[: nrows * nave * n].reshape(nrows, nave, n).astype(np.float64)
    x1 = p1[: nrows * nave * n].reshape(nrows, nave, n).astype(np.float64)
    P = (np.abs(np.fft.rfft(x0, axis=-1)) ** 2 + np.abs(np.fft.rfft(x1, axis=-1)) ** 2) / (n // 2 + 1)
    return P.mean(axis=1)[:, nskip:nskip + nch]


def test_filterbank_f64_matches_numpy(orc):
    rng = np.random.default_rng(3)
    n, nave, nrows = 256, 4, 5
    p0 = rng.integers(0, 256, nrows * nave * n).astype(np.uint8)
    p1 = rng.integers(0, 256, nrows * nave * n).astype(np.uint8)
    got = orc.filterbank_f64(p0, p1, nrows, n, nave, 1, n // 2)
    assert np.allclose(got, _fb_numpy(p0, p1, nrows, n, nave, 1, n // 2), rtol=1e-11)
    # non-power-of-two length through the mixed-radix path
    n = 250
    p0 = rng.integers(0, 256, nrows * nave * n).astype(np.uint8)
    p1 = rng.integers(0, 256, nrows * nave * n).astype(np.uint8)
    got = orc.filterbank_f64(p0, p1, nrows, n, nave, 3, 100)
    assert np.allclose(got, _fb_numpy(p0, p1, nrows, n, nave, 3, 100), rtol=1e-11)


def test_filterbank_closed_forms(orc):
    n, nave, nrows = 64, 2, 3
    nfreq = n // 2 + 1
    # constant input: only the DC bin (never selected) is non-zero
    c = np.full(nrows * nave * n, 128, np.uint8)
    assert np.max(np.abs(orc.filterbank_f64(c, c, nrows, n, nave, 1, n // 2))) < 1e-9
    # alternating +-64 on pol 0: all power in the Nyquist bin, X[N/2] = 64*N
    alt = (128 + 64 * (-1) ** np.arange(nrows * nave * n)).astype(np.uint8)
    fb = orc.filterbank_f64(alt, c, nrows, n, nave, 1, n // 2)
    expect = np.zeros_like(fb)
    expect[:, n // 2 - 1] = (64.0 * n) ** 2 / nfreq
    assert np.allclose(fb, expect, atol=1e-6)
    # period-4 pattern 128 + a*cos(pi n/2) on pol 1: bin N/4 with |X| = a*N/2
    a = 40
    pat = np.tile(np.array([128 + a, 128, 128 - a, 128], np.uint8), nrows * nave * n // 4)
    fb = orc.filterbank_f64(c, pat, nrows, n, nave, 1, n // 2)
    expect = np.zeros_like(fb)
    expect[:, n // 4 - 1] = (a * n / 2.0) ** 2 / nfreq
    assert np.allclose(fb, expect, atol=1e-6)
    # FP32 stage mirror agrees with FP64 to float accuracy
    fb32 = orc.filterbank_f32(c, pat, nrows, n, nave, 1, n // 2)
    assert np.allclose(fb32, expect, rtol=1e-6, atol=1e-3)


def test_filterbank_f32_mirror_close_to_f64(orc):
    rng = np.random.default_rng(11)
    n, nave, nrows = 2048, 4, 3
    p0 = np.clip(np.rint(rng.normal(128, 8, nrows * nave * n)), 0, 255).astype(np.uint8)
    p1 = np.clip(np.rint(rng.normal(128, 8, nrows * nave * n)), 0, 255).astype(np.uint8)
    a = orc.filterbank_f32(p0, p1, nrows, n, nave, 1, n // 2)
    b = orc.filterbank_f64(p0, p1, nrows, n, nave, 1, n // 2)
    assert np.max(np.abs(a - b) / np.maximum(b, 1.0)) < 2e-3


# --------------------------------------------------------- delay table ----
def _exact_delay(dm32, j, fch1, chanbw, tchan):
    """dt/tchan in 60-digit decimal arithmetic from the exact binary inputs."""
    getcontext().prec = 60
    D = Decimal
    f = D(fch1) - D(j) * D(chanbw)
    dt = D("4148.808") * D(float(dm32)) * (1 / (f * f) - 1 / (D(fch1) * D(fch1)))
    return dt / D(tchan)


def test_delay_table_against_high_precision(orc):
    nch, fch1, chanbw, tchan = 1024, 383.96875, 0.0625, 64 * 2048 / 128e6
    dms = (np.arange(256) * 0.1).astype(np.float32)
    off, maxd = orc.delay_table(dms, nch, fch1, chanbw, tchan)
    assert np.all(off[:, 0] == 0) and np.all(off[0] == 0)
    assert np.all(np.diff(off, axis=1) >= 0) and np.all(np.diff(off, axis=0) >= 0)
    rng = np.random.default_rng(0)
    for i, j in zip(rng.integers(1, 256, 300), rng.integers(1, nch, 300)):
        q = _exact_delay(dms[i], int(j), fch1, chanbw, tchan)
        fl = int(q)  # truncation toward zero, q > 0
        if abs(q - fl) > Decimal("1e-9") and abs(q - fl - 1) > Decimal("1e-9"):
            assert off[i, j] == fl, (i, j, q)
    # the window overlap quoted for this shape: 308 rows at DM 25.5 (j = NCHANOUT)
    assert maxd == 308
    assert off[-1, -1] == 307


def test_delay_table_c2_max_delay(orc):
    nch, chanbw, tchan = 4096, 64.0 / 4096, 16 * 8192 / 128e6
    fch1 = 320.0 + nch * chanbw - 0.5 * chanbw
    dms = np.linspace(0, 1000, 2048).astype(np.float32)
    off, maxd = orc.delay_table(dms, nch, fch1, chanbw, tchan)
    assert maxd == 12090
    q = _exact_delay(dms[-1], nch, fch1, chanbw, tchan)
    assert int(q) == maxd


# ------------------------------------------------------------ bandpass ----
def test_bandpass_hand_blocks(orc):
    nch = 3
    rows = 128 + 5  # one full block and a 5-row block
    fb = np.zeros((rows, nch), np.float32)
    # constant: variance 0 empties the strict clip window (LO < x < HI),
    # the count falls back to 1 with a zero sum, nothing is subtracted
    fb[:, 0] = 7.0
    ramp = np.arange(rows, dtype=np.float32)
    fb[:, 1] = ramp                                  # no clipping: minus the block mean
    base = np.where(np.arange(rows) % 2 == 0, 1000.25, 999.75).astype(np.float32)
    base[10] = 1010.0                                # outlier beyond mean + 3*var
    fb[:, 2] = base
    out = orc.bandpass_f32(fb)
    assert np.all(out[:, 0] == 7.0)
    assert np.allclose(out[:128, 1], ramp[:128] - ramp[:128].mean(), atol=1e-4)
    assert np.allclose(out[128:, 1], ramp[128:] - ramp[128:].mean(), atol=1e-4)
    keep = np.ones(128, bool)
    keep[10] = False
    clipped_mean = base[:128][keep].astype(np.float64).mean()
    assert abs((base[0] - out[0, 2]) - clipped_mean) < 1e-3
    assert abs((base[10] - out[10, 2]) - clipped_mean) < 1e-3


def test_bandpass_single_row_block_is_untouched(orc):
    # npts == 1: the variance is 0/0, every clip test fails, mean -> 0
    fb = np.arange(129 * 2, dtype=np.float32).reshape(129, 2)
    out = orc.bandpass_f32(fb)
    assert np.all(out[128] == fb[128])
    out64 = orc.bandpass_f64(fb.astype(np.float64))
    assert np.all(out64[128] == fb[128])


def test_bandpass_f32_vs_f64(orc):
    rng = np.random.default_rng(5)
    fb = rng.normal(256, 40, (300, 64)).astype(np.float32)
    fb[::17, 3] += 5000
    a = orc.bandpass_f32(fb)
    b = orc.bandpass_f64(fb.astype(np.float64))
    assert np.max(np.abs(a - b)) < 1e-3


# -------------------------------------------------------- dedispersion ----
def _dedisp_brute(fb, off, nt):
    ndm, nch = off.shape
    out = np.zeros((ndm, nt))
    for d in range(ndm):
        for t in range(nt):
            out[d, t] = sum(fb[t + off[d, c], c] for c in range(nch)) / nch
    return out


def test_dedisperse_bruteforce_and_known_answers(orc):
    rng = np.random.default_rng(2)
    nch, rows, ndm = 7, 40, 5
    fb = rng.normal(size=(rows, nch)).astype(np.float32)
    off = np.sort(rng.integers(0, 9, (ndm, nch)), axis=1).astype(np.int32)
    off[:, 0] = 0
    nt = rows - off.max()
    got = orc.dedisperse_f32(fb, off, nt)
    assert np.allclose(got, _dedisp_brute(fb.astype(np.float64), off, nt), atol=1e-6)
    got64 = orc.dedisperse_f64(fb.astype(np.float64), off, nt)
    assert np.allclose(got64, _dedisp_brute(fb.astype(np.float64), off, nt), atol=1e-12)
    # channel-major variant gives bit-identical results (same summation order)
    cm = np.ascontiguousarray(fb.T)
    assert np.array_equal(orc.dedisperse_cm_f32(cm, off, nt), got)
    # DM 0 -> channel mean of each row
    z = np.zeros((1, nch), np.int32)
    assert np.allclose(orc.dedisperse_f32(fb, z, rows)[0], fb.mean(axis=1), atol=1e-6)
    # single impulse at (r0, c) lands at t = r0 - off[d, c]
    imp = np.zeros((rows, nch), np.float32)
    r0, c = 30, 4
    imp[r0, c] = nch
    res = orc.dedisperse_f32(imp, off, nt)
    for d in range(ndm):
        exp = np.zeros(nt)
        if 0 <= r0 - off[d, c] < nt:
            exp[r0 - off[d, c]] = 1.0
        assert np.array_equal(res[d], exp.astype(np.float32))


def test_dedisperse_stats_known(orc):
    nch, rows = 4, 10
    fb = np.tile(np.array([1.0, 2.0, 3.0, 4.0], np.float32), (rows, 1))
    off = np.zeros((1, nch), np.int32)
    out, sig = orc.dedisperse_stats_f32(fb, off, rows)
    assert np.allclose(out, 2.5)
    # S1 = 0, S2 = 5 -> sqrt(5/(4*3))
    assert np.allclose(sig, np.sqrt(5.0 / 12.0), rtol=1e-6)


# ---------------------------------------------------------- statistics ----
def test_dm_stats_known_answers(orc):
    nt = 1000
    ramp = np.arange(nt, dtype=np.float32)[None, :]
    frb, ave, sig, peak, pt = orc.dm_stats(ramp)
    sd = np.sqrt(nt * (nt + 1) / 12.0)
    assert abs(ave[0] - (nt - 1) / 2) < 1e-3 and abs(sig[0] - sd) < 1e-2
    assert pt[0] == nt - 1 and abs(peak[0] - ((nt - 1) / 2) / sd) < 1e-4
    assert frb[0] == 0  # max SNR of a ramp is sqrt(3) < 3
    two = np.zeros((1, nt), np.float32)
    two[0, 500] = 100.0
    frb, ave, sig, peak, pt = orc.dm_stats(two)
    m = 100.0 / nt
    s = np.sqrt(((100 - m) ** 2 + (nt - 1) * m * m) / (nt - 1))
    assert frb[0] == 1 and pt[0] == 500 and abs(peak[0] - (100 - m) / s) < 1e-3
    neg = -np.abs(np.random.default_rng(0).normal(size=(1, nt))).astype(np.float32) - 100
    neg[0, :] = -100.0  # constant: sigma 0 -> SNR NaN, peak stays 0
    frb, ave, sig, peak, pt = orc.dm_stats(neg)
    assert peak[0] == 0.0 and pt[0] == -1 and frb[0] == 0
    a64, s64, p64, t64 = orc.dm_stats_f64(two.astype(np.float64))
    assert t64[0] == 500 and abs(p64[0] - (100 - m) / s) < 1e-9


def test_chan_stats_known_answers(orc):
    nch, ncol, nblk = 3, 50, 2
    fb = np.zeros((ncol * nblk, nch), np.float32)
    fb[:, 0] = 5.0
    fb[:, 1] = np.arange(ncol * nblk) % 2          # 0/1 alternating
    fb[:ncol, 2] = 1.0
    ave, sig, out = orc.chan_stats(fb, ncol, nblk)
    assert np.all(ave[:, 0] == 5.0) and np.all(sig[:, 0] == 0.0) and np.all(out == 0.0)
    assert np.allclose(ave[:, 1], 0.5) and np.allclose(sig[:, 1], 0.25 * ncol / (ncol - 1))
