/*
 * vfrb_oracle.c - CPU restatement of the VLITE FRB search chain.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or as the timed CPU baseline); the product path never calls it.
 *
 * Two flavours of every stage:
 *   *_f32 : stage mirror of the reference arithmetic in float, in the
 *           reference's own evaluation order wherever the order is defined
 *           (sequential sums, float divisions, double-promoted constants).
 *   *_f64 : the same algorithm in double (the FP64 oracle).
 * The FFT has no reference source (cuFFT is closed); it is restated from the
 * DFT definition X[k] = sum_n x[n] exp(-2 pi i n k / N), unnormalised.
 *
 * Reference files (paulray/vlite-frb): PIPELINE = vdifFILE2dedisp_PIPELINE.cu,
 * TIMESERIES = vdifDADA2timeseries.cu, STATS = vdifDADA2stats.cu.
 *
 * Pinning (tests/test_oracle.py): FFT against a brute-force DFT and closed
 * forms (DC, tones, Parseval); filterbank against closed-form spectra of
 * constant and single-tone inputs; delay table against hand-evaluated
 * entries and invariants; bandpass against hand-computed blocks; dedispersion
 * against brute force and DM 0; statistics against ramp / two-level series.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex zd;
typedef float complex zf;

/* ------------------------------------------------------------------------ */
/* Parameters: PIPELINE:289-309                                             */
/* ------------------------------------------------------------------------ */
/* out_i: twonchan nchan nfreq nave nbundle nloop nchanout nskip maxnt nhalfcyc
 * out_d: chanbw foff fch1 tchan tbundle */
int or_params(int twonchan, int nave, int nbundle, double maxtime, int fullband,
              int nchanout_override, int32_t* oi, double* od) {
    int nchan = (int)(0.5 * twonchan);
    int nfreq = nchan + 1;
    double chanbw = 64.0 / nchan;
    int nchanout, nskip;
    if (nchanout_override > 0) {
        nchanout = nchanout_override;
        nskip = (nchan - nchanout) + 1;
    } else if (fullband) {
        nchanout = nchan;
        nskip = 1;
    } else {
        nchanout = (int)((360.0 - 320.0) / chanbw) + 1;
        if ((nchanout % 2) == 1) nchanout++;
        nskip = (nchan - nchanout) + 1;
    }
    int nloop = nbundle * nave * twonchan;
    double foff = -1.0 * chanbw;
    double fch1 = (320.0 + (nchanout * chanbw)) - (0.5 * chanbw);
    double tchan = (double)(nave * twonchan) / 128e6;
    double tbundle = (double)(nave * twonchan * nbundle) / 128e6;
    float mt = (float)maxtime;
    int maxnt = (int)(mt / tchan);
    int nhalf = maxnt / 128;
    if (maxnt % 128 != 0) nhalf++;
    int32_t vi[10] = {twonchan, nchan, nfreq, nave, nbundle, nloop, nchanout, nskip, maxnt, nhalf};
    double vd[5] = {chanbw, foff, fch1, tchan, tbundle};
    memcpy(oi, vi, sizeof(vi));
    memcpy(od, vd, sizeof(vd));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Delay table: TIMESERIES:449-464                                          */
/* ------------------------------------------------------------------------ */
int or_delay_table(const float* dm, int ndm, int nchanout, double fch1, double chanbw,
                   double tchan, int32_t* off, int32_t* max_delay) {
    int i, j;
    double dt;
    for (i = 0; i < ndm; i++) {
        off[(int64_t)i * nchanout] = 0; /* highest freq chan is the reference */
        for (j = 1; j < nchanout; j++) {
            dt = 4.148808e3 * dm[i] * (pow(fch1 - (j * chanbw), -2) - pow(fch1, -2));
            off[(int64_t)i * nchanout + j] = (int32_t)(dt / tchan);
        }
    }
    *max_delay = (int32_t)((4.148808e3 * dm[ndm - 1] *
                            (pow(fch1 - ((nchanout)*chanbw), -2) - pow(fch1, -2))) / tchan);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FFT (double): mixed-radix recursive Cooley-Tukey over a twiddle table    */
/* ------------------------------------------------------------------------ */
static int smallest_factor(int n) {
    for (int p = 2; p * p <= n; p++)
        if (n % p == 0) return p;
    return n;
}

/* y[k] = sum_j x[j*xs] * w^(j*k*ts), w = tw[1] of an N0-point table. */
static void fft_rec(int n, const zd* x, int xs, zd* y, const zd* tw, int N0, int ts) {
    if (n == 1) {
        y[0] = x[0];
        return;
    }
    int p = smallest_factor(n);
    int m = n / p;
    if (m == 1) { /* prime length: direct DFT */
        for (int k = 0; k < n; k++) {
            zd s = 0;
            for (int j = 0; j < n; j++) s += x[j * xs] * tw[((int64_t)j * k % n) * ts % N0];
            y[k] = s;
        }
        return;
    }
    zd* sub = (zd*)malloc(sizeof(zd) * n);
    for (int r = 0; r < p; r++) fft_rec(m, x + r * xs, xs * p, sub + r * m, tw, N0, ts * p);
    for (int k = 0; k < m; k++) {
        for (int q = 0; q < p; q++) {
            const int kk = k + m * q;
            zd s = 0;
            for (int r = 0; r < p; r++) s += sub[r * m + k] * tw[((int64_t)r * kk % n) * ts % N0];
            y[kk] = s;
        }
    }
    free(sub);
}

static zd* twiddles_f64(int n, int sign) {
    zd* tw = (zd*)malloc(sizeof(zd) * n);
    for (int j = 0; j < n; j++) tw[j] = cexp(sign * 2.0 * M_PI * I * (double)j / (double)n);
    return tw;
}

/* Forward unnormalised complex FFT of length n (any n), double. */
void or_fft_f64(int n, const double* re, const double* im, double* ore, double* oim) {
    zd* x = (zd*)malloc(sizeof(zd) * n);
    zd* y = (zd*)malloc(sizeof(zd) * n);
    zd* tw = twiddles_f64(n, -1);
    for (int i = 0; i < n; i++) x[i] = re[i] + I * (im ? im[i] : 0.0);
    fft_rec(n, x, 1, y, tw, n, 1);
    for (int i = 0; i < n; i++) {
        ore[i] = creal(y[i]);
        oim[i] = cimag(y[i]);
    }
    free(x);
    free(y);
    free(tw);
}

/* Brute-force DFT straight from the definition (for pinning the FFT). */
void or_dft_f64(int n, const double* re, const double* im, double* ore, double* oim) {
    for (int k = 0; k < n; k++) {
        long double sr = 0, si = 0;
        for (int j = 0; j < n; j++) {
            const long double a = -2.0L * 3.141592653589793238462643383279503L * (long double)((int64_t)j * k % n) / n;
            const long double c = cosl(a), s = sinl(a);
            const long double xr = re[j], xi = im ? im[j] : 0.0;
            sr += xr * c - xi * s;
            si += xr * s + xi * c;
        }
        ore[k] = (double)sr;
        oim[k] = (double)si;
    }
}

/* ------------------------------------------------------------------------ */
/* FFT (float, power of two): iterative radix-2 for the FP32 stage mirror   */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n;
    zf* tw; /* exp(-2 pi i k/n), k < n/2, rounded from double */
    int* rev;
} fplan;

static fplan* fplan_new(int n) {
    fplan* p = (fplan*)malloc(sizeof(fplan));
    p->n = n;
    p->tw = (zf*)malloc(sizeof(zf) * (n / 2 + 1));
    p->rev = (int*)malloc(sizeof(int) * n);
    for (int k = 0; k < n / 2 + 1; k++) {
        const double a = -2.0 * M_PI * k / n;
        p->tw[k] = (float)cos(a) + I * (float)sin(a);
    }
    int lg = 0;
    while ((1 << lg) < n) lg++;
    for (int i = 0; i < n; i++) {
        int r = 0;
        for (int b = 0; b < lg; b++)
            if (i & (1 << b)) r |= 1 << (lg - 1 - b);
        p->rev[i] = r;
    }
    return p;
}
static void fplan_free(fplan* p) {
    free(p->tw);
    free(p->rev);
    free(p);
}
static void fft_f32_inplace(const fplan* p, zf* a) {
    const int n = p->n;
    for (int i = 0; i < n; i++)
        if (i < p->rev[i]) {
            zf t = a[i];
            a[i] = a[p->rev[i]];
            a[p->rev[i]] = t;
        }
    for (int len = 2; len <= n; len <<= 1) {
        const int step = n / len, h = len / 2;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < h; k++) {
                const zf w = p->tw[k * step];
                const zf u = a[i + k], v = a[i + k + h] * w;
                a[i + k] = u + v;
                a[i + k + h] = u - v;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Filterbank: convertarray, R2C, sumpower, avepower, channel select        */
/* (PIPELINE:582-632, 1292-1322).  Each pol is transformed on its own, from */
/* the raw byte value (no offset), exactly as the reference feeds cuFFT.    */
/* ------------------------------------------------------------------------ */
void or_filterbank_f64(const uint8_t* p0, const uint8_t* p1, int64_t nrows, int n, int nave,
                       int nskip, int nchanout, double* out) {
    const int nfreq = n / 2 + 1;
    zd* tw = twiddles_f64(n, -1);
#pragma omp parallel
    {
        zd* x = (zd*)malloc(sizeof(zd) * n);
        zd* X0 = (zd*)malloc(sizeof(zd) * n);
        zd* X1 = (zd*)malloc(sizeof(zd) * n);
        double* acc = (double*)malloc(sizeof(double) * nchanout);
#pragma omp for schedule(dynamic, 1)
        for (int64_t r = 0; r < nrows; r++) {
            for (int c = 0; c < nchanout; c++) acc[c] = 0.0;
            for (int a = 0; a < nave; a++) {
                const int64_t off = (r * nave + a) * (int64_t)n;
                for (int i = 0; i < n; i++) x[i] = (double)p0[off + i];
                fft_rec(n, x, 1, X0, tw, n, 1);
                for (int i = 0; i < n; i++) x[i] = (double)p1[off + i];
                fft_rec(n, x, 1, X1, tw, n, 1);
                for (int c = 0; c < nchanout; c++) {
                    const int k = nskip + c;
                    const double pw = creal(X0[k]) * creal(X0[k]) + cimag(X0[k]) * cimag(X0[k]) +
                                      creal(X1[k]) * creal(X1[k]) + cimag(X1[k]) * cimag(X1[k]);
                    acc[c] += pw / (double)nfreq;
                }
            }
            for (int c = 0; c < nchanout; c++) out[r * nchanout + c] = acc[c] / nave;
        }
        free(x);
        free(X0);
        free(X1);
        free(acc);
    }
    free(tw);
}

/* FP32 stage mirror (power-of-two n): float R2C via an n-point complex FFT
 * of the real input, sumpower's (|X0|^2+|X1|^2)/NFREQ in float, avepower's
 * sequential float sum then /NAVE. */
void or_filterbank_f32(const uint8_t* p0, const uint8_t* p1, int64_t nrows, int n, int nave,
                       int nskip, int nchanout, float* out) {
    fplan* pl = fplan_new(n);
    const int nfreq = n / 2 + 1;
#pragma omp parallel
    {
        zf* a0 = (zf*)malloc(sizeof(zf) * n);
        zf* a1 = (zf*)malloc(sizeof(zf) * n);
        float* acc = (float*)malloc(sizeof(float) * nchanout);
#pragma omp for schedule(dynamic, 1)
        for (int64_t r = 0; r < nrows; r++) {
            for (int c = 0; c < nchanout; c++) acc[c] = 0.0f;
            for (int a = 0; a < nave; a++) {
                const int64_t off = (r * nave + a) * (int64_t)n;
                for (int i = 0; i < n; i++) {
                    a0[i] = (float)p0[off + i];
                    a1[i] = (float)p1[off + i];
                }
                fft_f32_inplace(pl, a0);
                fft_f32_inplace(pl, a1);
                for (int c = 0; c < nchanout; c++) {
                    const int k = nskip + c;
                    const float pw = ((crealf(a0[k]) * crealf(a0[k])) + (cimagf(a0[k]) * cimagf(a0[k])) +
                                      (crealf(a1[k]) * crealf(a1[k])) + (cimagf(a1[k]) * cimagf(a1[k]))) /
                                     (float)nfreq;
                    acc[c] += pw;
                }
            }
            for (int c = 0; c < nchanout; c++) out[r * nchanout + c] = acc[c] / nave;
        }
        free(a0);
        free(a1);
        free(acc);
    }
    fplan_free(pl);
}

/* ------------------------------------------------------------------------ */
/* Bandpass: subtractbpCAL (PIPELINE:1427-1474)                             */
/* ------------------------------------------------------------------------ */
/* margin (optional, [nblocks][nch]): smallest |x - HI| or |x - LO| over the
 * block, relative to max(|HI|,|LO|,1); a value within rounding of a clip
 * threshold can legitimately land on either side in another summation order. */
void or_bandpass_f32(float* fb, int nrows, int nch, float* margin) {
    const int nhalf = (nrows + 127) / 128;
#pragma omp parallel for schedule(static)
    for (int row = 0; row < nch; row++) {
        for (int b = 0; b < nhalf; b++) {
            int i, j, istart, colmin, colmax, npts;
            float sum, sum2, ave, sig, HI, LO;
            colmin = 128 * b;
            colmax = colmin + 128;
            if (colmax > nrows) colmax = nrows;
            istart = row + (colmin * nch);
            npts = colmax - colmin;
            ave = 0.0f;
            for (j = 0; j < npts; j++) ave += fb[istart + (int64_t)j * nch];
            ave /= npts;
            sum = 0.0f;
            sum2 = 0.0f;
            for (j = 0; j < npts; j++) {
                sum2 += ((fb[istart + (int64_t)j * nch] - ave) * (fb[istart + (int64_t)j * nch] - ave));
                sum += (fb[istart + (int64_t)j * nch] - ave);
            }
            sig = (sum2 - (sum * sum) / npts) / (npts - 1);
            HI = ave + (3.0 * sig);
            LO = ave - (3.0 * sig);
            ave = 0.0f;
            i = 0;
            float mg = INFINITY;
            const float scale = fmaxf(fmaxf(fabsf(HI), fabsf(LO)), 1.0f);
            for (j = 0; j < npts; j++) {
                const float v = fb[istart + (int64_t)j * nch];
                if (v < HI && v > LO) {
                    ave += v;
                    i++;
                }
                mg = fminf(mg, fminf(fabsf(v - HI), fabsf(v - LO)) / scale);
            }
            if (margin) margin[(int64_t)b * nch + row] = mg;
            if (i == 0) i = 1;
            ave /= i;
            for (j = 0; j < npts; j++) fb[istart + (int64_t)j * nch] -= ave;
        }
    }
}

void or_bandpass_f64(double* fb, int nrows, int nch) {
    const int nhalf = (nrows + 127) / 128;
#pragma omp parallel for schedule(static)
    for (int row = 0; row < nch; row++) {
        for (int b = 0; b < nhalf; b++) {
            const int colmin = 128 * b;
            const int colmax = colmin + 128 > nrows ? nrows : colmin + 128;
            const int npts = colmax - colmin;
            double* x = fb + row + (int64_t)colmin * nch;
            double ave = 0.0, sum = 0.0, sum2 = 0.0;
            for (int j = 0; j < npts; j++) ave += x[(int64_t)j * nch];
            ave /= npts;
            for (int j = 0; j < npts; j++) {
                const double d = x[(int64_t)j * nch] - ave;
                sum2 += d * d;
                sum += d;
            }
            const double sig = (sum2 - (sum * sum) / npts) / (npts - 1);
            const double HI = ave + 3.0 * sig, LO = ave - 3.0 * sig;
            double cs = 0.0;
            int cnt = 0;
            for (int j = 0; j < npts; j++) {
                const double v = x[(int64_t)j * nch];
                if (v < HI && v > LO) {
                    cs += v;
                    cnt++;
                }
            }
            if (cnt == 0) cnt = 1;
            cs /= cnt;
            for (int j = 0; j < npts; j++) x[(int64_t)j * nch] -= cs;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Dedispersion: dedisperse (PIPELINE:1517-1537)                            */
/* out[dm][t] = (sum_c FB[t+off[dm][c]][c]) / nch; the channel loop is kept */
/* innermost per (dm,t) in spirit: the running sum of each output element   */
/* receives the channels in order 0..nch-1, as the reference's global RMW.  */
/* ------------------------------------------------------------------------ */
void or_dedisperse_f32(const float* fb, int nrows, int nch, const int32_t* off, int ndm, int nt,
                       float* out) {
    (void)nrows;
#pragma omp parallel for schedule(dynamic, 1)
    for (int dm = 0; dm < ndm; dm++) {
        float* o = out + (int64_t)dm * nt;
        for (int t = 0; t < nt; t++) o[t] = 0.0f;
        for (int c = 0; c < nch; c++) {
            const int d = off[(int64_t)dm * nch + c];
            const float* col = fb + (int64_t)d * nch + c;
            for (int t = 0; t < nt; t++) o[t] += col[(int64_t)t * nch];
        }
        for (int t = 0; t < nt; t++) o[t] /= nch;
    }
}

/* Same with a channel-major input [nch][ld] (fast CPU baseline layout). */
void or_dedisperse_cm_f32(const float* fbcm, int64_t ld, int nch, const int32_t* off, int ndm,
                          int nt, float* out) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int dm = 0; dm < ndm; dm++) {
        float* o = out + (int64_t)dm * nt;
        for (int t = 0; t < nt; t++) o[t] = 0.0f;
        for (int c = 0; c < nch; c++) {
            const float* col = fbcm + (int64_t)c * ld + off[(int64_t)dm * nch + c];
            for (int t = 0; t < nt; t++) o[t] += col[t];
        }
        for (int t = 0; t < nt; t++) o[t] /= nch;
    }
}

void or_dedisperse_f64(const double* fb, int nrows, int nch, const int32_t* off, int ndm, int nt,
                       double* out) {
    (void)nrows;
#pragma omp parallel for schedule(dynamic, 1)
    for (int dm = 0; dm < ndm; dm++) {
        double* o = out + (int64_t)dm * nt;
        for (int t = 0; t < nt; t++) o[t] = 0.0;
        for (int c = 0; c < nch; c++) {
            const double* col = fb + (int64_t)off[(int64_t)dm * nch + c] * nch + c;
            for (int t = 0; t < nt; t++) o[t] += col[(int64_t)t * nch];
        }
        for (int t = 0; t < nt; t++) o[t] /= nch;
    }
}

/* dedisperseSTATS (TIMESERIES:1450-1481). */
void or_dedisperse_stats_f32(const float* fb, int nch, const int32_t* off, int ndm, int nt,
                             float* out, float* sig) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int dm = 0; dm < ndm; dm++) {
        for (int t = 0; t < nt; t++) {
            float sum = 0.0f, sum2, ave, sigma;
            int i;
            for (i = 0; i < nch; i++) sum += fb[(int64_t)(off[(int64_t)dm * nch + i] + t) * nch + i];
            ave = sum / nch;
            sum = sum2 = 0.0f;
            for (i = 0; i < nch; i++) {
                const float v = fb[(int64_t)(off[(int64_t)dm * nch + i] + t) * nch + i];
                sum += (v - ave);
                sum2 += ((v - ave) * (v - ave));
            }
            sigma = (sum2 - (sum * sum)) / (nch * (nch - 1));
            out[(int64_t)dm * nt + t] = ave;
            sig[(int64_t)dm * nt + t] = sqrtf(sigma);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Per-DM statistics: calcstats (PIPELINE:1337-1370)                        */
/* ------------------------------------------------------------------------ */
void or_dm_stats(const float* out, int64_t ld, int ndm, int nt, int32_t* frb, float* ave,
                 float* sigma, float* peak, int32_t* peak_t) {
#pragma omp parallel for schedule(static)
    for (int dm = 0; dm < ndm; dm++) {
        const float* x = out + (int64_t)dm * ld;
        double tave, tsigma, sum, sum2, dout;
        int i;
        frb[dm] = 0;
        peak[dm] = 0.0f;
        peak_t[dm] = -1;
        sum = 0.0;
        for (i = 0; i < nt; i++) sum += x[i];
        tave = sum / nt;
        sum = sum2 = 0.0;
        for (i = 0; i < nt; i++) {
            dout = x[i];
            sum2 += ((dout - tave) * (dout - tave));
            sum += (dout - tave);
        }
        tsigma = (sum2 - (sum * sum) / nt) / (nt - 1);
        ave[dm] = (float)tave;
        sigma[dm] = (float)sqrt(tsigma);
        for (i = 0; i < nt; i++) {
            dout = (x[i] - ave[dm]) / sigma[dm];
            if (dout > 3.0) frb[dm] = 1;
            if (dout > peak[dm]) {
                peak[dm] = (float)dout;
                peak_t[dm] = i;
            }
        }
    }
}

/* FP64 oracle statistics of a double series: SNR in double throughout. */
void or_dm_stats_f64(const double* out, int64_t ld, int ndm, int nt, double* ave, double* sigma,
                     double* peak, int32_t* peak_t) {
#pragma omp parallel for schedule(static)
    for (int dm = 0; dm < ndm; dm++) {
        const double* x = out + (int64_t)dm * ld;
        double s = 0.0, s1 = 0.0, s2 = 0.0;
        for (int i = 0; i < nt; i++) s += x[i];
        const double m = s / nt;
        for (int i = 0; i < nt; i++) {
            s2 += (x[i] - m) * (x[i] - m);
            s1 += x[i] - m;
        }
        const double sd = sqrt((s2 - s1 * s1 / nt) / (nt - 1));
        double pk = 0.0;
        int pt = -1;
        for (int i = 0; i < nt; i++) {
            const double snr = (x[i] - m) / sd;
            if (snr > pk) {
                pk = snr;
                pt = i;
            }
        }
        ave[dm] = m;
        sigma[dm] = sd;
        peak[dm] = pk;
        peak_t[dm] = pt;
    }
}

/* ------------------------------------------------------------------------ */
/* Per-channel statistics: STATS calcstats (STATS:910-946)                  */
/* ------------------------------------------------------------------------ */
void or_chan_stats(const float* fb, int nrows, int nch, int ncol, int nblk, float* cave,
                   float* csig, float* cout) {
    for (int b = 0; b < nblk; b++) {
        for (int row = 0; row < nch; row++) {
            int i, j, colmin, colmax, npts, istart;
            float sum, sum2, ave, sig, HI;
            colmin = ncol * b;
            colmax = colmin + ncol;
            if (colmax > nrows) colmax = nrows;
            istart = row + (colmin * nch);
            npts = colmax - colmin;
            ave = 0.0f;
            for (j = 0; j < npts; j++) ave += fb[istart + (int64_t)j * nch];
            ave /= npts;
            sum = 0.0f;
            sum2 = 0.0f;
            for (j = 0; j < npts; j++) {
                sum2 += ((fb[istart + (int64_t)j * nch] - ave) * (fb[istart + (int64_t)j * nch] - ave));
                sum += (fb[istart + (int64_t)j * nch] - ave);
            }
            sig = (sum2 - (sum * sum) / npts) / (npts - 1);
            HI = ave + (5.0 * sig);
            i = 0;
            for (j = 0; j < npts; j++)
                if (fb[istart + (int64_t)j * nch] > HI) i++;
            cave[row + (b * nch)] = ave;
            csig[row + (b * nch)] = sig;
            cout[row + (b * nch)] = (i / npts);
        }
    }
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
