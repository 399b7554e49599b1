// Parameter derivation and the FP64 dispersion delay table (host only).
#include <cmath>
#include <cstring>

#include "../csrc/common.hpp"

extern "C" {

// Formulae of PIPELINE:289-309 (TIMESERIES:276-305, STATS:209-232).  The
// reference keeps MAXTIME in a float, so the window length is rounded to
// float before the division, exactly as `(int)(MAXTIME/tchan)` evaluates.
vfrb_status vfrb_params_derive(int32_t twonchan, int32_t nave, int32_t nbundle,
                               double maxtime, int32_t fullband,
                               int32_t nchanout_override, vfrb_params* p) {
    VFRB_REQUIRE(p, VFRB_ERR_INVALID, "params is NULL");
    VFRB_REQUIRE(twonchan >= 4 && twonchan % 2 == 0, VFRB_ERR_INVALID,
                 "twonchan=%d must be even and >= 4", twonchan);
    VFRB_REQUIRE(nave >= 1 && nbundle >= 1, VFRB_ERR_INVALID, "nave=%d nbundle=%d", nave, nbundle);
    std::memset(p, 0, sizeof(*p));
    p->twonchan = twonchan;
    p->nchan = (int32_t)(0.5 * twonchan);
    p->nfreq = p->nchan + 1;
    p->nave = nave;
    p->nbundle = nbundle;
    p->chanbw = VFRB_BW_MHZ / p->nchan;
    if (nchanout_override > 0) {
        p->nchanout = nchanout_override;
        p->nskip = (p->nchan - p->nchanout) + 1;
    } else if (fullband) {
        // Full 320-384 MHz band: bins 1..NCHAN (the DC bin is dropped as
        // in the reference's convention, PIPELINE:295).
        p->nchanout = p->nchan;
        p->nskip = 1;
    } else {
        p->nchanout = (int32_t)((VFRB_MUOSLOW_MHZ - VFRB_LOWFREQ_MHZ) / p->chanbw) + 1;
        if (p->nchanout % 2 == 1) p->nchanout++;
        p->nskip = (p->nchan - p->nchanout) + 1;
    }
    VFRB_REQUIRE(p->nskip >= 1 && p->nskip + p->nchanout - 1 <= p->nchan, VFRB_ERR_INVALID,
                 "channel selection nskip=%d nchanout=%d outside bins 1..%d", p->nskip,
                 p->nchanout, p->nchan);
    const int64_t nloop = (int64_t)nbundle * nave * twonchan;
    VFRB_REQUIRE(nloop <= INT32_MAX, VFRB_ERR_INVALID, "NLOOP overflows int32");
    p->nloop = (int32_t)nloop;
    p->foff = -1.0 * p->chanbw;
    p->fch1 = (VFRB_LOWFREQ_MHZ + (p->nchanout * p->chanbw)) - (0.5 * p->chanbw);
    p->tchan = (double)(nave * twonchan) / VFRB_SAMPLERATE;
    p->tbundle = (double)((int64_t)nave * twonchan * nbundle) / VFRB_SAMPLERATE;
    const float maxtime_f = (float)maxtime;
    p->maxtime = maxtime_f;
    p->maxnt = (int32_t)(maxtime_f / p->tchan);
    p->nhalfcyc = p->maxnt / 128;
    if (p->maxnt % 128 != 0) p->nhalfcyc++;
    return VFRB_OK;
}

// Delay table of TIMESERIES:449-464.  The expression is evaluated in the
// same order and precision as the reference: the DM is a float promoted to
// double, the channel frequency is fch1 - j*chanbw with j an int, and the
// quotient dt/tchan is truncated toward zero.
vfrb_status vfrb_delay_table(const float* dms, int32_t ndm, int32_t nchanout,
                             double fch1, double chanbw, double tchan,
                             int32_t* offsets, int32_t* max_delay, double* min_margin) {
    VFRB_REQUIRE(dms && ndm > 0 && nchanout > 0, VFRB_ERR_INVALID, "empty DM list or channel set");
    VFRB_REQUIRE(tchan > 0.0 && fch1 > 0.0, VFRB_ERR_INVALID, "tchan/fch1 must be positive");
    VFRB_REQUIRE(fch1 - nchanout * chanbw > 0.0, VFRB_ERR_INVALID, "band extends below 0 MHz");
    double margin = 1.0;
    const double inv_f1 = pow(fch1, -2);
    for (int32_t i = 0; i < ndm; i++) {
        VFRB_REQUIRE(std::isfinite(dms[i]) && dms[i] >= 0.0f, VFRB_ERR_INVALID,
                     "DM[%d]=%g must be finite and >= 0", i, (double)dms[i]);
        if (offsets) offsets[(int64_t)i * nchanout] = 0;
        for (int32_t j = 1; j < nchanout; j++) {
            const double dt = 4.148808e3 * dms[i] * (pow(fch1 - (j * chanbw), -2) - inv_f1);
            const double q = dt / tchan;
            VFRB_REQUIRE(q < 2.0e9, VFRB_ERR_INVALID, "delay overflows int32");
            if (offsets) offsets[(int64_t)i * nchanout + j] = (int32_t)q;
            const double frac = q - std::floor(q);
            const double m = frac < 1.0 - frac ? frac : 1.0 - frac;
            if (dms[i] > 0.0f && m < margin) margin = m;
        }
    }
    if (max_delay) {
        const double dt = 4.148808e3 * dms[ndm - 1] *
                          (pow(fch1 - ((nchanout)*chanbw), -2) - inv_f1);
        *max_delay = (int32_t)(dt / tchan);
    }
    if (min_margin) *min_margin = margin;
    return VFRB_OK;
}

}  // extern "C"
