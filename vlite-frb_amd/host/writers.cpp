// Byte formats of the reference outputs: SIGPROC header and PRESTO .inf text.
#include <cstdio>
#include <cstring>

#include "../csrc/common.hpp"

namespace {

// Growable view over the caller's buffer; `ok` drops to false on overflow.
struct Out {
    uint8_t* buf;
    int32_t cap;
    int32_t len = 0;
    bool ok = true;
    void put(const void* p, int32_t n) {
        if (len + n > cap) { ok = false; return; }
        std::memcpy(buf + len, p, n);
        len += n;
    }
    // send_string: int32 length, then the characters (PIPELINE:1177-1183).
    void str(const char* s) {
        int32_t n = (int32_t)std::strlen(s);
        put(&n, 4);
        put(s, n);
    }
    void i32(const char* name, int32_t v) { str(name); put(&v, 4); }
    void f64(const char* name, double v) { str(name); put(&v, 8); }
};

}  // namespace

extern "C" {

// Key order and types of PIPELINE:529-540.
vfrb_status vfrb_sigproc_header(uint8_t* buf, int32_t cap, int32_t* len, double fch1,
                                double foff, int32_t nchans, int32_t nbits, double tstart,
                                double tsamp) {
    VFRB_REQUIRE(buf && len && cap > 0, VFRB_ERR_INVALID, "bad buffer");
    Out o{buf, cap};
    o.str("HEADER_START");
    o.i32("barycentric", 1);
    o.i32("telescope_id", 12);
    o.i32("data_type", 1);
    o.f64("fch1", fch1);
    o.f64("foff", foff);
    o.i32("nchans", nchans);
    o.i32("nbits", nbits);
    o.f64("tstart", tstart);
    o.f64("tsamp", tsamp);
    o.i32("nifs", 1);
    o.str("HEADER_END");
    VFRB_REQUIRE(o.ok, VFRB_ERR_CAPACITY, "SIGPROC header needs more than %d bytes", cap);
    *len = o.len;
    return VFRB_OK;
}

// Text of writeprestoheader (TIMESERIES:1483-1515).
vfrb_status vfrb_presto_inf(char* buf, int32_t cap, const char* base, double mjdstart,
                            int32_t ntimes, double tsamp, float dm, double fchanlo, double bw,
                            int32_t nchan, double chanbw) {
    VFRB_REQUIRE(buf && base && cap > 0, VFRB_ERR_INVALID, "bad buffer");
    int n = snprintf(
        buf, cap,
        " Data file name without suffix          = %s\n"
        " Telescope used                         = VLA\n"
        " Instrument used                        = VLITE-FAST\n"
        " Object being observed                  = Unknown\n"
        " J2000 Right Ascension (hh:mm:ss.ssss)  = 00:00:00.0000\n"
        " J2000 Declination     (dd:mm:ss.ssss)  = 00:00:00.0000\n"
        " Data observed by                       = vdifDADA2timeseries\n"
        " Epoch of observation (MJD)             = %.15lf\n"
        " Barycentered?                          = 1\n"
        " Number of bins in the time series      = %d\n"
        " Width of each time series bin (sec)    = %.9lf\n"
        " Any breaks in the data? (1=yes, 0=no)  = 0\n"
        " Type of observation (EM band)          = Radio\n"
        " Beam diameter (arcsec)                 = 3600\n"
        " Dispersion measure (cm-3 pc)           = %f\n"
        " Central freq of low channel (MHz)      = %f\n"
        " Total bandwidth (MHz)                  = %f\n"
        " Number of channels                     = %d\n"
        " Channel bandwidth (MHz)                = %f\n"
        " Data analyzed by                       = vdifDADA2timeseries\n"
        " Any additional notes:\n    \n\n",
        base, mjdstart, ntimes, tsamp, (double)dm, fchanlo, bw, nchan, chanbw);
    VFRB_REQUIRE(n >= 0 && n < cap, VFRB_ERR_CAPACITY, "PRESTO .inf needs %d bytes", n + 1);
    return VFRB_OK;
}

}  // extern "C"
