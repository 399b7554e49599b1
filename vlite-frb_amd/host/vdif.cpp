// VDIF header codec and the two-polarisation frame sequencer (host only).
//
// The sequencer keeps the reference's observable behaviour (which frames are
// consumed, which payload bytes come out, what is logged, when input ends)
// while owning its state in an object instead of globals.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../csrc/common.hpp"

namespace {

inline uint32_t rd32(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
inline void wr32(uint8_t* p, uint32_t v) {
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
    p[2] = (uint8_t)(v >> 16);
    p[3] = (uint8_t)(v >> 24);
}
inline uint32_t bits(uint32_t w, int lo, int n) { return (w >> lo) & ((n == 32) ? 0xffffffffu : ((1u << n) - 1u)); }

// Days from 1970-01-01 to the given proleptic Gregorian civil date.
int64_t days_from_civil(int64_t y, int64_t m, int64_t d) {
    y -= m <= 2;
    const int64_t era = (y >= 0 ? y : y - 399) / 400;
    const int64_t yoe = y - era * 400;
    const int64_t doy = (153 * (m + (m > 2 ? -3 : 9)) + 2) / 5 + d - 1;
    const int64_t doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
    return era * 146097 + doe - 719468;
}

}  // namespace

extern "C" {

// Bit layout of vdifio.h:53-72 as the reference reads it with fread into a
// little-endian bitfield struct: fields fill each 32-bit word from bit 0.
vfrb_status vfrb_vdif_decode(const uint8_t* raw, vfrb_vdif_header* h) {
    VFRB_REQUIRE(raw && h, VFRB_ERR_INVALID, "NULL argument");
    const uint32_t w0 = rd32(raw), w1 = rd32(raw + 4), w2 = rd32(raw + 8), w3 = rd32(raw + 12);
    const uint32_t w4 = rd32(raw + 16);
    h->seconds = bits(w0, 0, 30);
    h->legacymode = bits(w0, 30, 1);
    h->invalid = bits(w0, 31, 1);
    h->frame = bits(w1, 0, 24);
    h->epoch = bits(w1, 24, 6);
    h->unassigned = bits(w1, 30, 2);
    h->framelength8 = bits(w2, 0, 24);
    h->nchan = bits(w2, 24, 5);
    h->version = bits(w2, 29, 3);
    h->stationid = bits(w3, 0, 16);
    h->threadid = bits(w3, 16, 10);
    h->nbits = bits(w3, 26, 5);
    h->iscomplex = bits(w3, 31, 1);
    h->eversion = bits(w4, 0, 8);
    h->extended1 = bits(w4, 8, 24);
    h->extended2 = rd32(raw + 20);
    h->extended3 = rd32(raw + 24);
    h->extended4 = rd32(raw + 28);
    return VFRB_OK;
}

vfrb_status vfrb_vdif_encode(const vfrb_vdif_header* h, uint8_t* raw) {
    VFRB_REQUIRE(raw && h, VFRB_ERR_INVALID, "NULL argument");
    auto f = [](uint32_t v, int lo, int n) { return (v & ((1u << n) - 1u)) << lo; };
    wr32(raw, f(h->seconds, 0, 30) | f(h->legacymode, 30, 1) | f(h->invalid, 31, 1));
    wr32(raw + 4, f(h->frame, 0, 24) | f(h->epoch, 24, 6) | f(h->unassigned, 30, 2));
    wr32(raw + 8, f(h->framelength8, 0, 24) | f(h->nchan, 24, 5) | f(h->version, 29, 3));
    wr32(raw + 12, f(h->stationid, 0, 16) | f(h->threadid, 16, 10) | f(h->nbits, 26, 5) |
                       f(h->iscomplex, 31, 1));
    wr32(raw + 16, f(h->eversion, 0, 8) | f(h->extended1, 8, 24));
    wr32(raw + 20, h->extended2);
    wr32(raw + 24, h->extended3);
    wr32(raw + 28, h->extended4);
    return VFRB_OK;
}

// VDIF epoch e counts half years from 2000-01-01 00:00 UTC (e even: Jan 1,
// e odd: Jul 1 of year 2000 + e/2).  The frame time is epoch start +
// seconds + frame/framespersec; MJD = days since 1858-11-17.
double vfrb_vdif_dmjd(const uint8_t* raw, int32_t framespersec) {
    if (!raw || framespersec <= 0) return 0.0;
    vfrb_vdif_header h;
    vfrb_vdif_decode(raw, &h);
    const int64_t year = 2000 + h.epoch / 2;
    const int64_t month = (h.epoch % 2) ? 7 : 1;
    const int64_t mjd_epoch = days_from_civil(year, month, 1) + 40587;  // 1970-01-01 = MJD 40587
    const int64_t mjd_int = mjd_epoch + h.seconds / 86400;
    return (double)mjd_int + (double)(h.seconds % 86400) / 86400.0 +
           (double)h.frame / ((double)framespersec * 86400.0);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Frame sequencer
// ---------------------------------------------------------------------------
struct vfrb_seq {
    // input
    FILE* fp = nullptr;
    const uint8_t* mem = nullptr;
    size_t mem_size = 0, mem_pos = 0;
    bool eof = false;
    // log
    vfrb_log_fn log_fn = nullptr;
    void* log_user = nullptr;
    // per-pol frame state
    uint8_t hdr[2][VFRB_HDRLEN];
    uint8_t data[2][VFRB_DATALEN];
    uint8_t held[2][VFRB_DATALEN];  // real frame parked while zeros are emitted
    int32_t expected[2] = {0, 0};
    int32_t expected_sec[2] = {0, 0};
    int32_t pending_zero[2] = {0, 0};
    bool sec_check = false;
    // bundle carry
    int32_t part_fill = 0, part_left = 0;
    int64_t frames_read = 0, frames_filled = 0;
    bool initialised = false;

    void log(const char* fmt, ...) __attribute__((format(printf, 2, 3))) {
        if (!log_fn) return;
        char line[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(line, sizeof(line), fmt, ap);
        va_end(ap);
        log_fn(line, log_user);
    }

    // Reads exactly n bytes; false on short read (end of input).
    bool read_bytes(uint8_t* dst, size_t n) {
        if (fp) {
            size_t got = fread(dst, 1, n, fp);
            if (got != n) { eof = true; return false; }
            return true;
        }
        if (mem_pos + n > mem_size) { mem_pos = mem_size; eof = true; return false; }
        std::memcpy(dst, mem + mem_pos, n);
        mem_pos += n;
        return true;
    }

    // initframePOL0/1 (PIPELINE:829-883): header then payload; 1 = end of input.
    int read_raw(int pol) {
        static const char* names[2] = {"0", "1"};
        if (!read_bytes(hdr[pol], VFRB_HDRLEN)) {
            log("EOF while trying to read header%s\n", names[pol]);
            return 1;
        }
        if (!read_bytes(data[pol], VFRB_DATALEN)) {
            log("EOF while trying to read data%s\n", names[pol]);
            return 1;
        }
        frames_read++;
        return 0;
    }

    int frame_number(int pol) const {
        vfrb_vdif_header h;
        vfrb_vdif_decode(hdr[pol], &h);
        return (int)h.frame;
    }
    int full_second(int pol) const {
        vfrb_vdif_header h;
        vfrb_vdif_decode(hdr[pol], &h);
        return (int)h.seconds;
    }
    int thread_id(int pol) const {
        vfrb_vdif_header h;
        vfrb_vdif_decode(hdr[pol], &h);
        return (int)h.threadid;
    }

    // checkframeskip0/1 (PIPELINE:888-930): a frame number other than the
    // expected one starts a run of zero-filled payloads whose length is the
    // gap (mod FRAMESPERSEC); the real frame is parked until the run ends.
    //
    // Seconds-aware rule of TIMESERIES:893-937: the gap also counts whole
    // seconds, gap = fnum - expected + (sec - expected_sec) * FRAMESPERSEC.
    void check_gap(int pol) {
        const int fnum = frame_number(pol);
        const int fsec = sec_check ? full_second(pol) : 0;
        if (fnum != expected[pol] || (sec_check && fsec != expected_sec[pol])) {
            int gap = fnum - expected[pol];
            if (sec_check) {
                // The reference forms this in int (TIMESERIES:901), which a
                // damaged seconds field overflows; its result there is the
                // 32-bit two's-complement wrap, formed here without overflow.
                const int64_t g = (int64_t)gap + ((int64_t)fsec - expected_sec[pol]) * VFRB_FRAMESPERSEC;
                gap = (int)(int32_t)(uint32_t)(uint64_t)g;
                if (gap < 0) gap = 0;  // frames going backwards: no zero run
            } else if (gap < 0) {
                gap += VFRB_FRAMESPERSEC;
            }
            pending_zero[pol] = gap;
            log("FRAME SKIP %d!!! EXPECTED FRAME NUMBER%d= %d, BUT READ %d! (NSKIP%d= %d)\n", pol, pol,
                expected[pol], fnum, pol, gap);
            if (gap > 0) {
                std::memcpy(held[pol], data[pol], VFRB_DATALEN);
                std::memset(data[pol], 0, VFRB_DATALEN);
                frames_filled++;
            }
        }
        expected[pol] = (fnum + 1) % VFRB_FRAMESPERSEC;
        expected_sec[pol] = fsec + (sec_check && expected[pol] == 0 ? 1 : 0);
    }

    // readoneframePOL0/1 (PIPELINE:934-1025).
    int next_frame(int pol) {
        if (pending_zero[pol]) {
            pending_zero[pol]--;
            if (pending_zero[pol] == 0) {
                std::memcpy(data[pol], held[pol], VFRB_DATALEN);
            } else {
                frames_filled++;
            }
            return 0;
        }
        if (read_raw(pol)) return 1;
        check_gap(pol);
        return 0;
    }

    // readtwoframe (PIPELINE:1031-1040).
    int next_pair() {
        if (next_frame(0)) return 1;
        if (next_frame(1)) return 1;
        return 0;
    }
};

extern "C" {

vfrb_status vfrb_seq_open_file(vfrb_seq** seq, const char* path) {
    VFRB_REQUIRE(seq && path, VFRB_ERR_INVALID, "NULL argument");
    FILE* fp = fopen(path, "rb");
    VFRB_REQUIRE(fp, VFRB_ERR_IO, "Unable to open input VDIF file: %s", path);
    vfrb_seq* s = new vfrb_seq();
    s->fp = fp;
    *seq = s;
    return VFRB_OK;
}

vfrb_status vfrb_seq_open_memory(vfrb_seq** seq, const uint8_t* data, size_t bytes) {
    VFRB_REQUIRE(seq && (data || bytes == 0), VFRB_ERR_INVALID, "NULL argument");
    vfrb_seq* s = new vfrb_seq();
    s->mem = data;
    s->mem_size = bytes;
    *seq = s;
    return VFRB_OK;
}

vfrb_status vfrb_seq_skip_pairs(vfrb_seq* seq, int64_t npairs) {
    VFRB_REQUIRE(seq && npairs >= 0, VFRB_ERR_INVALID, "bad skip arguments");
    for (int64_t i = 0; i < npairs; i++) {
        if (seq->read_raw(0) || seq->read_raw(1)) return VFRB_ERR_EOF;
    }
    return VFRB_OK;
}

vfrb_status vfrb_seq_set_seconds_check(vfrb_seq* seq, int32_t enable) {
    VFRB_REQUIRE(seq, VFRB_ERR_INVALID, "NULL sequencer");
    seq->sec_check = enable != 0;
    return VFRB_OK;
}

vfrb_status vfrb_seq_set_log(vfrb_seq* seq, vfrb_log_fn fn, void* user) {
    VFRB_REQUIRE(seq, VFRB_ERR_INVALID, "NULL sequencer");
    seq->log_fn = fn;
    seq->log_user = user;
    return VFRB_OK;
}

// initframe01 (PIPELINE:1058-1131) + initskip01 (PIPELINE:1044-1052).
vfrb_status vfrb_seq_init(vfrb_seq* s) {
    VFRB_REQUIRE(s, VFRB_ERR_INVALID, "NULL sequencer");
    s->pending_zero[0] = s->pending_zero[1] = 0;
    s->part_fill = s->part_left = 0;
    if (s->read_raw(0)) return VFRB_ERR_EOF;
    if (s->thread_id(0) != 0) {
        s->log("VDIF Threads misaligned, skipping forward 1 frame\n");
        if (s->read_raw(0)) return VFRB_ERR_EOF;
    }
    int id0 = s->thread_id(0);
    if (id0 != 0) {
        s->log("ERROR, initframe01! VDIF Thread ID= %d, != 0. Abort!\n", id0);
        vfrb::set_error("first two frames are not thread 0 (thread %d)", id0);
        return VFRB_ERR_FORMAT;
    }
    if (s->read_raw(1)) return VFRB_ERR_EOF;
    int id1 = s->thread_id(1);
    if (id1 != 1) {
        s->log("VDIF Thread ID= %d, != 1. Abort!\n", id1);
        vfrb::set_error("second frame is thread %d, expected 1", id1);
        return VFRB_ERR_FORMAT;
    }
    int fnum0 = s->frame_number(0), fnum1 = s->frame_number(1);
    if (fnum0 != fnum1 || (s->sec_check && s->full_second(0) != s->full_second(1))) {
        s->log("ERROR! polarization VDIF frame numbers do not agree! 0: %d  !=  1: %d\n", fnum0, fnum1);
        vfrb::set_error("pol frame numbers differ: %d != %d", fnum0, fnum1);
        return VFRB_ERR_FORMAT;
    }
    // Start of the 10 Hz noise-cal cycle: 2560 frames per cycle.  The
    // reference tests fmod((float)fnum0, 2560.f) < 1e-4, i.e. an exact
    // multiple for the 24-bit frame numbers that occur.
    int ncal = 0;
    while (std::fmod((float)fnum0, (float)2560) >= 1e-4f) {
        if (s->read_raw(0)) return VFRB_ERR_EOF;
        if (s->read_raw(1)) return VFRB_ERR_EOF;
        fnum0 = s->frame_number(0);
        ncal++;
    }
    s->log("Skipped %d frames to begin at noise cal cycle\n", ncal);
    fnum1 = s->frame_number(1);
    if (fnum0 != fnum1 || (s->sec_check && s->full_second(0) != s->full_second(1))) {
        s->log("ERROR! polarization VDIF frame numbers do not agree! 0: %d  !=  1: %d\n", fnum0, fnum1);
        vfrb::set_error("pol frame numbers differ after cal alignment: %d != %d", fnum0, fnum1);
        return VFRB_ERR_FORMAT;
    }
    s->expected[0] = (fnum0 + 1) % VFRB_FRAMESPERSEC;
    s->expected[1] = (fnum1 + 1) % VFRB_FRAMESPERSEC;
    for (int pol = 0; pol < 2; pol++)
        s->expected_sec[pol] = s->full_second(pol) + (s->expected[pol] == 0 ? 1 : 0);
    s->initialised = true;
    return VFRB_OK;
}

vfrb_status vfrb_seq_header(const vfrb_seq* s, uint8_t* raw32) {
    VFRB_REQUIRE(s && raw32, VFRB_ERR_INVALID, "NULL argument");
    std::memcpy(raw32, s->hdr[0], VFRB_HDRLEN);
    return VFRB_OK;
}

// filltimebufferHOST (PIPELINE:1135-1171).
vfrb_status vfrb_seq_fill(vfrb_seq* s, int32_t nloop, uint8_t* out0, uint8_t* out1) {
    VFRB_REQUIRE(s && out0 && out1 && nloop > 0, VFRB_ERR_INVALID, "bad fill arguments");
    VFRB_REQUIRE(s->initialised, VFRB_ERR_INVALID, "vfrb_seq_init not called");
    int64_t ifill = 0;
    if (s->part_left) {
        const int32_t n = std::min(s->part_left, nloop);
        std::memcpy(out0, s->data[0] + s->part_fill, n);
        std::memcpy(out1, s->data[1] + s->part_fill, n);
        ifill += n;
        if (s->part_left > nloop) {
            // Bundle shorter than the rest of the frame: keep the carry.
            s->part_fill += nloop;
            s->part_left -= nloop;
            return VFRB_OK;
        }
        if (s->next_pair()) return VFRB_ERR_EOF;
    }
    const int32_t ntot = nloop - (int32_t)ifill;
    const int32_t nwhole = ntot / VFRB_DATALEN;
    s->part_fill = ntot % VFRB_DATALEN;
    s->part_left = VFRB_DATALEN - s->part_fill;
    for (int32_t q = 0; q < nwhole; q++) {
        std::memcpy(out0 + ifill, s->data[0], VFRB_DATALEN);
        std::memcpy(out1 + ifill, s->data[1], VFRB_DATALEN);
        ifill += VFRB_DATALEN;
        if (s->next_pair()) return VFRB_ERR_EOF;
    }
    if (s->part_fill) {
        std::memcpy(out0 + ifill, s->data[0], s->part_fill);
        std::memcpy(out1 + ifill, s->data[1], s->part_fill);
    }
    return VFRB_OK;
}

vfrb_status vfrb_seq_consumed(const vfrb_seq* s, int64_t* bytes) {
    VFRB_REQUIRE(s && bytes, VFRB_ERR_INVALID, "NULL argument");
    *bytes = s->frames_read * (int64_t)VFRB_FRAMELEN;
    return VFRB_OK;
}

vfrb_status vfrb_seq_counters(const vfrb_seq* s, int64_t* frames_read, int64_t* frames_filled) {
    VFRB_REQUIRE(s, VFRB_ERR_INVALID, "NULL sequencer");
    if (frames_read) *frames_read = s->frames_read;
    if (frames_filled) *frames_filled = s->frames_filled;
    return VFRB_OK;
}

vfrb_status vfrb_seq_close(vfrb_seq* s) {
    if (!s) return VFRB_OK;
    if (s->fp) fclose(s->fp);
    delete s;
    return VFRB_OK;
}

}  // extern "C"
