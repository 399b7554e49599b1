// libvfrb's host code under AddressSanitizer and UndefinedBehaviorSanitizer:
// parameter derivation, the delay table, the VDIF header codec, the frame
// sequencer on valid, edge-case and randomly damaged frame streams (both gap
// rules), and the header writers with short buffers.  Every call must return
// a status without a sanitizer report; a few outcomes the reference fixes
// (start on the cal boundary, zero-filled gaps, the dropped last bundle) are
// asserted as well.  Exit status 0 on success.
//
// Built by tests/test_host_sanitize.py with g++ from vlite-frb_amd/host/*.cpp
// (no device code); this file supplies the library's error recorder, which
// otherwise lives with the HIP runtime wrappers.
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vfrb.h"

namespace vfrb {
void set_error(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
}
}  // namespace vfrb

namespace {

int failures = 0;

void check(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "CHECK FAILED: %s\n", what);
        failures++;
    }
}

struct Rng {
    uint64_t s;
    uint32_t next() {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return (uint32_t)(s >> 33);
    }
    uint32_t below(uint32_t n) { return next() % n; }
};

constexpr int kFrame = 5032, kPayload = 5000, kPerSec = 25600;

// One VDIF frame: header fields as the instrument writes them, payload bytes
// derived from (pol, pair) so that the sequencer's output can be traced.
void put_frame(std::vector<uint8_t>& v, uint32_t sec, uint32_t frame, uint32_t thread, uint8_t fill) {
    vfrb_vdif_header h{};
    h.seconds = sec;
    h.frame = frame;
    h.epoch = 30;
    h.framelength8 = kFrame / 8;
    h.threadid = thread;
    h.nbits = 7;
    const size_t o = v.size();
    v.resize(o + kFrame, fill);
    vfrb_vdif_encode(&h, v.data() + o);
}

// Pairs k0 .. k0+n-1 (frame index k = second * 25600 + frame), thread 0 then 1.
void put_pairs(std::vector<uint8_t>& v, int64_t k0, int64_t n) {
    for (int64_t k = k0; k < k0 + n; k++) {
        const uint32_t sec = 1000 + (uint32_t)(k / kPerSec), fr = (uint32_t)(k % kPerSec);
        put_frame(v, sec, fr, 0, (uint8_t)(1 + k % 200));
        put_frame(v, sec, fr, 1, (uint8_t)(2 + k % 200));
    }
}

struct RunResult {
    vfrb_status init = VFRB_OK;
    int bundles = 0;
    int64_t read = 0, filled = 0;
    std::vector<uint8_t> first0;  // first bundle of pol 0
};

RunResult run(const std::vector<uint8_t>& data, int seconds_rule, int nloop, int max_bundles = 64) {
    RunResult r;
    vfrb_seq* seq = nullptr;
    if (vfrb_seq_open_memory(&seq, data.data(), data.size()) != VFRB_OK) {
        r.init = VFRB_ERR_INVALID;
        return r;
    }
    vfrb_seq_set_seconds_check(seq, seconds_rule);
    r.init = vfrb_seq_init(seq);
    if (r.init == VFRB_OK) {
        uint8_t hdr[32];
        vfrb_seq_header(seq, hdr);
        std::vector<uint8_t> o0(nloop), o1(nloop);
        while (r.bundles < max_bundles && vfrb_seq_fill(seq, nloop, o0.data(), o1.data()) == VFRB_OK) {
            if (r.bundles == 0) r.first0 = o0;
            r.bundles++;
        }
    }
    vfrb_seq_counters(seq, &r.read, &r.filled);
    vfrb_seq_close(seq);
    return r;
}

void sequencer_cases() {
    const int nloop = 3 * kPayload + 1234;  // bundles end mid-frame: partial-frame carry
    {
        std::vector<uint8_t> v;
        put_pairs(v, 0, 60);
        const RunResult r = run(v, 0, nloop);
        check(r.init == VFRB_OK && r.bundles == (60 * kPayload) / nloop, "whole stream: every complete bundle");
        check(r.filled == 0, "whole stream: nothing zero-filled");
        check(!r.first0.empty() && r.first0[0] == 1, "whole stream: first payload byte of pair 0");
    }
    {
        std::vector<uint8_t> v;  // off the cal boundary: starts at frame 2560
        put_pairs(v, 2500, 100);
        const RunResult r = run(v, 0, nloop);
        check(r.init == VFRB_OK && !r.first0.empty() && r.first0[0] == (uint8_t)(1 + 2560 % 200),
              "off-cal start skips to frame 2560");
    }
    {
        std::vector<uint8_t> v;  // starts on a thread-1 frame
        put_frame(v, 1000, 0, 1, 9);
        put_pairs(v, 1, 40);
        check(run(v, 0, nloop).init != VFRB_ERR_INVALID, "thread-1 start handled");
    }
    {
        std::vector<uint8_t> v;  // gap of 7 pairs: zero-filled
        put_pairs(v, 0, 20);
        put_pairs(v, 27, 20);
        const RunResult r = run(v, 0, nloop);
        check(r.init == VFRB_OK && r.filled == 2 * 7, "7-pair gap zero-filled in both pols");
    }
    for (int rule = 0; rule < 2; rule++) {
        std::vector<uint8_t> v;  // across a second boundary
        put_pairs(v, kPerSec - 2560, 2600);
        check(run(v, rule, nloop, 4).init == VFRB_OK, "second wrap");
    }
    {
        std::vector<uint8_t> v;  // a whole missing second: frame numbers repeat
        put_pairs(v, 0, 10);
        put_pairs(v, kPerSec + 10, 10);
        const RunResult r0 = run(v, 0, nloop), r1 = run(v, 1, 8 * kPayload, 4096);
        check(r0.filled == 0, "frame-number rule does not see a whole-second gap");
        check(r1.filled == 2 * kPerSec, "seconds-aware rule fills the whole second in both pols");
    }
    {
        // a damaged seconds field 2^29 s ahead under the seconds-aware rule:
        // the gap (in int at TIMESERIES:901) overflows there; here it wraps
        // the same way without undefined behaviour
        std::vector<uint8_t> v;
        put_pairs(v, 0, 5);
        put_frame(v, 1000 + (1u << 29), 5, 0, 3);
        put_frame(v, 1000 + (1u << 29), 5, 1, 3);
        put_pairs(v, 6, 5);
        run(v, 1, nloop, 4);
    }
    {
        std::vector<uint8_t> v;  // pol frame numbers disagree
        put_pairs(v, 0, 5);
        put_frame(v, 1000, 5, 0, 3);
        put_frame(v, 1000, 6, 1, 3);
        put_pairs(v, 7, 5);
        run(v, 0, nloop);
        run(v, 1, nloop);
    }
    {
        std::vector<uint8_t> v;  // truncated last frame: the last bundle is dropped
        put_pairs(v, 0, 12);
        v.resize(v.size() - 1777);
        run(v, 0, nloop);
        run(v, 0, kPayload);
    }
    for (size_t n : {size_t(0), size_t(1), size_t(31), size_t(32), size_t(kFrame - 1), size_t(kFrame),
                     size_t(2 * kFrame + 5)}) {
        std::vector<uint8_t> v;
        put_pairs(v, 0, 3);
        v.resize(n);
        run(v, 0, nloop);
        run(v, 1, nloop);
    }
}

// Random damage to a valid stream: header bytes overwritten, frames cut,
// chunks removed or duplicated, random tails.  Zero-filling is bounded by the
// number of bundles asked for.
void sequencer_fuzz(int iters) {
    Rng rng{12345};
    for (int it = 0; it < iters; it++) {
        std::vector<uint8_t> v;
        put_pairs(v, rng.below(3 * kPerSec), 8 + rng.below(40));
        const int edits = 1 + rng.below(6);
        for (int e = 0; e < edits && !v.empty(); e++) {
            const size_t f = rng.below((uint32_t)(v.size() / kFrame + 1)) * kFrame;
            switch (rng.below(5)) {
                case 0:  // header byte
                    if (f + 32 <= v.size()) v[f + rng.below(32)] = (uint8_t)rng.next();
                    break;
                case 1:  // truncate anywhere
                    v.resize(rng.below((uint32_t)v.size() + 1));
                    break;
                case 2:  // remove some bytes in the middle
                    if (v.size() > 100) {
                        const size_t a = rng.below((uint32_t)v.size() - 50);
                        v.erase(v.begin() + a, v.begin() + a + 1 + rng.below(49));
                    }
                    break;
                case 3:  // duplicate a frame
                    if (f + kFrame <= v.size()) {
                        std::vector<uint8_t> fr(v.begin() + f, v.begin() + f + kFrame);
                        v.insert(v.begin() + f, fr.begin(), fr.end());
                    }
                    break;
                default:  // random tail
                    for (int k = 0; k < 100; k++) v.push_back((uint8_t)rng.next());
            }
        }
        const int nloop = 1 + (int)rng.below(4 * kPayload);
        run(v, (int)(it & 1), nloop, 16);
    }
}

void params_and_tables() {
    vfrb_params p;
    const int32_t tw[] = {0, -2, 2, 3, 2048, 8192, 12500, 12501, 1 << 20};
    const int32_t na[] = {0, -1, 1, 4, 64};
    const int32_t nb[] = {0, 1, 32};
    const double mt[] = {0.0, -1.0, 0.001, 1.0, 70.0, 1e12};
    for (int32_t a : tw)
        for (int32_t b : na)
            for (int32_t c : nb)
                for (double d : mt)
                    for (int fb = 0; fb < 2; fb++)
                        for (int32_t ov : {0, 3840, -5, 1 << 30}) vfrb_params_derive(a, b, c, d, fb, ov, &p);
    check(vfrb_params_derive(12500, 4, 32, 70.0, 0, 0, nullptr) != VFRB_OK, "params: NULL output rejected");

    const float dms[] = {0.0f, 0.1f, 25.5f, 1000.0f, -3.0f};
    std::vector<int32_t> off(5 * 4096);
    int32_t maxd = 0;
    double margin = 0.0;
    check(vfrb_delay_table(dms, 4, 4096, 383.9921875, 0.015625, 0.001024, off.data(), &maxd, &margin) == VFRB_OK,
          "delay table 4096 ch");
    check(off[0] == 0 && maxd > 0, "delay table: DM 0 row is zero, max_delay positive");
    vfrb_delay_table(dms, 5, 4096, 383.9921875, 0.015625, 0.001024, off.data(), &maxd, &margin);
    vfrb_delay_table(dms, 0, 4096, 383.9921875, 0.015625, 0.001024, off.data(), &maxd, nullptr);
    vfrb_delay_table(dms, 2, 0, 383.9921875, 0.015625, 0.001024, off.data(), &maxd, nullptr);
    vfrb_delay_table(dms, 2, 16, 0.0, 0.015625, 0.001024, off.data(), &maxd, nullptr);
    vfrb_delay_table(dms, 2, 16, 383.9, 0.015625, 0.0, off.data(), &maxd, nullptr);
    vfrb_delay_table(nullptr, 2, 16, 383.9, 0.015625, 0.001, off.data(), &maxd, nullptr);
    vfrb_delay_table(dms, 4, 4096, 383.9921875, 0.015625, 0.001024, nullptr, &maxd, nullptr);
}

void vdif_codec() {
    Rng rng{7};
    for (int i = 0; i < 2000; i++) {
        uint8_t raw[32], raw2[32];
        for (uint8_t& b : raw) b = (uint8_t)rng.next();
        vfrb_vdif_header h;
        check(vfrb_vdif_decode(raw, &h) == VFRB_OK, "decode");
        check(vfrb_vdif_encode(&h, raw2) == VFRB_OK && std::memcmp(raw, raw2, 32) == 0, "decode/encode round trip");
        const double mjd = vfrb_vdif_dmjd(raw, kPerSec);
        check(mjd == mjd, "dmjd is a number");
    }
    check(vfrb_vdif_decode(nullptr, nullptr) != VFRB_OK, "decode NULL rejected");
    check(vfrb_vdif_dmjd(nullptr, kPerSec) == 0.0, "dmjd NULL");
}

void writers() {
    uint8_t buf[600];
    int32_t len = 0;
    for (int cap = 0; cap <= 600; cap += 7)
        vfrb_sigproc_header(buf, cap, &len, 383.99, -0.0156, 4096, 32, 60000.5, 0.001024);
    check(vfrb_sigproc_header(buf, 600, &len, 383.99, -0.0156, 4096, 32, 60000.5, 0.001024) == VFRB_OK && len > 0,
          "sigproc header");
    char txt[4096];
    const std::string longname(3000, 'x');
    for (int cap = 0; cap <= 4096; cap += 61) {
        vfrb_presto_inf(txt, cap, "DM12.3_20261203_120000", 60000.5, 19531, 0.001024, 12.3f, 320.0, 64.0, 4096,
                        0.015625);
        vfrb_presto_inf(txt, cap, longname.c_str(), 60000.5, 19531, 0.001024, 12.3f, 320.0, 64.0, 4096, 0.015625);
    }
}

}  // namespace

int main() {
    params_and_tables();
    vdif_codec();
    writers();
    sequencer_cases();
    sequencer_fuzz(2000);
    std::printf("host checks: %d failure(s)\n", failures);
    return failures ? 1 : 0;
}
