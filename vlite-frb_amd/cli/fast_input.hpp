// Fast input path of vdifFILE2dedisp_PIPELINE: the raw VDIF frame pairs of
// the input file are read in large chunks (parallel pread into pinned double
// buffers, the next chunk read while the current one is on the GPU), copied
// to the device, validated and gap-filled by the device ingest, and gathered
// into the per-pol payload stream the host frame sequencer would deliver.
// Bundles are cut from that stream by the sequencer's own rule: bundle b
// (bytes [b*NLOOP, (b+1)*NLOOP)) is delivered only when a payload slot exists
// past its last byte, because filltimebufferHOST reads the next frame pair
// after each frame it finishes copying (PIPELINE:1135-1171) and a bundle whose
// read-ahead hits the end of the file is dropped.
#pragma once

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <future>
#include <string>
#include <thread>
#include <vector>

#include "cli_common.hpp"

namespace cli {

constexpr int kFastFallback = 3;

// Reads n bytes at offset off into dst with `nthreads` parallel preads.
inline bool pread_all(int fd, uint8_t* dst, int64_t n, int64_t off, int nthreads) {
    std::vector<std::thread> th;
    std::vector<int> ok(nthreads, 1);
    const int64_t part = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; t++) {
        const int64_t a = std::min(n, t * part), b = std::min(n, a + part);
        th.emplace_back([&, t, a, b]() {
            int64_t done = a;
            while (done < b) {
                const ssize_t r = pread(fd, dst + done, (size_t)(b - done), (off_t)(off + done));
                if (r <= 0) {
                    ok[t] = 0;
                    return;
                }
                done += r;
            }
        });
    }
    for (auto& x : th) x.join();
    return std::all_of(ok.begin(), ok.end(), [](int v) { return v != 0; });
}

// Consumes the whole input through `consume(pol0, pol1, nbundles)` (device
// pointers to consecutive bundles).  `seq` has been through vfrb_seq_init,
// which found the pair the stream starts on.  Returns 0 at the end of the
// input, kFastFallback (with the reason in *why) when the device ingest
// rejects the stream or the shape does not suit this path, 2 on an I/O or
// device error.
template <class Consume>
int fast_frames(const char* path, vfrb_seq* seq, int64_t nloop, Consume&& consume, FILE* flog, std::string* why) {
    constexpr int64_t kPair = 2 * (int64_t)VFRB_FRAMELEN;
    // 41 MB of frames (0.16 s of data) per chunk, read by 8 threads: on one
    // MI355X host the loop then runs at ~22 GB/s of file data from the page
    // cache (2048-16384 pairs and 8-16 readers within 5 % of each other)
    constexpr int64_t kChunkPairs = 4096;
    constexpr int64_t kTableCap = kChunkPairs + VFRB_FRAMESPERSEC;      // room for a one-second gap per chunk
    constexpr int kReaders = 8;
    if (nloop % 4 != 0) {
        *why = "bundle size not a multiple of 4 bytes";
        return kFastFallback;
    }
    int64_t consumed = 0;
    vfrb_seq_consumed(seq, &consumed);
    const int64_t start = consumed - kPair;
    const int fd = open(path, O_RDONLY);
    struct stat sb {};
    if (fd < 0 || fstat(fd, &sb) != 0 || !S_ISREG(sb.st_mode) || start < 0) {
        if (fd >= 0) close(fd);
        *why = "input is not a regular file";
        return kFastFallback;
    }
    const int64_t npairs_total = std::max<int64_t>(0, (int64_t)sb.st_size - start) / kPair;
    const int64_t pay_bytes = nloop + kTableCap * VFRB_DATALEN;
    uint8_t* host[2] = {nullptr, nullptr};
    uint8_t *frames = nullptr, *pay[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    int32_t* table = nullptr;
    vfrb_ingest* ig = nullptr;
    int rc = 0;
    auto fail = [&](int code, const std::string& w) {
        if (rc == 0) {
            rc = code;
            *why = w;
        }
    };
    for (int k = 0; k < 2 && rc == 0; k++)
        if (vfrb_host_alloc((void**)&host[k], (size_t)(kChunkPairs * kPair)) != VFRB_OK) fail(2, vfrb_last_error());
    if (rc == 0 && vfrb_malloc((void**)&frames, (size_t)(kChunkPairs * kPair)) != VFRB_OK) fail(2, vfrb_last_error());
    if (rc == 0 && vfrb_malloc((void**)&table, (size_t)kTableCap * sizeof(int32_t)) != VFRB_OK)
        fail(2, vfrb_last_error());
    for (int pol = 0; pol < 2 && rc == 0; pol++)
        for (int k = 0; k < 2 && rc == 0; k++)
            if (vfrb_malloc((void**)&pay[pol][k], (size_t)pay_bytes) != VFRB_OK) fail(2, vfrb_last_error());
    if (rc == 0 && vfrb_ingest_create(&ig, kChunkPairs, 0) != VFRB_OK) fail(2, vfrb_last_error());

    int64_t slots = 0, bdone = 0, base = 0, fill = 0;
    int cur = 0;
    std::future<bool> next;
    auto read_chunk = [&](int k, int64_t pair0, int64_t n) {
        return std::async(std::launch::async, [=]() {
            return pread_all(fd, host[k], n * kPair, start + pair0 * kPair, kReaders);
        });
    };
    if (rc == 0 && npairs_total > 0) next = read_chunk(0, 0, std::min(kChunkPairs, npairs_total));
    for (int64_t c = 0, pair0 = 0; rc == 0 && pair0 < npairs_total; c++) {
        const int64_t n = std::min(kChunkPairs, npairs_total - pair0);
        if (!next.get()) {
            fail(2, "read error on the input file");
            break;
        }
        // chunk c+1 is read while chunk c is on the device; its buffer was
        // last read by the (synchronous) copy of chunk c-1
        if (pair0 + n < npairs_total)
            next = read_chunk((int)((c + 1) & 1), pair0 + n, std::min(kChunkPairs, npairs_total - pair0 - n));
        if (vfrb_memcpy_h2d(frames, host[c & 1], (size_t)(n * kPair), nullptr) != VFRB_OK) {
            fail(2, vfrb_last_error());
            break;
        }
        int64_t ns = 0;
        const vfrb_status st = vfrb_ingest_frames(ig, frames, n, table, kTableCap, &ns, nullptr);
        if (st == VFRB_ERR_FORMAT || st == VFRB_ERR_CAPACITY) {
            fail(kFastFallback, vfrb_last_error());
            break;
        }
        if (st != VFRB_OK) {
            fail(2, vfrb_last_error());
            break;
        }
        for (int pol = 0; pol < 2 && rc == 0; pol++)
            if (vfrb_ingest_gather(frames, table, ns, pol, pay[pol][cur] + fill, nullptr) != VFRB_OK)
                fail(2, vfrb_last_error());
        if (rc) break;
        fill += ns * VFRB_DATALEN;
        slots += ns;
        // bundles whose read-ahead slot exists: (b+1)*NLOOP < 5000*slots
        const int64_t safe = (VFRB_DATALEN * slots - 1) / nloop;
        if (safe > bdone) {
            const int64_t off = bdone * nloop - base;
            consume(pay[0][cur] + off, pay[1][cur] + off, safe - bdone);
            bdone = safe;
        }
        // the bytes of bundles not yet complete (< NLOOP) move to the other buffer
        const int64_t off = bdone * nloop - base, left = fill - off;
        for (int pol = 0; pol < 2 && left > 0 && rc == 0; pol++)
            if (vfrb_memcpy_d2d(pay[pol][1 - cur], pay[pol][cur] + off, (size_t)left, nullptr) != VFRB_OK)
                fail(2, vfrb_last_error());
        cur = 1 - cur;
        base = bdone * nloop;
        fill = left;
        pair0 += n;
    }
    if (next.valid()) next.wait();
    if (rc == 0) std::fprintf(flog, "fast input path: %lld frame pairs, %lld payload slots, %lld bundles\n",
                              (long long)npairs_total, (long long)slots, (long long)bdone);
    vfrb_device_synchronize();
    if (ig) vfrb_ingest_destroy(ig);
    for (int pol = 0; pol < 2; pol++)
        for (int k = 0; k < 2; k++)
            if (pay[pol][k]) vfrb_free(pay[pol][k]);
    if (table) vfrb_free(table);
    if (frames) vfrb_free(frames);
    for (int k = 0; k < 2; k++)
        if (host[k]) vfrb_host_free(host[k]);
    close(fd);
    return rc;
}

}  // namespace cli
