// Device-side VDIF frame ingest: raw frame pairs already in HBM (a NIC-to-GPU
// path) -> slot table of the two per-polarisation payload streams, with the
// reference's zero-filled runs for missing frames (PIPELINE:888-1025,
// seconds-aware rule TIMESERIES:893-937).
//
// A pair k is frames 2k (thread 0, pol 0) and 2k+1 (thread 1, pol 1) with
// the same frame number (and second, for the seconds-aware rule).  Frame k's
// run of zero payloads has length gap_k, from the frame number expected after
// pair k-1 (or the carried state for k = 0); its payload lands in slot
// P_k = sum_{i<=k}(gap_i + 1) - 1 after the batch's first slot, the gap_k
// slots before it stay zero.  Both pols advance together, so one table serves
// both.  A pair whose threads, frame numbers or seconds disagree is a format
// error: the host sequencer (vfrb_seq_*) reproduces the reference on such
// streams.
#include <algorithm>
#include <mutex>

#include "common.hpp"

namespace {

constexpr int kFrame = 5032;
constexpr int kFps = 25600;
constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;                     // pairs per thread in the block scan
constexpr int kScanTile = kScanBlock * kScanItems;  // pairs per scan block

struct IngestState {
    int expected;      // frame number expected next, -1 = take the first pair as is
    int expected_sec;  // second expected next (seconds-aware rule)
};

__device__ __forceinline__ void decode(const uint8_t* f, int* sec, int* fnum, int* thread) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(f);
    *sec = (int)(w[0] & 0x3fffffffu);
    *fnum = (int)(w[1] & 0xffffffu);
    *thread = (int)((w[3] >> 16) & 0x3ffu);
}

// gap[k] = zero payloads inserted before pair k; bad = 1 on a malformed pair.
__global__ void __launch_bounds__(256) ingest_gap_kernel(const uint8_t* __restrict__ raw, int64_t npairs,
                                                         int sec_check, const IngestState* __restrict__ st,
                                                         int* __restrict__ gap, int* __restrict__ bad) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npairs) return;
    const uint8_t* f0 = raw + 2 * k * kFrame;
    int s0, n0, t0, s1, n1, t1;
    decode(f0, &s0, &n0, &t0);
    decode(f0 + kFrame, &s1, &n1, &t1);
    if (t0 != 0 || t1 != 1 || n0 != n1 || (sec_check && s0 != s1)) atomicOr(bad, 1);
    int exp, exps;
    if (k == 0) {
        exp = st->expected;
        exps = st->expected_sec;
    } else {
        int sp, np, tp;
        decode(f0 - 2 * kFrame, &sp, &np, &tp);
        exp = (np + 1) % kFps;
        exps = sp + (sec_check && exp == 0 ? 1 : 0);
    }
    int g = 0;
    if (exp >= 0 && (n0 != exp || (sec_check && s0 != exps))) {
        g = n0 - exp;
        if (sec_check) {
            // 32-bit two's-complement wrap of the reference's int arithmetic
            // (TIMESERIES:901) on a damaged seconds field, without overflow
            const int64_t g64 = (int64_t)g + ((int64_t)s0 - exps) * kFps;
            g = (int)(int32_t)(uint32_t)(uint64_t)g64;
            if (g < 0) g = 0;  // frames going backwards: no zero run
        } else if (g < 0) {
            g += kFps;
        }
    }
    gap[k] = g;
}

// Inclusive scan of (gap + 1) within blocks of kScanTile pairs.
__global__ void __launch_bounds__(kScanBlock) ingest_scan_kernel(const int* __restrict__ gap, int64_t npairs,
                                                                  int64_t* __restrict__ incl,
                                                                  int64_t* __restrict__ blk_sum) {
    __shared__ int64_t part[kScanBlock];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int64_t v[kScanItems];
    int64_t run = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        run += base + i < npairs ? (int64_t)gap[base + i] + 1 : 0;
        v[i] = run;
    }
    part[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        const int64_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    const int64_t prefix = threadIdx.x > 0 ? part[threadIdx.x - 1] : 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++)
        if (base + i < npairs) incl[base + i] = prefix + v[i];
    if (threadIdx.x == kScanBlock - 1) blk_sum[blockIdx.x] = part[kScanBlock - 1];
}

// Exclusive scan of the block sums (one workgroup, any count) and the total.
__global__ void __launch_bounds__(kScanBlock) ingest_blocks_kernel(int64_t* __restrict__ blk_sum, int nblk,
                                                                    int64_t* __restrict__ total) {
    __shared__ int64_t part[kScanBlock];
    int64_t carry = 0;
    for (int b0 = 0; b0 < nblk; b0 += kScanBlock) {
        const int b = b0 + (int)threadIdx.x;
        const int64_t mine = b < nblk ? blk_sum[b] : 0;
        part[threadIdx.x] = mine;
        __syncthreads();
        for (int off = 1; off < kScanBlock; off <<= 1) {
            const int64_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (b < nblk) blk_sum[b] = carry + part[threadIdx.x] - mine;  // exclusive
        const int64_t tile = part[kScanBlock - 1];
        __syncthreads();
        carry += tile;
    }
    if (threadIdx.x == 0) total[0] = carry;
}

// table[s] = -1 for every slot of the batch (bounded by the capacity).
__global__ void __launch_bounds__(256) ingest_clear_kernel(int32_t* __restrict__ table,
                                                            const int64_t* __restrict__ total, int64_t cap) {
    const int64_t n = min(total[0], cap);
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x)
        table[s] = -1;
}

// table[P_k] = k; the last pair carries the expected frame into the state.
__global__ void __launch_bounds__(256) ingest_place_kernel(const uint8_t* __restrict__ raw, int64_t npairs,
                                                            int sec_check, const int64_t* __restrict__ incl,
                                                            const int64_t* __restrict__ blk_off,
                                                            int32_t* __restrict__ table, int64_t cap,
                                                            IngestState* __restrict__ st) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npairs) return;
    const int64_t slot = blk_off[k / kScanTile] + incl[k] - 1;
    if (slot < cap) table[slot] = (int32_t)k;
    if (k == npairs - 1) {
        int s, n, t;
        decode(raw + 2 * k * kFrame, &s, &n, &t);
        st->expected = (n + 1) % kFps;
        st->expected_sec = s + (sec_check && (n + 1) % kFps == 0 ? 1 : 0);
    }
}

// Payload byte o of pol p: slot o/5000 through the table, 0 for a zero slot.
__global__ void __launch_bounds__(256) ingest_gather_kernel(const uint8_t* __restrict__ raw,
                                                             const int32_t* __restrict__ table, int64_t ndw,
                                                             int pol, uint32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndw) return;
    const int64_t o = 4 * i;
    const int64_t s = o / 5000;
    const int w = (int)(o - s * 5000);
    const int k = table[s];
    out[i] = k < 0 ? 0u : *reinterpret_cast<const uint32_t*>(raw + ((int64_t)2 * k + pol) * kFrame + 32 + w);
}

}  // namespace

struct vfrb_ingest {
    int64_t max_pairs = 0;
    int sec_check = 0;
    // Two parse states: [cur] is the one the next batch continues from; a
    // batch writes its successor into [1 - cur], and the host switches only
    // after the batch passed its format and capacity checks, so a failed
    // call leaves the stream's state as it was.
    IngestState* d_state = nullptr;
    int cur = 0;
    int* d_gap = nullptr;
    int64_t* d_incl = nullptr;
    int64_t* d_blk = nullptr;
    int64_t* d_total = nullptr;
    int* d_bad = nullptr;
    int64_t* h_total = nullptr;  // pinned
    int* h_bad = nullptr;        // pinned
};

namespace {

void ingest_free(vfrb_ingest* ig) {
    (void)hipFree(ig->d_state);
    (void)hipFree(ig->d_gap);
    (void)hipFree(ig->d_incl);
    (void)hipFree(ig->d_blk);
    (void)hipFree(ig->d_total);
    (void)hipFree(ig->d_bad);
    (void)hipHostFree(ig->h_total);
    (void)hipHostFree(ig->h_bad);
    delete ig;
}

__global__ void ingest_zero_kernel(int* p) { *p = 0; }

// ---------------------------------------------------------------------------
// Ingest without allocation: the same kernels, with their scratch (gap runs,
// scan partials, slot total, format flag, parse states) in one file-scope
// device array instead of buffers of a vfrb_ingest.  Each GPU holds its own
// copy until the process ends; calls on one GPU share it.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocMaxPairs = int64_t(1) << 20;  // 41 s of VLITE frame pairs
constexpr size_t align256(size_t x) { return (x + 255) / 256 * 256; }

struct NoAllocTail {         // read back to the host in one copy
    int64_t total;           // payload slots of the batch
    int bad;                 // a malformed pair was seen
    int pad;
    IngestState state[2];    // [0] the caller's state, [1] the successor
};

constexpr size_t kNaGap = 0;
constexpr size_t kNaIncl = align256(kNaGap + sizeof(int) * kNoAllocMaxPairs);
constexpr size_t kNaBlk = align256(kNaIncl + sizeof(int64_t) * kNoAllocMaxPairs);
constexpr size_t kNaTail = align256(kNaBlk + sizeof(int64_t) * ((kNoAllocMaxPairs + kScanTile - 1) / kScanTile));
constexpr size_t kScratchBytes = align256(kNaTail + sizeof(NoAllocTail));  // 12.6 MB

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The device address of `scratch` on each GPU, looked up once per device
// with that device current (host code never passes the symbol itself).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};

vfrb_status scratch_of_current_device(unsigned char** out) {
    int dev = -1;
    VFRB_HIP(hipGetDevice(&dev));
    VFRB_REQUIRE(dev >= 0 && dev < kMaxDevices, VFRB_ERR_UNSUPPORTED, "device %d", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        VFRB_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        VFRB_REQUIRE(ptr != nullptr, VFRB_ERR_HIP, "no device address for the ingest scratch");
        g_scratch[dev] = static_cast<unsigned char*>(ptr);
    }
    *out = g_scratch[dev];
    return VFRB_OK;
}

__global__ void ingest_state_set_kernel(NoAllocTail* t, int expected, int expected_sec) {
    t->total = 0;
    t->bad = 0;
    t->state[0].expected = expected;
    t->state[0].expected_sec = expected_sec;
}

}  // namespace

extern "C" {

vfrb_status vfrb_ingest_create(vfrb_ingest** out, int64_t max_pairs, int32_t seconds_check) {
    VFRB_REQUIRE(out && max_pairs > 0, VFRB_ERR_INVALID, "bad ingest arguments");
    *out = nullptr;
    vfrb_ingest* ig = new vfrb_ingest();
    ig->max_pairs = max_pairs;
    ig->sec_check = seconds_check != 0;
    const int64_t nblk = vfrb::ceil_div(max_pairs, kScanTile);
    hipError_t e = hipMalloc(&ig->d_state, 2 * sizeof(IngestState));
    if (e == hipSuccess) e = hipMalloc(&ig->d_gap, sizeof(int) * (size_t)max_pairs);
    if (e == hipSuccess) e = hipMalloc(&ig->d_incl, sizeof(int64_t) * (size_t)max_pairs);
    if (e == hipSuccess) e = hipMalloc(&ig->d_blk, sizeof(int64_t) * (size_t)nblk);
    if (e == hipSuccess) e = hipMalloc(&ig->d_total, sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc(&ig->d_bad, sizeof(int));
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&ig->h_total), sizeof(int64_t));
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&ig->h_bad), sizeof(int));
    if (e == hipSuccess) {
        const IngestState init{-1, 0};
        e = hipMemcpy(ig->d_state, &init, sizeof init, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        const vfrb_status st = vfrb::hip_fail(e, "vfrb_ingest_create", __FILE__, __LINE__);
        ingest_free(ig);
        return st;
    }
    *out = ig;
    return VFRB_OK;
}

vfrb_status vfrb_ingest_destroy(vfrb_ingest* ig) {
    if (ig) ingest_free(ig);
    return VFRB_OK;
}

vfrb_status vfrb_ingest_reset(vfrb_ingest* ig, int32_t expected_frame, int32_t expected_second) {
    VFRB_REQUIRE(ig && expected_frame >= -1 && expected_frame < kFps, VFRB_ERR_INVALID, "bad reset");
    const IngestState st{expected_frame, expected_second};
    VFRB_HIP(hipMemcpy(ig->d_state + ig->cur, &st, sizeof st, hipMemcpyHostToDevice));
    return VFRB_OK;
}

vfrb_status vfrb_ingest_frames(vfrb_ingest* ig, const uint8_t* raw, int64_t npairs, int32_t* table,
                               int64_t table_cap, int64_t* nslots, void* stream) {
    VFRB_REQUIRE(ig && raw && table && nslots, VFRB_ERR_INVALID, "NULL argument");
    VFRB_REQUIRE(npairs > 0 && npairs <= ig->max_pairs, VFRB_ERR_INVALID, "npairs=%lld outside 1..%lld",
                 (long long)npairs, (long long)ig->max_pairs);
    VFRB_REQUIRE(((uintptr_t)raw % 4) == 0, VFRB_ERR_INVALID, "frames must be 4-byte aligned");
    hipStream_t s = vfrb::as_stream(stream);
    const unsigned g = (unsigned)vfrb::ceil_div(npairs, 256);
    const int nblk = (int)vfrb::ceil_div(npairs, kScanTile);
    hipLaunchKernelGGL(ingest_zero_kernel, dim3(1), dim3(1), 0, s, ig->d_bad);
    hipLaunchKernelGGL(ingest_gap_kernel, dim3(g), dim3(256), 0, s, raw, npairs, ig->sec_check,
                       ig->d_state + ig->cur, ig->d_gap, ig->d_bad);
    hipLaunchKernelGGL(ingest_scan_kernel, dim3((unsigned)nblk), dim3(kScanBlock), 0, s, ig->d_gap, npairs,
                       ig->d_incl, ig->d_blk);
    hipLaunchKernelGGL(ingest_blocks_kernel, dim3(1), dim3(kScanBlock), 0, s, ig->d_blk, nblk, ig->d_total);
    hipLaunchKernelGGL(ingest_clear_kernel, dim3(1024), dim3(256), 0, s, table, ig->d_total, table_cap);
    hipLaunchKernelGGL(ingest_place_kernel, dim3(g), dim3(256), 0, s, raw, npairs, ig->sec_check, ig->d_incl,
                       ig->d_blk, table, table_cap, ig->d_state + (1 - ig->cur));
    VFRB_LAUNCH_CHECK("ingest kernels");
    VFRB_HIP(hipMemcpyAsync(ig->h_total, ig->d_total, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    VFRB_HIP(hipMemcpyAsync(ig->h_bad, ig->d_bad, sizeof(int), hipMemcpyDeviceToHost, s));
    VFRB_HIP(hipStreamSynchronize(s));
    VFRB_REQUIRE(*ig->h_bad == 0, VFRB_ERR_FORMAT,
                 "a frame pair is not thread 0/1 with equal frame numbers (and seconds)");
    VFRB_REQUIRE(*ig->h_total <= table_cap, VFRB_ERR_CAPACITY, "%lld payload slots exceed the table of %lld",
                 (long long)*ig->h_total, (long long)table_cap);
    *nslots = *ig->h_total;
    ig->cur = 1 - ig->cur;
    return VFRB_OK;
}

vfrb_status vfrb_ingest_frames_noalloc(const uint8_t* raw, int64_t npairs, int32_t seconds_check,
                                       int32_t* expected_frame, int32_t* expected_second, int32_t* table,
                                       int64_t table_cap, int64_t* nslots, void* stream) {
    VFRB_REQUIRE(raw && table && nslots && expected_frame && expected_second, VFRB_ERR_INVALID, "NULL argument");
    VFRB_REQUIRE(npairs > 0, VFRB_ERR_INVALID, "npairs=%lld", (long long)npairs);
    VFRB_REQUIRE(npairs <= kNoAllocMaxPairs, VFRB_ERR_CAPACITY,
                 "npairs=%lld exceeds the %lld frame pairs the no-allocation scratch holds "
                 "(vfrb_ingest_create + vfrb_ingest_frames take any count)",
                 (long long)npairs, (long long)kNoAllocMaxPairs);
    VFRB_REQUIRE(*expected_frame >= -1 && *expected_frame < kFps, VFRB_ERR_INVALID, "expected_frame=%d",
                 *expected_frame);
    VFRB_REQUIRE(((uintptr_t)raw % 4) == 0, VFRB_ERR_INVALID, "frames must be 4-byte aligned");
    unsigned char* sp = nullptr;
    const vfrb_status st = scratch_of_current_device(&sp);
    if (st != VFRB_OK) return st;
    int* gap = reinterpret_cast<int*>(sp + kNaGap);
    int64_t* incl = reinterpret_cast<int64_t*>(sp + kNaIncl);
    int64_t* blk = reinterpret_cast<int64_t*>(sp + kNaBlk);
    NoAllocTail* tail = reinterpret_cast<NoAllocTail*>(sp + kNaTail);
    hipStream_t s = vfrb::as_stream(stream);
    const int sec = seconds_check != 0;
    const unsigned g = (unsigned)vfrb::ceil_div(npairs, 256);
    const int nblk = (int)vfrb::ceil_div(npairs, kScanTile);
    // the previous call's contents are stale: total, flag and state start anew
    hipLaunchKernelGGL(ingest_state_set_kernel, dim3(1), dim3(1), 0, s, tail, *expected_frame, *expected_second);
    hipLaunchKernelGGL(ingest_gap_kernel, dim3(g), dim3(256), 0, s, raw, npairs, sec, &tail->state[0], gap,
                       &tail->bad);
    hipLaunchKernelGGL(ingest_scan_kernel, dim3((unsigned)nblk), dim3(kScanBlock), 0, s, gap, npairs, incl, blk);
    hipLaunchKernelGGL(ingest_blocks_kernel, dim3(1), dim3(kScanBlock), 0, s, blk, nblk, &tail->total);
    hipLaunchKernelGGL(ingest_clear_kernel, dim3(1024), dim3(256), 0, s, table, &tail->total, table_cap);
    hipLaunchKernelGGL(ingest_place_kernel, dim3(g), dim3(256), 0, s, raw, npairs, sec, incl, blk, table, table_cap,
                       &tail->state[1]);
    VFRB_LAUNCH_CHECK("ingest kernels (no-allocation scratch)");
    NoAllocTail h{};
    VFRB_HIP(hipMemcpyAsync(&h, tail, sizeof h, hipMemcpyDeviceToHost, s));
    VFRB_HIP(hipStreamSynchronize(s));
    VFRB_REQUIRE(h.bad == 0, VFRB_ERR_FORMAT,
                 "a frame pair is not thread 0/1 with equal frame numbers (and seconds)");
    VFRB_REQUIRE(h.total <= table_cap, VFRB_ERR_CAPACITY, "%lld payload slots exceed the table of %lld",
                 (long long)h.total, (long long)table_cap);
    *nslots = h.total;
    *expected_frame = h.state[1].expected;
    *expected_second = h.state[1].expected_sec;
    return VFRB_OK;
}

vfrb_status vfrb_ingest_gather(const uint8_t* raw, const int32_t* table, int64_t nslots, int32_t pol,
                               uint8_t* out, void* stream) {
    VFRB_REQUIRE(raw && table && out && nslots >= 0 && (pol == 0 || pol == 1), VFRB_ERR_INVALID,
                 "bad gather arguments");
    VFRB_REQUIRE(((uintptr_t)out % 4) == 0 && ((uintptr_t)raw % 4) == 0, VFRB_ERR_INVALID,
                 "buffers must be 4-byte aligned");
    const int64_t ndw = nslots * 5000 / 4;
    if (ndw == 0) return VFRB_OK;
    hipLaunchKernelGGL(ingest_gather_kernel, dim3((unsigned)vfrb::ceil_div(ndw, 256)), dim3(256), 0,
                       vfrb::as_stream(stream), raw, table, ndw, pol, reinterpret_cast<uint32_t*>(out));
    VFRB_LAUNCH_CHECK("ingest_gather_kernel");
    return VFRB_OK;
}

}  // extern "C"
