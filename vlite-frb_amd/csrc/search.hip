// Back end: window load (bandpass + channel-major pack in the storage type)
// and brute-force dedispersion over LDS-tiled channel x time blocks.
//
// Data layout in HBM
//   fb ring   : f32 [ring_rows][nch], time-major, written by the front end.
//   window    : storage type [nch][ld], channel-major residuals of one window.
//   out       : f32 [ndm][ld_out].
// The dedispersion of DM d and sample t reads window[c][t + off[d][c]] for
// all channels c.  A workgroup owns a block of DMs x a tile of T samples;
// for each chunk of channels it stages window[c][t0+dmin_c .. t0+T+dmax_c)
// in LDS once and every wave accumulates its DMs from there.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include <hip/hip_fp16.h>

#include "bandpass.hpp"
#include "common.hpp"

using vfrb::BP_CH;
using vfrb::BP_PER;
using vfrb::BP_Q;
using vfrb::BP_ROWS;

namespace {

// ---------------------------------------------------------------------------
// Window load
// ---------------------------------------------------------------------------
enum LoadMode { LOAD_F32 = 0, LOAD_STATS = 1, LOAD_APPLY_F16 = 2, LOAD_APPLY_U8 = 3 };

__device__ __forceinline__ unsigned float_order(float f) {
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float float_unorder(unsigned u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// grid = (ceil(nch/64), ceil(nrows/128)), block = 256.
template <int MODE>
__global__ void __launch_bounds__(256)
load_kernel(const float* __restrict__ fb, int64_t ring_rows, int64_t row0, int nrows, int nch,
            int apply_bp, void* __restrict__ win, int64_t ld, float* __restrict__ bp_mean,
            unsigned* __restrict__ range /* [2][nch] ordered min/max, or [1] global max|.| */,
            const int* __restrict__ f16_exp, const float* __restrict__ u8_scale,
            const float* __restrict__ u8_off) {
    __shared__ float red[BP_Q * BP_CH];
    __shared__ int ired[BP_Q * BP_CH];
    __shared__ float tile[BP_CH][BP_ROWS + 1];
    const int lane = threadIdx.x & 63;
    const int q = threadIdx.x >> 6;
    const int c = blockIdx.x * BP_CH + lane;
    const int blk = blockIdx.y;
    const int rbase = blk * BP_ROWS;
    const int npts = min(BP_ROWS, nrows - rbase);
    const int mine0 = q * BP_PER;
    const int n_mine = max(0, min(BP_PER, npts - mine0));
    const bool cvalid = c < nch;

    float x[BP_PER];
#pragma unroll
    for (int i = 0; i < BP_PER; i++) {
        x[i] = 0.0f;
        if (i < n_mine && cvalid) {
            int64_t r = row0 + rbase + mine0 + i;
            if (ring_rows > 0) r %= ring_rows;
            x[i] = fb[r * nch + c];
        }
    }
    float m = 0.0f;
    if (MODE == LOAD_F32 || MODE == LOAD_STATS) {
        if (apply_bp) m = vfrb::bp_clipped_mean(x, cvalid ? n_mine : 0, npts, red, ired, lane, q);
        if (MODE == LOAD_STATS) {
            if (q == 0 && cvalid) bp_mean[(int64_t)blk * nch + c] = m;
            if (cvalid && n_mine > 0) {
                float lo = INFINITY, hi = -INFINITY;
#pragma unroll
                for (int i = 0; i < BP_PER; i++) {
                    if (i < n_mine) {
                        const float y = x[i] - m;
                        lo = fminf(lo, y);
                        hi = fmaxf(hi, y);
                    }
                }
                if (range) {
                    // range[0..nch): ordered min, range[nch..2nch): ordered max
                    atomicMin(&range[c], float_order(lo));
                    atomicMax(&range[nch + c], float_order(hi));
                }
            }
            return;
        }
    } else {
        if (apply_bp && cvalid) m = bp_mean[(int64_t)blk * nch + c];
    }
    // residuals -> LDS tile [channel][row]
#pragma unroll
    for (int i = 0; i < BP_PER; i++) tile[lane][mine0 + i] = x[i] - m;
    __syncthreads();
    // channel-major store: each wave writes whole channel rows, 2 rows/lane
    const int w = threadIdx.x >> 6;
    const int r0 = 2 * lane;
    for (int cl = w; cl < BP_CH; cl += 4) {
        const int cc = blockIdx.x * BP_CH + cl;
        if (cc >= nch) break;
        const float v0 = tile[cl][r0], v1 = tile[cl][r0 + 1];
        const int64_t o = (int64_t)cc * ld + rbase + r0;
        if (MODE == LOAD_F32) {
            float* dst = reinterpret_cast<float*>(win) + o;
            if (r0 + 1 < npts) {
                *reinterpret_cast<float2*>(dst) = make_float2(v0, v1);
            } else if (r0 < npts) {
                dst[0] = v0;
            }
        } else if (MODE == LOAD_APPLY_F16) {
            const float s = ldexpf(1.0f, -f16_exp[0]);
            __half* dst = reinterpret_cast<__half*>(win) + o;
            if (r0 + 1 < npts) {
                *reinterpret_cast<__half2*>(dst) = __floats2half2_rn(v0 * s, v1 * s);
            } else if (r0 < npts) {
                dst[0] = __float2half_rn(v0 * s);
            }
        } else {  // LOAD_APPLY_U8
            const float sc = u8_scale[cc], of = u8_off[cc];
            const float inv = 1.0f / sc;
            auto qz = [&](float v) {
                float t = rintf((v - of) * inv);
                t = fminf(fmaxf(t, 0.0f), 255.0f);
                return (unsigned)t;
            };
            uint8_t* dst = reinterpret_cast<uint8_t*>(win) + o;
            if (r0 + 1 < npts) {
                dst[0] = (uint8_t)qz(v0);
                dst[1] = (uint8_t)qz(v1);
            } else if (r0 < npts) {
                dst[0] = (uint8_t)qz(v0);
            }
        }
    }
}

// Turns the ordered per-channel ranges into the storage scales.
//  f16: one power-of-two exponent e with max|y| * 2^-e <= 32768.
//  u8 : per channel offset = min, scale = (max-min)/255 (1 if flat).
__global__ void finish_range_kernel(const unsigned* __restrict__ range, int nch, int storage,
                                    int* __restrict__ f16_exp, float* __restrict__ u8_scale,
                                    float* __restrict__ u8_off, float* __restrict__ u8_offsum) {
    __shared__ float part[256];
    __shared__ float partsum[256];
    float mx = 0.0f, osum = 0.0f;
    for (int c = threadIdx.x; c < nch; c += blockDim.x) {
        const float lo = float_unorder(range[c]);
        const float hi = float_unorder(range[nch + c]);
        if (storage == VFRB_STORE_F16) {
            mx = fmaxf(mx, fmaxf(fabsf(lo), fabsf(hi)));
        } else {
            float sc = (hi - lo) / 255.0f;
            if (!(sc > 0.0f) || !isfinite(sc)) sc = 1.0f;
            const float of = isfinite(lo) ? lo : 0.0f;
            u8_scale[c] = sc;
            u8_off[c] = of;
            osum += of;
        }
    }
    part[threadIdx.x] = mx;
    partsum[threadIdx.x] = osum;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            part[threadIdx.x] = fmaxf(part[threadIdx.x], part[threadIdx.x + s]);
            partsum[threadIdx.x] += partsum[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (storage == VFRB_STORE_F16) {
            int e = 0;
            const float m = part[0];
            if (isfinite(m) && m > 32768.0f) e = (int)ceilf(log2f(m / 32768.0f));
            f16_exp[0] = e;
        } else {
            u8_offsum[0] = partsum[0];
        }
    }
}

__global__ void range_init_kernel(unsigned* range, int nch) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nch) range[i] = 0xffffffffu;      // ordered +max -> atomicMin start
    else if (i < 2 * nch) range[i] = 0u;      // ordered -max -> atomicMax start
}

// ---------------------------------------------------------------------------
// Dedispersion
// ---------------------------------------------------------------------------
typedef _Float16 h2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr h2_t kSelLo = {(_Float16)1.0f, (_Float16)0.0f};
constexpr h2_t kSelHi = {(_Float16)0.0f, (_Float16)1.0f};

struct Chunk {
    int c0, count, stride;  // stride: staged samples per copy, multiple of 4
    int roff;               // first entry of this chunk in the LDS byte-offset table
    int nd;                 // pair tiles: presum rows per channel pair; group tiles: rows in the chunk
    int pad[3];             // group tiles: pad[0] = first int of the chunk's row records
};

// Tables read through pointers of unknown address space (the batch kernel's
// job records) load with flat vector loads; these helpers put the uniform
// values back into scalar registers so that loop bounds and strides stay
// uniform.
__device__ __forceinline__ int2 uniform_load(const int2* p) {
    const int2 v = *p;
    return make_int2(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y));
}
__device__ __forceinline__ Chunk uniform_load(const Chunk* p) {
    const Chunk v = *p;
    Chunk u;
    u.c0 = __builtin_amdgcn_readfirstlane(v.c0);
    u.count = __builtin_amdgcn_readfirstlane(v.count);
    u.stride = __builtin_amdgcn_readfirstlane(v.stride);
    u.roff = __builtin_amdgcn_readfirstlane(v.roff);
    u.nd = __builtin_amdgcn_readfirstlane(v.nd);
    u.pad[0] = __builtin_amdgcn_readfirstlane(v.pad[0]);
    u.pad[1] = u.pad[2] = 0;
    return u;
}

// At most kChunkMax channels per LDS chunk; staged rows take at most
// kDataBudget bytes, so with the offset table a workgroup stays under 80 KB
// and two fit in one CU's 160 KB of LDS.
constexpr int kChunkMax = 32;
constexpr size_t kDataBudget = 60 * 1024;
constexpr int kBufferFlags = 0x00020000;  // word 3 of a gfx9 raw buffer resource

// Group tiles: presum rows per chunk at most, and the record of each row
// (8 ints: for each of the group's 4 channels the byte offset, from the
// chunk's first channel row at time 0, of the dword that holds the channel's
// first element; the elements' phases in those dwords, one per byte; the
// row's LDS byte offset; the group's first channel in the chunk).
constexpr int kRowMax = 48;
constexpr int kRowInts = 8;

// LDS ahead of the staged rows: the chunk's [kChunkMax][DMB] u32 read offsets,
// then two buffers, one per chunk parity, of the chunk's channel start rows
// (dmin; group tiles: its row records), so that chunk k+1's are loaded while
// chunk k is staged.
__host__ __device__ constexpr size_t lds_head_bytes(int dmb) {
    return (size_t)kChunkMax * dmb * 4 + 2 * kRowMax * kRowInts * 4;
}

// Staging items (channel, 4-dword group) each thread loads per pass: enough
// for the largest chunk the data budget admits (kDataBudget / (16*EPD)
// groups) in one round trip to memory, as far as the registers allow: an
// item holds NSRC data dwords and two offsets, and staging may take a
// quarter of the per-lane VGPR budget of the tile's occupancy (1024/W).
template <int EPD, int W>
struct StageItems {
    static constexpr int groups = (int)(kDataBudget / (16 * EPD));
    static constexpr int need = (groups + 64 * W - 1) / (64 * W);
    static constexpr int fit = (1024 / W / 4) / ((EPD > 1 ? 5 : 4) + 2);
    static constexpr int cap = fit < 1 ? 1 : (fit > 8 ? 8 : fit);
    static constexpr int value = need < cap ? need : cap;
};

// LDS (<= 77 KB per workgroup) limits a CU to two workgroups, i.e. W/2 waves
// per SIMD.  The kernels ask for exactly that occupancy (both bounds of
// amdgpu_waves_per_eu): the register allocator then fills the VGPR budget it
// implies (512 / (W/2) per lane), so that many LDS reads are in flight, and
// never exceeds it, which would leave room for only one workgroup per CU.
template <int W>
struct WavesPerEu {
    static constexpr int value = W / 2 > 1 ? W / 2 : 1;
};

// One tile: DM block b (W*DW DMs) x time tile tx (T = 64*SL samples) of one
// window.  SL samples per lane, DW DMs per wave, W waves.  The tile ends
// with a barrier, so the caller may reuse the LDS at once.
// ST (f32 only): dedisperseSTATS as well, a second accumulator per sample
// sums the squares and sig receives the spread over channels.
template <typename TS, int SL, int DW, int W, bool ST = false>
__device__ __forceinline__ void
dedisp_tile(const TS* __restrict__ win, int64_t ld, int nch, int nrows, int nt, int ndm,
            const uint32_t* __restrict__ roffs,  // per chunk: [count][W*DW] LDS byte offsets
            const int* __restrict__ dmin,        // [nblk][nch]
            const Chunk* __restrict__ chunks, const int2* __restrict__ blk_chunks,
            const float* __restrict__ chan_scale,  // u8 only
            const int* __restrict__ f16_exp, const float* __restrict__ offsum,
            float* __restrict__ out, int64_t ld_out, int b, int tx, unsigned char* lds_raw,
            float* __restrict__ sig = nullptr, const float* __restrict__ shift = nullptr) {
    // LDS: [kChunkMax][DMB] u32 byte offsets of each (channel, DM) read,
    // then the staged rows.  A row holds COPIES copies of the channel's
    // samples, copy p shifted by p elements, so that the dword a lane reads
    // (EPD consecutive samples) is always 4-byte aligned whatever the delay.
    constexpr int ESZ = (int)sizeof(TS);
    constexpr int EPD = 4 / ESZ;           // samples per dword = copies per row
    constexpr int T = 64 * SL;
    constexpr int DMB = W * DW;
    constexpr int NDW = SL / EPD;          // dwords per lane per (DM, channel)
    static_assert(SL % EPD == 0, "SL must hold whole dwords");
    constexpr int NT = 64 * W;              // threads per workgroup
    constexpr int NSRC = EPD > 1 ? 5 : 4;   // source dwords per staged group
    constexpr int NI = StageItems<EPD, W>::value;
    constexpr int NR = (kChunkMax * DMB / 4 + NT - 1) / NT;  // offset-table uint4 per thread
    uint32_t* roff = reinterpret_cast<uint32_t*>(lds_raw);
    int* dbuf = reinterpret_cast<int*>(lds_raw + (size_t)kChunkMax * DMB * 4);
    unsigned char* data = lds_raw + lds_head_bytes(DMB);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int t0 = tx * T;
    const int dm0 = b * DMB + wave * DW;
    const int2 cr = uniform_load(blk_chunks + b);
    const int* dmin_b = dmin + (int64_t)b * nch;
    const uint32_t ldw = (uint32_t)(ld / EPD);  // dwords per window row

    static_assert(!ST || ESZ == 4, "dedisperseSTATS reads an f32 window");
    // dedisperseSTATS: a window-wide shift K (0 unless the window's level is
    // large against its spread) is subtracted from every staged sample, so
    // the one-pass sums S1 = sum(x - K), S2 = sum((x - K)^2) do not cancel.
    const float kshift = ST ? shift[0] : 0.0f;
    float acc[DW][SL];
    float acc2[ST ? DW : 1][ST ? SL : 1];
#pragma unroll
    for (int d = 0; d < DW; d++)
#pragma unroll
        for (int j = 0; j < SL; j++) {
            acc[d][j] = 0.0f;
            if constexpr (ST) acc2[d][j] = 0.0f;
        }

    // first chunk's channel start rows
    if (cr.y > 0) {
        const Chunk c0 = uniform_load(chunks + cr.x);
        if (tid < c0.count) dbuf[tid] = dmin_b[c0.c0 + tid];
    }
    __syncthreads();

    for (int k = 0; k < cr.y; k++) {
        const Chunk ch = uniform_load(chunks + cr.x + k);
        const int rowb = 4 * ch.stride;    // bytes per staged channel (all copies)
        const int ng = ch.stride / EPD / 4;  // 4-dword groups per copy
        const int items = ch.count * ng;
        const int* dcur = dbuf + (k & 1) * kChunkMax;
        // Issue every global load of the chunk before the first LDS write, so
        // that staging costs one round trip: the chunk's offset table, the
        // next chunk's start rows and the staged groups.
        // rv: byte offset of the first dword of each (channel, DM) read, built
        // on the host: a contiguous copy of count*DMB words (DMB % 4 == 0)
        u32x4 rv[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int i = tid + NT * r;
            if (i < ch.count * DMB / 4) rv[r] = reinterpret_cast<const u32x4*>(roffs + ch.roff)[i];
        }
        int dnext = 0;
        const bool more = k + 1 < cr.y;
        Chunk nx = ch;
        if (more) {
            nx = uniform_load(chunks + cr.x + k + 1);
            if (tid < nx.count) dnext = dmin_b[nx.c0 + tid];
        }
        // Item j = (channel ci, group g), ci = j / ng: a float quotient, off by
        // at most one for j < 2^15, corrected.  Group g of channel c covers
        // source dwords sd0 + 4g .. sd0 + 4g + 3 (+ one more for the shifted
        // copies), sd0 the dword-aligned start of the rows the tile reads:
        // (t0 + dmin) / EPD; the phase, dmin mod EPD, is folded into the
        // offset table.  It is written to dwords 4g..4g+3 of every copy, copy
        // p being the source shifted by p elements (v_alignbyte across
        // neighbouring dwords).  The loads are buffer loads from the chunk's
        // first window row (32-bit offsets, one VGPR per item) bounded by the
        // window's end: dwords past a row read the next row, or zero past the
        // last, and only feed samples t >= nt, never stored.
        const uint32_t* base = reinterpret_cast<const uint32_t*>(win) + (int64_t)ch.c0 * ldw;
        const int64_t left = ((int64_t)nch - ch.c0) * ldw * 4;
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<uint32_t*>(base), 0, (int)(left < 0x7fffffff ? left : 0x7fffffff), kBufferFlags);
        const float rng = 1.0f / (float)ng;
        auto stage = [&](int j0) {
            uint32_t vo[NI], sv[NI][NSRC];
            int wo[NI];
#pragma unroll
            for (int m = 0; m < NI; m++) {
                const int j = j0 + tid + NT * m;
                int ci = (int)((float)j * rng);
                int g = j - ci * ng;
                const int fix = (g >> 31) + (g >= ng ? 1 : 0);  // -1, 0 or +1, branch-free
                ci += fix;
                g -= fix * ng;
                const bool ok = j < items;
                const uint32_t r0 = (uint32_t)(t0 + dcur[ok ? ci : 0]);
                // an idle item reads at offset 2^31, past the resource: zeros
                vo[m] = ok ? 4u * ((uint32_t)ci * ldw + r0 / EPD + 4u * (uint32_t)g) : 0x80000000u;
                wo[m] = ok ? ci * rowb + 16 * g : -1;
            }
#pragma unroll
            for (int m = 0; m < NI; m++)
#pragma unroll
                for (int e = 0; e < NSRC; e++)
                    sv[m][e] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rsrc, vo[m] + 4 * e, 0, 0);
#pragma unroll
            for (int m = 0; m < NI; m++) {
                if (wo[m] < 0) continue;
#pragma unroll
                for (int p = 0; p < EPD; p++) {
                    uint4 w = make_uint4(sv[m][0], sv[m][1], sv[m][2], sv[m][3]);
                    if constexpr (ST) {  // dedisperseSTATS sums x - K (see the epilogue)
                        w.x = __float_as_uint(__uint_as_float(w.x) - kshift);
                        w.y = __float_as_uint(__uint_as_float(w.y) - kshift);
                        w.z = __float_as_uint(__uint_as_float(w.z) - kshift);
                        w.w = __float_as_uint(__uint_as_float(w.w) - kshift);
                    }
                    if constexpr (EPD > 1) {  // shifted copies (then NSRC == 5)
                        if (p > 0) {
                            w.x = __builtin_amdgcn_alignbyte(sv[m][1], sv[m][0], p * ESZ);
                            w.y = __builtin_amdgcn_alignbyte(sv[m][2], sv[m][1], p * ESZ);
                            w.z = __builtin_amdgcn_alignbyte(sv[m][3], sv[m][2], p * ESZ);
                            w.w = __builtin_amdgcn_alignbyte(sv[m][NSRC - 1], sv[m][3], p * ESZ);
                        }
                    }
                    *reinterpret_cast<uint4*>(data + wo[m] + p * ch.stride * ESZ) = w;
                }
            }
        };
        // the first pass shares its round trip with the offsets; one pass
        // covers any chunk when the item count allows it
        stage(0);
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int i = tid + NT * r;
            if (i < ch.count * DMB / 4) reinterpret_cast<u32x4*>(roff)[i] = rv[r];
        }
        if constexpr (StageItems<EPD, W>::need > StageItems<EPD, W>::value) {
            for (int j0 = NT * NI; j0 < items; j0 += NT * NI) stage(j0);
        }
        if (more && tid < nx.count) dbuf[((k + 1) & 1) * kChunkMax + tid] = dnext;
        __syncthreads();
        for (int ci = 0; ci < ch.count; ci++) {
            const uint32_t* ro = roff + ci * DMB + wave * DW;
            float sc = 1.0f;
            if constexpr (ESZ == 1) sc = chan_scale[ch.c0 + ci];
#pragma unroll
            for (int d = 0; d < DW; d++) {
                const uint32_t* p = reinterpret_cast<const uint32_t*>(data + ro[d] + 4 * lane);
#pragma unroll
                for (int q = 0; q < NDW; q++) {
                    const uint32_t w = p[64 * q];
                    if constexpr (ESZ == 4) {
                        const float v = __uint_as_float(w);
                        acc[d][q] += v;
                        if constexpr (ST) acc2[d][q] = fmaf(v, v, acc2[d][q]);
                    } else if constexpr (ESZ == 2) {
                        // one v_dot2c_f32_f16 per sample: acc += h.lo*1 + h.hi*0
                        const h2_t h = *reinterpret_cast<const h2_t*>(&w);
                        acc[d][2 * q] = __builtin_amdgcn_fdot2(h, kSelLo, acc[d][2 * q], false);
                        acc[d][2 * q + 1] = __builtin_amdgcn_fdot2(h, kSelHi, acc[d][2 * q + 1], false);
                    } else {
#pragma unroll
                        for (int e = 0; e < 4; e++)
                            acc[d][4 * q + e] = fmaf(sc, (float)((w >> (8 * e)) & 0xffu), acc[d][4 * q + e]);
                    }
                }
            }
        }
        __syncthreads();
    }
    // f32: sum / nch, the reference's final division (PIPELINE:1535);
    // f16: undo the power-of-two storage scale; u8: add the channel offsets.
    // Lane element (q, e) is sample t0 + EPD*lane + 64*EPD*q + e.
    const float fnch = (float)nch;
    float mul = 1.0f, add = 0.0f;
    if constexpr (ESZ == 2) mul = ldexpf(1.0f, f16_exp[0]);
    if constexpr (ESZ == 1) add = offsum[0];
#pragma unroll
    for (int d = 0; d < DW; d++) {
        const int dm = dm0 + d;
        if (dm >= ndm) break;
        float* o = out + (int64_t)dm * ld_out;
#pragma unroll
        for (int q = 0; q < NDW; q++) {
#pragma unroll
            for (int e = 0; e < EPD; e++) {
                const int t = t0 + EPD * lane + 64 * EPD * q + e;
                const float a = acc[d][EPD * q + e];
                if (t < nt) {
                    if constexpr (ST) o[t] = a / fnch + kshift;
                    else if constexpr (ESZ == 4) o[t] = a / fnch;
                    else if constexpr (ESZ == 2) o[t] = (a * mul) / fnch;
                    else o[t] = (a + add) / fnch;
                    if constexpr (ST) {
                        // dedisperseSTATS (TIMESERIES:1462-1477): ave = sum(x)/N in
                        // float, sum = sum(x - ave), sum2 = sum((x - ave)^2),
                        // sig = sqrt((sum2 - sum*sum) / (N(N-1))).  With the
                        // shifted sums S1 = sum(x - K), S2 = sum((x - K)^2) and
                        // d = ave - K, sum = S1 - N d and sum2 = S2 - 2 d S1 + N d^2
                        // exactly in real arithmetic; they are formed in double.
                        const double ave = (double)o[t], s1 = a, n = fnch;
                        const double dd = ave - (double)kshift;
                        const double sum = s1 - n * dd;
                        const double sum2 = (double)acc2[d][EPD * q + e] - 2.0 * dd * s1 + n * dd * dd;
                        const float num = fmaxf((float)(sum2 - sum * sum), 0.0f);
                        sig[(int64_t)dm * ld_out + t] = sqrtf(num / (float)(nch * (nch - 1)));
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Pair tiles.  Channels c and c+1 of a chunk are summed once into f32 presum
// rows, P_i[s] = x[c][S + s] + x[c+1][S + s + dlo + i], one row per relative
// delay dlo + i that the block's DMs use for the pair (off(d,c+1) - off(d,c)
// varies by a sample or two across 64 adjacent DMs).  Every DM then adds one
// presum per pair, out[d][t] += P_{delta(d)-dlo}[t - t0 + off(d,c) - S + t0]:
// half the adds of the per-channel tile, and an f32 read per sample whatever
// the window storage.  The terms are the same; only the summation tree
// differs (each pair is added first).
// ---------------------------------------------------------------------------
constexpr int kPairMax = kChunkMax / 2;  // channel pairs per chunk
constexpr int kPairMaxND = 4;            // presum rows per pair at most

// Staging items (pair, 4-sample group) per thread per pass: each holds up to
// 2 + (4 + kPairMaxND - 1)/EPD + 2 source dwords and two offsets; staging may
// take a quarter of the tile's per-lane VGPR budget (1024/W).
template <int EPD, int W>
struct PairStageItems {
    static constexpr int groups = (int)(kDataBudget / 16);  // 16 B per item and presum row
    static constexpr int need = (groups + 64 * W - 1) / (64 * W);
    static constexpr int regs = 4 + (4 + kPairMaxND - 1) + 2;  // f32 elements of both channels, offsets
    static constexpr int fit = (1024 / W / 4) / regs;
    static constexpr int cap = fit < 1 ? 1 : (fit > 8 ? 8 : fit);
    static constexpr int value = need < cap ? need : cap;
};

// K consecutive window elements of one channel as f32 (u8: times the channel
// scale), from element e of the row whose first dword is at byte voff of the
// resource.  A missing channel (voff >= 2^31) reads as zeros.
template <typename TS, int K>
__device__ __forceinline__ void load_elems(__amdgpu_buffer_rsrc_t rs, uint32_t voff, int phase, float sc,
                                           float (&v)[K]) {
    constexpr int ESZ = (int)sizeof(TS);
    constexpr int EPD = 4 / ESZ;
    constexpr int NW = (K + EPD - 1) / EPD;  // shifted dwords needed
    uint32_t dw[NW + 1];
#pragma unroll
    for (int i = 0; i < NW + 1; i++) dw[i] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rs, voff + 4 * i, 0, 0);
    if constexpr (ESZ == 4) {
#pragma unroll
        for (int i = 0; i < K; i++) v[i] = __uint_as_float(dw[i]);
    } else {
#pragma unroll
        for (int i = 0; i < NW; i++) {
            const uint32_t w = __builtin_amdgcn_alignbyte(dw[i + 1], dw[i], phase * ESZ);
#pragma unroll
            for (int e = 0; e < EPD; e++) {
                const int k = EPD * i + e;
                if (k >= K) break;
                if constexpr (ESZ == 2) {
                    const h2_t h = *reinterpret_cast<const h2_t*>(&w);
                    v[k] = (float)h[e];
                } else {
                    v[k] = sc * (float)((w >> (8 * e)) & 0xffu);
                }
            }
        }
    }
}

template <typename TS, int SL, int DW, int W>
__device__ __forceinline__ void
dedisp_pair_tile(const TS* __restrict__ win, int64_t ld, int nch, int nrows, int nt, int ndm,
                 const uint32_t* __restrict__ roffs,  // per chunk: [pairs][W*DW] LDS byte offsets
                 const int* __restrict__ dmin,        // [nblk][nch]: (start row, dlo) per pair
                 const Chunk* __restrict__ chunks, const int2* __restrict__ blk_chunks,
                 const float* __restrict__ chan_scale, const int* __restrict__ f16_exp,
                 const float* __restrict__ offsum, float* __restrict__ out, int64_t ld_out, int b, int tx,
                 unsigned char* lds_raw) {
    constexpr int ESZ = (int)sizeof(TS);
    constexpr int EPD = 4 / ESZ;
    constexpr int T = 64 * SL;
    constexpr int DMB = W * DW;
    constexpr int NT = 64 * W;
    constexpr int K2 = 4 + kPairMaxND - 1;  // elements of the second channel per group
    constexpr int NI = PairStageItems<EPD, W>::value;
    constexpr int NR = (kPairMax * DMB / 4 + NT - 1) / NT;
    uint32_t* roff = reinterpret_cast<uint32_t*>(lds_raw);
    int* dbuf = reinterpret_cast<int*>(lds_raw + (size_t)kChunkMax * DMB * 4);
    unsigned char* data = lds_raw + lds_head_bytes(DMB);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int t0 = tx * T;
    const int dm0 = b * DMB + wave * DW;
    const int2 cr = uniform_load(blk_chunks + b);
    const int* dmin_b = dmin + (int64_t)b * nch;
    const uint32_t ldw = (uint32_t)(ld / EPD);

    float acc[DW][SL];
#pragma unroll
    for (int d = 0; d < DW; d++)
#pragma unroll
        for (int j = 0; j < SL; j++) acc[d][j] = 0.0f;

    // dbuf[parity][2p] = start row of pair p's first channel, [2p+1] = dlo
    if (cr.y > 0) {
        const Chunk c0 = uniform_load(chunks + cr.x);
        if (tid < 2 * c0.count && c0.c0 + tid < nch) dbuf[tid] = dmin_b[c0.c0 + tid];
    }
    __syncthreads();

    for (int k = 0; k < cr.y; k++) {
        const Chunk ch = uniform_load(chunks + cr.x + k);
        const int np = ch.count, nd = ch.nd;
        const int rowb = nd * ch.stride * 4;  // bytes per pair (nd presum rows)
        const int ng = ch.stride / 4;         // 4-sample groups per row
        const int items = np * ng;
        const int* dcur = dbuf + (k & 1) * kChunkMax;
        int dnext = 0;
        const bool more = k + 1 < cr.y;
        Chunk nx = ch;
        if (more) {
            nx = uniform_load(chunks + cr.x + k + 1);
            if (tid < 2 * nx.count && nx.c0 + tid < nch) dnext = dmin_b[nx.c0 + tid];
        }
        const TS* base = win + (int64_t)ch.c0 * ld;
        const int64_t left = ((int64_t)nch - ch.c0) * ld * ESZ;
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<TS*>(base), 0, (int)(left < 0x7fffffff ? left : 0x7fffffff), kBufferFlags);
        const float rng = 1.0f / (float)ng;
        // Item j = (pair p, group g): presum samples 4g..4g+3 of each of the
        // pair's nd rows.  The first channel's elements e1..e1+3 and the
        // second's e1+dlo .. e1+dlo+nd+2 are loaded once (buffer loads,
        // zeros past the window or for a missing second channel) and summed
        // for every row.  Elements past a row only feed samples t >= nt.
        auto stage = [&](int j0) {
            uint32_t va[NI], vb[NI];
            int pha[NI], phb[NI], wo[NI];
            float sa[NI], sb[NI];
#pragma unroll
            for (int m = 0; m < NI; m++) {
                const int j = j0 + tid + NT * m;
                int p = (int)((float)j * rng);
                int g = j - p * ng;
                const int fix = (g >> 31) + (g >= ng ? 1 : 0);
                p += fix;
                g -= fix * ng;
                const bool ok = j < items;
                const int pp = ok ? p : 0;
                const int c1 = 2 * pp, c2 = 2 * pp + 1;
                const uint32_t e1 = (uint32_t)(t0 + dcur[c1] + 4 * g);
                const bool two = ch.c0 + c2 < nch;
                const uint32_t e2 = e1 + (uint32_t)(two ? dcur[c2] : 0);
                va[m] = ok ? 4u * ((uint32_t)c1 * ldw + e1 / EPD) : 0x80000000u;
                vb[m] = ok && two ? 4u * ((uint32_t)c2 * ldw + e2 / EPD) : 0x80000000u;
                pha[m] = (int)(e1 % EPD);
                phb[m] = (int)(e2 % EPD);
                wo[m] = ok ? pp * rowb + 16 * g : -1;
                sa[m] = 1.0f;
                sb[m] = 1.0f;
                if constexpr (ESZ == 1) {
                    sa[m] = chan_scale[ch.c0 + c1];
                    sb[m] = two ? chan_scale[ch.c0 + c2] : 0.0f;
                }
            }
            float x1[NI][4], x2[NI][K2];
#pragma unroll
            for (int m = 0; m < NI; m++) {
                load_elems<TS, 4>(rsrc, va[m], pha[m], sa[m], x1[m]);
                load_elems<TS, K2>(rsrc, vb[m], phb[m], sb[m], x2[m]);
            }
#pragma unroll
            for (int m = 0; m < NI; m++) {
                if (wo[m] < 0) continue;
#pragma unroll
                for (int i = 0; i < kPairMaxND; i++) {
                    if (i >= nd) break;
                    float4 w;
                    w.x = x1[m][0] + x2[m][i];
                    w.y = x1[m][1] + x2[m][i + 1];
                    w.z = x1[m][2] + x2[m][i + 2];
                    w.w = x1[m][3] + x2[m][i + 3];
                    *reinterpret_cast<float4*>(data + wo[m] + i * ch.stride * 4) = w;
                }
            }
        };
        stage(0);
        // the chunk's offset table, loaded after the staged groups so that
        // its registers are not live while they are
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int i = tid + NT * r;
            if (i < np * DMB / 4)
                reinterpret_cast<u32x4*>(roff)[i] = reinterpret_cast<const u32x4*>(roffs + ch.roff)[i];
        }
        if constexpr (PairStageItems<EPD, W>::need > PairStageItems<EPD, W>::value) {
            for (int j0 = NT * NI; j0 < items; j0 += NT * NI) stage(j0);
        }
        if (more && tid < 2 * nx.count && nx.c0 + tid < nch) dbuf[((k + 1) & 1) * kChunkMax + tid] = dnext;
        __syncthreads();
        for (int pi = 0; pi < np; pi++) {
            const uint32_t* ro = roff + pi * DMB + wave * DW;
#pragma unroll
            for (int d = 0; d < DW; d++) {
                const float* src = reinterpret_cast<const float*>(data + ro[d]) + lane;
#pragma unroll
                for (int q = 0; q < SL; q++) acc[d][q] += src[64 * q];
            }
        }
        __syncthreads();
    }
    // lane element q is sample t0 + lane + 64q
    const float fnch = (float)nch;
    float mul = 1.0f, add = 0.0f;
    if constexpr (ESZ == 2) mul = ldexpf(1.0f, f16_exp[0]);
    if constexpr (ESZ == 1) add = offsum[0];
#pragma unroll
    for (int d = 0; d < DW; d++) {
        const int dm = dm0 + d;
        if (dm >= ndm) break;
        float* o = out + (int64_t)dm * ld_out;
#pragma unroll
        for (int q = 0; q < SL; q++) {
            const int t = t0 + lane + 64 * q;
            const float a = acc[d][q];
            if (t < nt) {
                if constexpr (ESZ == 4) o[t] = a / fnch;
                else if constexpr (ESZ == 2) o[t] = (a * mul) / fnch;
                else o[t] = (a + add) / fnch;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Group tiles (4 channels).  The DMs of a block see a group's channels at a
// few distinct delay tuples (off(d,c+k) - off(d,c), k = 1..3; about 4 per
// group of 4 at 4096 ch x 2048 DM).  Each distinct tuple is one f32 presum
// row, R_r[s] = sum_k x[c+k][S + s + d_k], and every DM adds one presum per
// group: a quarter of the per-channel adds and half the pair tile's LDS reads
// per sample, for the same staged bytes (4 B per channel-sample).  The host
// lists a chunk's rows as records in the dmin array.
// ---------------------------------------------------------------------------
constexpr int kGroupCh = 4;
constexpr int kGroupMax = kChunkMax / kGroupCh;


// f16 group presum of 4 samples: after the phase shift, channel c's dwords
// hold elements (0,1) and (2,3).  v_perm pairs element e of two channels in
// one half2, and v_dot2c_f32_f16 with (1,1) adds both into an f32 sum:
// convert and add in one instruction per pair.
__device__ __forceinline__ float4 group_sum_f16(const uint32_t (&al)[kGroupCh][2]) {
    constexpr h2_t kOnes = {(_Float16)1.0f, (_Float16)1.0f};
    float v[4];
#pragma unroll
    for (int i = 0; i < 2; i++) {
#pragma unroll
        for (int hsel = 0; hsel < 2; hsel++) {
            const uint32_t sel = hsel ? 0x07060302u : 0x05040100u;
            const uint32_t q01 = __builtin_amdgcn_perm(al[1][i], al[0][i], sel);
            const uint32_t q23 = __builtin_amdgcn_perm(al[3][i], al[2][i], sel);
            const float s01 = __builtin_amdgcn_fdot2(*reinterpret_cast<const h2_t*>(&q01), kOnes, 0.0f, false);
            v[2 * i + hsel] = __builtin_amdgcn_fdot2(*reinterpret_cast<const h2_t*>(&q23), kOnes, s01, false);
        }
    }
    return make_float4(v[0], v[1], v[2], v[3]);
}

// Samples per time tile: 64 per lane-sample for the per-channel, pair and
// group tiles; 63 per lane-sample for the wide group tiles (MODE 3), whose
// last lane only supplies its neighbour.
__host__ __device__ constexpr int tile_samples(int sl, int mode) { return (mode == 3 ? 63 : 64) * sl; }

// Wide group tiles.  Lane l holds presum quads Q+l (and Q+63+l) of a DM's
// row and adds the DM's samples S+4l .. S+4l+3 (S = 4Q + R) of each to
// a[4k..4k+3]: elements R..3 of its own quad and 0..R-1 of lane l+1's (DPP
// wave_shl:1 on the add's first operand; lane 63 gets zeros, and its sums
// are not used).  R is the same for the whole wave.  The branch over the four
// phases sits inside one asm block: as four branches in C++ the compiler
// merges them into one add sequence whose operands it selects, or copies the
// sums in and out of each, and spills.  The s_nop gives a DPP operand written
// by a VALU instruction its two wait states.
template <int NB>
__device__ __forceinline__ void wide_add(float* a, const float4 (&v)[NB], uint32_t r) {
    static_assert(NB == 1 || NB == 2, "one or two quads per lane");
    if constexpr (NB == 2)
        asm volatile(
        "s_bitcmp1_b32 %16, 1\n"
        "s_cbranch_scc1 .Lwide_hi%=\n"
        "s_bitcmp1_b32 %16, 0\n"
        "s_cbranch_scc1 .Lwide_1%=\n"
        "v_add_f32 %0, %8, %0\n"
        "v_add_f32 %1, %9, %1\n"
        "v_add_f32 %2, %10, %2\n"
        "v_add_f32 %3, %11, %3\n"
        "v_add_f32 %4, %12, %4\n"
        "v_add_f32 %5, %13, %5\n"
        "v_add_f32 %6, %14, %6\n"
        "v_add_f32 %7, %15, %7\n"
        "s_branch .Lwide_end%=\n"
        ".Lwide_1%=:\n"
        "s_nop 1\n"
        "v_add_f32 %0, %9, %0\n"
        "v_add_f32 %1, %10, %1\n"
        "v_add_f32 %2, %11, %2\n"
        "v_add_f32_dpp %3, %8, %3 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_add_f32 %4, %13, %4\n"
        "v_add_f32 %5, %14, %5\n"
        "v_add_f32 %6, %15, %6\n"
        "v_add_f32_dpp %7, %12, %7 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "s_branch .Lwide_end%=\n"
        ".Lwide_hi%=:\n"
        "s_bitcmp1_b32 %16, 0\n"
        "s_cbranch_scc1 .Lwide_3%=\n"
        "s_nop 1\n"
        "v_add_f32 %0, %10, %0\n"
        "v_add_f32 %1, %11, %1\n"
        "v_add_f32_dpp %2, %8, %2 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_add_f32_dpp %3, %9, %3 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_add_f32 %4, %14, %4\n"
        "v_add_f32 %5, %15, %5\n"
        "v_add_f32_dpp %6, %12, %6 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_add_f32_dpp %7, %13, %7 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "s_branch .Lwide_end%=\n"
        ".Lwide_3%=:\n"
        "s_nop 1\n"
        "v_add_f32 %0, %11, %0\n"
        "v_add_f32_dpp %1, %8, %1 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_add_f32_dpp %2, %9, %2 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_add_f32_dpp %3, %10, %3 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_add_f32 %4, %15, %4\n"
        "v_add_f32_dpp %5, %12, %5 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_add_f32_dpp %6, %13, %6 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_add_f32_dpp %7, %14, %7 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        ".Lwide_end%=:\n"
        : "+v"(a[0]), "+v"(a[1]), "+v"(a[2]), "+v"(a[3]), "+v"(a[4]), "+v"(a[5]), "+v"(a[6]), "+v"(a[7])
        : "v"(v[0].x), "v"(v[0].y), "v"(v[0].z), "v"(v[0].w), "v"(v[1].x), "v"(v[1].y), "v"(v[1].z), "v"(v[1].w), "s"(r)
        : "scc");
    else
        asm volatile(
        "s_bitcmp1_b32 %8, 1\n"
        "s_cbranch_scc1 .Lwide_hi%=\n"
        "s_bitcmp1_b32 %8, 0\n"
        "s_cbranch_scc1 .Lwide_1%=\n"
        "v_add_f32 %0, %4, %0\n"
        "v_add_f32 %1, %5, %1\n"
        "v_add_f32 %2, %6, %2\n"
        "v_add_f32 %3, %7, %3\n"
        "s_branch .Lwide_end%=\n"
        ".Lwide_1%=:\n"
        "s_nop 1\n"
        "v_add_f32 %0, %5, %0\n"
        "v_add_f32 %1, %6, %1\n"
        "v_add_f32 %2, %7, %2\n"
        "v_add_f32_dpp %3, %4, %3 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "s_branch .Lwide_end%=\n"
        ".Lwide_hi%=:\n"
        "s_bitcmp1_b32 %8, 0\n"
        "s_cbranch_scc1 .Lwide_3%=\n"
        "s_nop 1\n"
        "v_add_f32 %0, %6, %0\n"
        "v_add_f32 %1, %7, %1\n"
        "v_add_f32_dpp %2, %4, %2 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_add_f32_dpp %3, %5, %3 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "s_branch .Lwide_end%=\n"
        ".Lwide_3%=:\n"
        "s_nop 1\n"
        "v_add_f32 %0, %7, %0\n"
        "v_add_f32_dpp %1, %4, %1 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_add_f32_dpp %2, %5, %2 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_add_f32_dpp %3, %6, %3 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        ".Lwide_end%=:\n"
        : "+v"(a[0]), "+v"(a[1]), "+v"(a[2]), "+v"(a[3])
        : "v"(v[0].x), "v"(v[0].y), "v"(v[0].z), "v"(v[0].w), "s"(r)
        : "scc");
}

template <typename TS, int SL, int DW, int W, bool WIDE = false>
__device__ __forceinline__ void
dedisp_group_tile(const TS* __restrict__ win, int64_t ld, int nch, int nrows, int nt, int ndm,
                  const uint32_t* __restrict__ roffs,  // per chunk: [groups][W*DW] LDS byte offsets
                  const int* __restrict__ recs_g,      // row records, kRowInts ints each
                  const Chunk* __restrict__ chunks, const int2* __restrict__ blk_chunks,
                  const float* __restrict__ chan_scale, const int* __restrict__ f16_exp,
                  const float* __restrict__ offsum, float* __restrict__ out, int64_t ld_out, int b, int tx,
                  unsigned char* lds_raw) {
    constexpr int ESZ = (int)sizeof(TS);
    constexpr int EPD = 4 / ESZ;
    constexpr int T = tile_samples(SL, WIDE ? 3 : 2);
    constexpr int DMB = W * DW;
    constexpr int NT = 64 * W;
    constexpr int NB = SL / 4;  // wide tiles: presum quads per lane and DM
    static_assert(!WIDE || SL % 4 == 0, "wide tiles read whole quads");
    constexpr int OPW = WIDE ? 2 : 1;  // DM offsets per table word (wide tiles: two 16-bit halves)
    static_assert(!WIDE || (DW % 2 == 0 && DMB % 8 == 0), "wide tiles pair the DMs of a wave");
    static_assert(kDataBudget < 65536, "16-bit LDS offsets");
    constexpr int NR = (kGroupMax * DMB / (4 * OPW) + NT - 1) / NT;
    constexpr int NQ = (kRowMax * kRowInts + NT - 1) / NT;  // record ints per thread
    uint32_t* roff = reinterpret_cast<uint32_t*>(lds_raw);
    int* rbuf = reinterpret_cast<int*>(lds_raw + (size_t)kChunkMax * DMB * 4);
    unsigned char* data = lds_raw + lds_head_bytes(DMB);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int t0 = tx * T;
    const int dm0 = b * DMB + wave * DW;
    const int2 cr = uniform_load(blk_chunks + b);

    float acc[DW][SL];
#pragma unroll
    for (int d = 0; d < DW; d++)
#pragma unroll
        for (int j = 0; j < SL; j++) acc[d][j] = 0.0f;

    if (cr.y > 0) {
        const Chunk c0 = uniform_load(chunks + cr.x);
        for (int i = tid; i < c0.nd * kRowInts; i += NT) rbuf[i] = recs_g[c0.pad[0] + i];
    }
    __syncthreads();

    for (int k = 0; k < cr.y; k++) {
        const Chunk ch = uniform_load(chunks + cr.x + k);
        const int ngr = ch.count, nrw = ch.nd;
        const int ng = ch.stride / 4;  // 4-sample groups per row
        const int* rcur = rbuf + (k & 1) * kRowMax * kRowInts;
        const bool more = k + 1 < cr.y;
        Chunk nx = ch;
        int qn[NQ];
        if (more) {
            nx = uniform_load(chunks + cr.x + k + 1);
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int i = tid + NT * q;
                qn[q] = i < nx.nd * kRowInts ? recs_g[nx.pad[0] + i] : 0;
            }
        }
        const TS* base = win + (int64_t)ch.c0 * ld;
        const int64_t left = ((int64_t)nch - ch.c0) * ld * ESZ;
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<TS*>(base), 0, (int)(left < 0x7fffffff ? left : 0x7fffffff), kBufferFlags);
        const uint32_t toff = 4u * (uint32_t)(t0 / EPD);  // t0 is a multiple of EPD
        // Item j = (row r, 4-sample group g), j = r*ng + g: the row's four
        // channels' elements S + d_k + 4g .. +3 (buffer loads; zeros past the
        // window or for a channel the last group lacks) summed into presum
        // samples 4g..4g+3.  The row record holds each channel's byte offset
        // and phase.  A thread takes items tid, tid + NT, ...: (r, g) advance
        // by (NT / ng, NT mod ng) with one carry.
        int rq = NT / ng, rr = NT - rq * ng, ngv = ng;
        // Wide f32/f16 tiles keep these in vector registers for the loop: its
        // carry step would otherwise copy them from scalar registers on every
        // item (one scalar operand per vector instruction).  The other tiles
        // have no three registers to spare.
        if constexpr (WIDE && ESZ != 1) asm volatile("" : "+v"(rq), "+v"(rr), "+v"(ngv));
        int r = (int)((float)tid * (1.0f / (float)ng));
        int g = tid - r * ng;
        {
            const int fix = (g >> 31) + (g >= ng ? 1 : 0);
            r += fix;
            g -= fix * ng;
        }
        for (int j0 = 0; j0 < nrw * ng; j0 += NT) {
            // A thread past the last row: wide tiles skip its item; the
            // others (fewer registers to spare) load row 0's elements and drop
            // them, which keeps their staging branch-free.
            const bool ok = r < nrw;
            if (!WIDE || ok) {
                const int* rec = rcur + (WIDE || ok ? r : 0) * kRowInts;
                const int4 rbase = *reinterpret_cast<const int4*>(rec);
                const int4 rmeta = *reinterpret_cast<const int4*>(rec + 4);
                const int bases[kGroupCh] = {rbase.x, rbase.y, rbase.z, rbase.w};
                const uint32_t lo = toff + (uint32_t)(16 * g / EPD);
                uint32_t vo[kGroupCh];
                int ph[kGroupCh];
                float sc[kGroupCh];
#pragma unroll
                for (int c = 0; c < kGroupCh; c++) {
                    vo[c] = (uint32_t)bases[c] + lo;
                    ph[c] = (rmeta.x >> (8 * c)) & 3;
                    sc[c] = 1.0f;
                    if constexpr (ESZ == 1) sc[c] = bases[c] >= 0 ? chan_scale[ch.c0 + rmeta.z + c] : 0.0f;
                }
                unsigned char* dst = data + rmeta.y + 16 * g;
                if constexpr (ESZ == 2) {
                    // f16: three raw dwords per channel, shifted by the phase and
                    // summed by group_sum_f16
                    uint32_t raw[kGroupCh][3];
#pragma unroll
                    for (int c = 0; c < kGroupCh; c++)
#pragma unroll
                        for (int i = 0; i < 3; i++)
                            raw[c][i] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rsrc, vo[c] + 4 * i, 0, 0);
                    // all twelve loads in flight before the first use: the
                    // scheduler otherwise waits for each channel in turn
                    __builtin_amdgcn_sched_barrier(0);
                    uint32_t al[kGroupCh][2];
#pragma unroll
                    for (int c = 0; c < kGroupCh; c++)
#pragma unroll
                        for (int i = 0; i < 2; i++) al[c][i] = __builtin_amdgcn_alignbyte(raw[c][i + 1], raw[c][i], 2 * ph[c]);
                    if (WIDE || ok) *reinterpret_cast<float4*>(dst) = group_sum_f16(al);
                } else {
                    // f32: four dwords per channel; u8: two, shifted by the phase
                    // and scaled (the elements of load_elems<TS, 4>), all loads
                    // in flight before the first use
                    constexpr int NW = ESZ == 4 ? 4 : 2;
                    uint32_t raw[kGroupCh][NW];
#pragma unroll
                    for (int c = 0; c < kGroupCh; c++)
#pragma unroll
                        for (int i = 0; i < NW; i++)
                            raw[c][i] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rsrc, vo[c] + 4 * i, 0, 0);
                    __builtin_amdgcn_sched_barrier(0);
                    float x[kGroupCh][4];
#pragma unroll
                    for (int c = 0; c < kGroupCh; c++) {
                        if constexpr (ESZ == 4) {
#pragma unroll
                            for (int e = 0; e < 4; e++) x[c][e] = __uint_as_float(raw[c][e]);
                        } else {
                            const uint32_t wd = __builtin_amdgcn_alignbyte(raw[c][1], raw[c][0], ph[c]);
#pragma unroll
                            for (int e = 0; e < 4; e++) x[c][e] = sc[c] * (float)((wd >> (8 * e)) & 0xffu);
                        }
                    }
                    float4 w;
                    w.x = (x[0][0] + x[1][0]) + (x[2][0] + x[3][0]);
                    w.y = (x[0][1] + x[1][1]) + (x[2][1] + x[3][1]);
                    w.z = (x[0][2] + x[1][2]) + (x[2][2] + x[3][2]);
                    w.w = (x[0][3] + x[1][3]) + (x[2][3] + x[3][3]);
                    if (WIDE || ok) *reinterpret_cast<float4*>(dst) = w;
                }
            }
            r += rq;
            g += rr;
            if (g >= ngv) {
                g -= ngv;
                r++;
            }
        }
        // the chunk's offset table, and the next chunk's row records
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int i = tid + NT * r;
            if (i < ngr * DMB / (4 * OPW))
                reinterpret_cast<u32x4*>(roff)[i] = reinterpret_cast<const u32x4*>(roffs + ch.roff)[i];
        }
        if (more) {
            int* rnext = rbuf + ((k + 1) & 1) * kRowMax * kRowInts;
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int i = tid + NT * q;
                if (i < nx.nd * kRowInts) rnext[i] = qn[q];
            }
        }
        __syncthreads();
        if constexpr (WIDE) {
            // Lane l reads quads Q+l and Q+63+l of a DM's row (Q = its start
            // sample / 4) with 16-byte LDS reads, which run at nearly twice
            // the rate of the two-sample reads below; wide_add picks the add
            // sequence of the start's phase.  Lane 63's sums are not stored.
            const uint32_t* ro = roff + wave * DW / 2;
            for (int gi = 0; gi < ngr; gi++) {
                uint32_t o[DW];
#pragma unroll
                for (int d = 0; d < DW; d += 2) {
                    const uint32_t w = __builtin_amdgcn_readfirstlane(ro[gi * DMB / 2 + d / 2]);
                    o[d] = w & 0xffffu;
                    o[d + 1] = w >> 16;
                }
                // the next DM's quads are read before this one's are added
                float4 v[2][NB];
                auto quads = [&](int d) {
                    const unsigned char* src = data + (o[d] & ~15u) + 16 * lane;
#pragma unroll
                    for (int k = 0; k < NB; k++) v[d & 1][k] = *reinterpret_cast<const float4*>(src + 16 * 63 * k);
                };
                quads(0);
#pragma unroll
                for (int d = 0; d < DW; d++) {
                    if (d + 1 < DW) quads(d + 1);
                    wide_add<NB>(acc[d], v[d & 1], (o[d] >> 2) & 3u);
                }
            }
        } else if constexpr (W <= 8 && DW % 2 == 0) {
            // Tiles of 8 waves (4 per SIMD) keep two DMs' presum samples in
            // flight before adding them: the compiler otherwise reads, waits
            // and adds one pair of samples at a time, which these few waves
            // cannot hide (G512_dm64_w8 6.0 -> 4.9 ms at 4096 ch x 2048 DM).
            // 16-wave tiles hide it with their waves and keep the plain loop,
            // which their 64 VGPRs hold without the buffer.
            for (int gi = 0; gi < ngr; gi++) {
                const uint32_t* ro = roff + gi * DMB + wave * DW;
#pragma unroll
                for (int d0 = 0; d0 < DW; d0 += 2) {
                    float v[2][SL];
#pragma unroll
                    for (int e = 0; e < 2; e++) {
                        const float* src = reinterpret_cast<const float*>(data + ro[d0 + e]) + lane;
#pragma unroll
                        for (int q = 0; q < SL; q++) v[e][q] = src[64 * q];
                    }
#pragma unroll
                    for (int e = 0; e < 2; e++)
#pragma unroll
                        for (int q = 0; q < SL; q++) acc[d0 + e][q] += v[e][q];
                }
            }
        } else {
            for (int gi = 0; gi < ngr; gi++) {
                const uint32_t* ro = roff + gi * DMB + wave * DW;
#pragma unroll
                for (int d = 0; d < DW; d++) {
                    const float* src = reinterpret_cast<const float*>(data + ro[d]) + lane;
#pragma unroll
                    for (int q = 0; q < SL; q++) acc[d][q] += src[64 * q];
                }
            }
        }
        __syncthreads();
    }
    const float fnch = (float)nch;
    float mul = 1.0f, add = 0.0f;
    if constexpr (ESZ == 2) mul = ldexpf(1.0f, f16_exp[0]);
    if constexpr (ESZ == 1) add = offsum[0];
#pragma unroll
    for (int d = 0; d < DW; d++) {
        const int dm = dm0 + d;
        if (dm >= ndm) break;
        float* o = out + (int64_t)dm * ld_out;
#pragma unroll
        for (int q = 0; q < SL; q++) {
            const int t = WIDE ? t0 + 252 * (q / 4) + 4 * lane + (q & 3) : t0 + lane + 64 * q;
            const float a = acc[d][q];
            if (t < nt && (!WIDE || lane < 63)) {
                if constexpr (ESZ == 4) o[t] = a / fnch;
                else if constexpr (ESZ == 2) o[t] = (a * mul) / fnch;
                else o[t] = (a + add) / fnch;
            }
        }
    }
}

// Workgroups are dealt round-robin to the 8 XCDs, each with its own L2.
// Remap the 1-D grid so that every XCD gets a contiguous range of tile ids
// (ids 0..n-1 in XCD-major order), and order the ids DM-block fastest: the
// workgroups an XCD runs together then share a time tile, and the window
// rows they stage (the same rows for all DM blocks wherever the delays are
// small) come from that XCD's L2 instead of HBM.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int bid, int n) {
    const int q = n / kXcds, r = n % kXcds;
    const int xcd = bid % kXcds, slot = bid / kXcds;
    return xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// One workgroup per tile: 1-D grid of nblk * (time tiles) workgroups.  MODE
// 0: per-channel tile; 1: channel-pair presums; 2: 4-channel group presums.
template <typename TS, int SL, int DW, int W, int MODE>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(WavesPerEu<W>::value, WavesPerEu<W>::value)))
dedisp_kernel(const TS* __restrict__ win, int64_t ld, int nch, int nrows, int nt, int ndm,
              const uint32_t* __restrict__ roffs, const int* __restrict__ dmin,
              const Chunk* __restrict__ chunks, const int2* __restrict__ blk_chunks,
              const float* __restrict__ chan_scale, const int* __restrict__ f16_exp,
              const float* __restrict__ offsum, float* __restrict__ out, int64_t ld_out, int nblk) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const int id = xcd_tile(blockIdx.x, gridDim.x);
    const int b = id % nblk, tx = id / nblk;
    if constexpr (MODE >= 2)
        dedisp_group_tile<TS, SL, DW, W, MODE == 3>(win, ld, nch, nrows, nt, ndm, roffs, dmin, chunks, blk_chunks,
                                         chan_scale, f16_exp, offsum, out, ld_out, b, tx, lds_raw);
    else if constexpr (MODE == 1)
        dedisp_pair_tile<TS, SL, DW, W>(win, ld, nch, nrows, nt, ndm, roffs, dmin, chunks, blk_chunks,
                                        chan_scale, f16_exp, offsum, out, ld_out, b, tx, lds_raw);
    else
        dedisp_tile<TS, SL, DW, W>(win, ld, nch, nrows, nt, ndm, roffs, dmin, chunks, blk_chunks,
                                   chan_scale, f16_exp, offsum, out, ld_out, b, tx, lds_raw);
}

// ---------------------------------------------------------------------------
// Batched dedispersion: one persistent launch over the tiles of many windows
// (jobs), each with its own DM list, length and storage scales.  The work
// list lives in HBM, so a captured launch can be replayed on new jobs.
// ---------------------------------------------------------------------------
struct BatchJob {
    const void* win;
    int64_t ld;
    int nch, nrows, nt, ndm;
    const uint32_t* roffs;
    const int* dmin;
    const Chunk* chunks;
    const int2* blk_chunks;
    const float* chan_scale;
    const int* f16_exp;
    const float* offsum;
    float* out;
    int64_t ld_out;
};

struct BatchTile {
    int job, blk, tx, pad;
};

// ctr[0]: next tile, ctr[1]: workgroups done.  The last workgroup to finish
// puts both back to zero, so the next launch (or graph replay) starts clean
// without a memset; ntiles[0] = tiles in the list.
template <typename TS, int SL, int DW, int W, int MODE>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(WavesPerEu<W>::value, WavesPerEu<W>::value)))
dedisp_batch_kernel(const BatchJob* __restrict__ jobs, const BatchTile* __restrict__ tiles,
                    const int* __restrict__ ntiles, unsigned* __restrict__ ctr) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    __shared__ int s_tile;
    const int total = ntiles[0];
    for (;;) {
        if (threadIdx.x == 0) s_tile = (int)atomicAdd(&ctr[0], 1u);
        __syncthreads();
        const int i = __builtin_amdgcn_readfirstlane(s_tile);
        __syncthreads();
        if (i >= total) break;
        const BatchTile tl = tiles[i];
        const BatchJob& j = jobs[tl.job];
        if constexpr (MODE >= 2)
            dedisp_group_tile<TS, SL, DW, W, MODE == 3>(static_cast<const TS*>(j.win), j.ld, j.nch, j.nrows, j.nt, j.ndm,
                                             j.roffs, j.dmin, j.chunks, j.blk_chunks, j.chan_scale, j.f16_exp,
                                             j.offsum, j.out, j.ld_out, tl.blk, tl.tx, lds_raw);
        else if constexpr (MODE == 1)
            dedisp_pair_tile<TS, SL, DW, W>(static_cast<const TS*>(j.win), j.ld, j.nch, j.nrows, j.nt, j.ndm,
                                            j.roffs, j.dmin, j.chunks, j.blk_chunks, j.chan_scale, j.f16_exp,
                                            j.offsum, j.out, j.ld_out, tl.blk, tl.tx, lds_raw);
        else
            dedisp_tile<TS, SL, DW, W>(static_cast<const TS*>(j.win), j.ld, j.nch, j.nrows, j.nt, j.ndm,
                                       j.roffs, j.dmin, j.chunks, j.blk_chunks, j.chan_scale, j.f16_exp,
                                       j.offsum, j.out, j.ld_out, tl.blk, tl.tx, lds_raw);
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&ctr[1], 1u) == gridDim.x - 1) {
            atomicExch(&ctr[0], 0u);
            atomicExch(&ctr[1], 0u);
        }
    }
}

// dedisperseSTATS (TIMESERIES:1450-1481) on the per-channel tile of an f32
// window: one workgroup per (DM block, time tile), out = mean over channels,
// sig = their spread, both [ndm][ld_out].
template <int SL, int DW, int W>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(WavesPerEu<W>::value, WavesPerEu<W>::value)))
dedisp_stats_kernel(const float* __restrict__ win, int64_t ld, int nch, int nrows, int nt, int ndm,
                    const uint32_t* __restrict__ roffs, const int* __restrict__ dmin,
                    const Chunk* __restrict__ chunks, const int2* __restrict__ blk_chunks,
                    float* __restrict__ out, float* __restrict__ sig, int64_t ld_out, int nblk,
                    const float* __restrict__ shift) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const int id = xcd_tile(blockIdx.x, gridDim.x);
    const int b = id % nblk, tx = id / nblk;
    dedisp_tile<float, SL, DW, W, true>(win, ld, nch, nrows, nt, ndm, roffs, dmin, chunks, blk_chunks, nullptr,
                                        nullptr, nullptr, out, ld_out, b, tx, lds_raw, sig, shift);
}

// The dedisperseSTATS shift K of a channel-major f32 window, from three rows
// (first, middle, last) of every channel: their mean m rounded to 12
// significant bits when |m| exceeds twice their spread, else 0.  A bandpassed
// (zero-mean) window keeps K = 0, i.e. plain sums; a window with a large
// level gets sums of x - K that do not cancel.
__global__ void __launch_bounds__(256) stats_shift_kernel(const float* __restrict__ win, int64_t ld, int nch,
                                                          int nrows, float* __restrict__ shift) {
    __shared__ double s1[256], s2[256];
    double a = 0.0, b = 0.0;
    const int64_t rows[3] = {0, nrows / 2, nrows - 1};
    for (int c = threadIdx.x; c < nch; c += 256)
        for (int k = 0; k < 3; k++) {
            const double x = win[(int64_t)c * ld + rows[k]];
            a += x;
            b += x * x;
        }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            s1[threadIdx.x] += s1[threadIdx.x + h];
            s2[threadIdx.x] += s2[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double n = 3.0 * nch, m = s1[0] / n;
        const double sd = sqrt(fmax(s2[0] / n - m * m, 0.0));
        float k = 0.0f;
        if (isfinite(m) && fabs(m) > 2.0 * sd && m != 0.0) {
            const int e = ilogb(m);
            k = (float)ldexp(rint(ldexp(m, 11 - e)), e - 11);
        }
        shift[0] = k;
    }
}

// ---------------------------------------------------------------------------
// Tile configurations
// ---------------------------------------------------------------------------
struct TileCfg {
    int sl, dw, w;
    int mode;  // 0 per channel; 1 channel-pair presums ("P"); 2 4-channel group presums ("G");
               // 3 group presums read 16 bytes per lane ("W")
    const char* name;
};
// samples per lane (multiple of 4 so u8 lanes read whole dwords), DMs per
// wave, waves per workgroup.  Tile = 64*SL samples (W tiles: 63*SL) x W*DW
// DMs.  "P" tiles sum channel pairs first, "G" and "W" tiles groups of 4
// channels.
constexpr TileCfg kTiles[] = {
    {4, 16, 4, 0, "T256_dm64_w4"},    {8, 16, 4, 0, "T512_dm64_w4"},
    {8, 8, 4, 0, "T512_dm32_w4"},     {8, 8, 8, 0, "T512_dm64_w8"},
    {4, 8, 8, 0, "T256_dm64_w8"},     {8, 4, 8, 0, "T512_dm32_w8"},
    {16, 4, 8, 0, "T1024_dm32_w8"},   {4, 16, 8, 0, "T256_dm128_w8"},
    {8, 4, 16, 0, "T512_dm64_w16"},   {4, 4, 16, 0, "T256_dm64_w16"},
    {4, 8, 16, 0, "T256_dm128_w16"},  {16, 8, 4, 0, "T1024_dm32_w4"},
    {8, 4, 16, 1, "P512_dm64_w16"},   {8, 8, 8, 1, "P512_dm64_w8"},
    {4, 8, 16, 1, "P256_dm128_w16"},  {16, 4, 8, 1, "P1024_dm32_w8"},
    {8, 4, 16, 2, "G512_dm64_w16"},   {8, 8, 8, 2, "G512_dm64_w8"},
    {4, 8, 16, 2, "G256_dm128_w16"},  {4, 4, 16, 2, "G256_dm64_w16"},
    {8, 4, 16, 3, "W504_dm64_w16"},   {8, 8, 8, 3, "W504_dm64_w8"},
    {4, 8, 16, 3, "W252_dm128_w16"},  {4, 4, 16, 3, "W252_dm64_w16"},
    {4, 4, 8, 2, "G256_dm32_w8"},     {4, 2, 8, 2, "G256_dm16_w8"},
};
constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);
// Default tile: the 4-channel group tile G256_dm64_w16.  Group and pair tiles
// need the block's DMs to see each group (pair) at few relative delays; for
// DM lists too coarse for them the default falls back along kTileFallback:
// the pair tile, then the per-channel T512_dm64_w16, which fits any list
// that a per-channel tile of 512 samples x 64 DMs fits (T512_dm64_w8 stages
// the same rows, 14.5 against 11.0 ms in f32 at 4096 ch x 2048 DMs).
constexpr int kDefaultTile = 19;                // G256_dm64_w16
constexpr int kTileFallback[] = {19, 12, 8};    // G256_dm64_w16, P512_dm64_w16, T512_dm64_w16
// dedisperseSTATS tile: per-channel, two accumulators per sample, so half the
// DMs per wave of T512_dm64_w8 within the same 128 VGPRs (fastest of
// 256x64, 512x32 and 1024x16 samples x DMs at 4096 ch x 2048 DMs).
constexpr TileCfg kStatsTile = {8, 4, 8, 0, "S512_dm32_w8"};

template <typename TS, int I>
const void* tile_kernel_ptr(bool batch) {
    constexpr TileCfg c = kTiles[I];
    return batch ? reinterpret_cast<const void*>(&dedisp_batch_kernel<TS, c.sl, c.dw, c.w, c.mode>)
                 : reinterpret_cast<const void*>(&dedisp_kernel<TS, c.sl, c.dw, c.w, c.mode>);
}
template <typename TS, size_t... I>
const void* tile_kernel_at(int i, bool batch, std::index_sequence<I...>) {
    const void* k[] = {tile_kernel_ptr<TS, (int)I>(batch)...};
    return k[i];
}
template <typename TS>
const void* tile_kernel(int i, bool batch) {
    return tile_kernel_at<TS>(i, batch, std::make_index_sequence<kNumTiles>{});
}
const void* tile_kernel_for(int storage, int i, bool batch) {
    if (storage == VFRB_STORE_F32) return tile_kernel<float>(i, batch);
    if (storage == VFRB_STORE_F16) return tile_kernel<__half>(i, batch);
    return tile_kernel<uint8_t>(i, batch);
}

size_t storage_bytes(int storage) {
    return storage == VFRB_STORE_F32 ? 4 : storage == VFRB_STORE_F16 ? 2 : 1;
}

// Device tables of one tile configuration over a plan's DM list.
struct TileSet {
    int tile = -1;
    uint32_t* d_roffs = nullptr;
    int* d_dmin = nullptr;
    Chunk* d_chunks = nullptr;
    int2* d_blk_chunks = nullptr;
    int nblk = 0;
    size_t lds_bytes = 0;
    // estimated lane-operations per tile of each DM block (batch ordering)
    std::vector<int64_t> blk_cost;
};

}  // namespace

struct vfrb_search_plan {
    vfrb_search_desc desc{};
    std::vector<float> dms;
    std::vector<int32_t> off;  // host [ndm][nch]
    int32_t max_delay = 0;
    int nch = 0, ndm = 0, storage = 0;
    int64_t ld = 0;            // window row length (elements)
    // device
    int32_t* d_off = nullptr;
    void* d_win = nullptr;
    float* d_bp_mean = nullptr;
    unsigned* d_range = nullptr;
    int* d_f16_exp = nullptr;
    float* d_u8_scale = nullptr;
    float* d_u8_off = nullptr;
    float* d_u8_offsum = nullptr;
    float* d_tmp_fb = nullptr;  // time-major staging for vfrb_search_execute
    float* d_shift = nullptr;   // dedisperseSTATS window shift, allocated with its tables
    TileSet ts;   // tables of the dedispersion tile in use
    TileSet sts;  // tables of the dedisperseSTATS tile, built on first use
    int loaded_rows = 0;
};

struct vfrb_batch {
    int storage = 0, tile = -1;
    bool tile_default = false;  // created with tile_config -1: may fall back
    int pref_tile = -1;         // default-tiled: tried first on every job set (default or autotuned)
    int max_jobs = 0, max_tiles = 0;
    int ntiles = -1;  // -1: no jobs set
    std::vector<vfrb_search_plan*> plans;
    std::vector<int32_t> nrows;
    std::vector<float*> outs;
    std::vector<int64_t> ld_outs;
    // pinned staging of the work list and its device copy
    BatchJob* h_jobs = nullptr;
    BatchTile* h_tiles = nullptr;
    int* h_ntiles = nullptr;
    BatchJob* d_jobs = nullptr;
    BatchTile* d_tiles = nullptr;
    int* d_ntiles = nullptr;
    unsigned* d_ctr = nullptr;
    hipEvent_t staged = nullptr;  // recorded after the staging upload
    bool staged_pending = false;
    int grid = 0;
    size_t lds_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void free_tiles(TileSet& t) {
    (void)hipFree(t.d_roffs);
    (void)hipFree(t.d_dmin);
    (void)hipFree(t.d_chunks);
    (void)hipFree(t.d_blk_chunks);
    t = TileSet{};
}

// Builds the per-DM-block tables for tile configuration i: minimum delay per
// channel, channel chunks whose staged rows fit the LDS budget, and for every
// (chunk channel, DM of the block) the LDS byte offset of the first dword the
// kernel reads: channel row + copy (delay mod samples-per-dword) + aligned
// position of the delay inside that copy.
vfrb_status build_tile_set(const vfrb_search_plan* p, const TileCfg& tc, int i, TileSet& out) {
    const int T = tile_samples(tc.sl, tc.mode), DMB = tc.w * tc.dw;
    const int nch = p->nch, ndm = p->ndm;
    const int nblk = (ndm + DMB - 1) / DMB;
    const int esz = (int)storage_bytes(p->storage);
    const int epd = 4 / esz;
    size_t max_data = 0;
    std::vector<int> dmin((size_t)nblk * nch, 0);
    std::vector<int> recs;  // group tiles: row records, uploaded in place of dmin
    std::vector<Chunk> chunks;
    std::vector<uint32_t> roffs;
    std::vector<int2> blk(nblk);
    std::vector<int64_t> cost(nblk, 0);
    // A staged channel takes 4*stride bytes in every storage type: one f32
    // copy, two shifted f16 copies or four shifted u8 copies.  A copy starts
    // at the dword-aligned element below the first one needed (phase up to
    // epd-1) and a read dword reaches epd-1 elements past its sample, so a
    // copy holds T + span + 2*epd elements, rounded to 16 (16-byte rows).
    auto stride_of = [&](int s) { return (T + s + 2 * epd + 15) / 16 * 16; };
    for (int b = 0; b < nblk; b++) {
        const int d0 = b * DMB, d1 = std::min(ndm, d0 + DMB);
        std::vector<int> span(nch);
        for (int c = 0; c < nch; c++) {
            int lo = INT32_MAX, hi = INT32_MIN;
            for (int d = d0; d < d1; d++) {
                lo = std::min(lo, p->off[(size_t)d * nch + c]);
                hi = std::max(hi, p->off[(size_t)d * nch + c]);
            }
            dmin[(size_t)b * nch + c] = lo;
            span[c] = hi - lo;
        }
        blk[b].x = (int)chunks.size();
        if (tc.mode >= 2) {
            // Group q = channels 4q..4q+3 (fewer in the last group).  Each
            // distinct tuple of relative delays of the group's channels over
            // the block's DMs is one presum row; a chunk's rows are listed as
            // records (kRowInts ints) appended to recs.
            const int ngrp = (nch + kGroupCh - 1) / kGroupCh;
            std::vector<std::vector<std::array<int, 3>>> tup(ngrp);
            std::vector<std::vector<int>> rowof(ngrp);
            for (int g = 0; g < ngrp; g++) {
                const int c0 = kGroupCh * g, nchg = std::min(kGroupCh, nch - c0);
                for (int d = d0; d < d1; d++) {
                    std::array<int, 3> t{0, 0, 0};
                    for (int k = 1; k < nchg; k++)
                        t[k - 1] = p->off[(size_t)d * nch + c0 + k] - p->off[(size_t)d * nch + c0];
                    int idx = -1;
                    for (int r = 0; r < (int)tup[g].size() && idx < 0; r++)
                        if (tup[g][r] == t) idx = r;
                    if (idx < 0) {
                        idx = (int)tup[g].size();
                        tup[g].push_back(t);
                    }
                    rowof[g].push_back(idx);
                }
            }
            // A row covers T + span samples; a wide tile's lane 63 reads up
            // to 3 samples further (the quad after the last one used).
            const int extra = tc.mode == 3 ? 4 : 0;
            auto gstride = [&](int sp) { return (T + sp + extra + 3) / 4 * 4; };
            int g0 = 0;
            while (g0 < ngrp) {
                int count = 0, rows = 0, smax = 0;
                while (g0 + count < ngrp && count < kGroupMax) {
                    const int r2 = rows + (int)tup[g0 + count].size();
                    const int s2 = std::max(smax, span[kGroupCh * (g0 + count)]);
                    if (r2 > kRowMax || (size_t)r2 * 4 * (size_t)gstride(s2) > kDataBudget) break;
                    rows = r2;
                    smax = s2;
                    count++;
                }
                VFRB_REQUIRE(count > 0, VFRB_ERR_UNSUPPORTED,
                             "one channel group's presum rows exceed the LDS budget; use a tile with fewer DMs "
                             "per block");
                const int stride = gstride(smax);
                chunks.push_back(Chunk{kGroupCh * g0, count, stride, (int)roffs.size(), rows,
                                       {(int)recs.size(), 0, 0}});
                int rowstart = 0;
                for (int q = 0; q < count; q++) {
                    const int g = g0 + q, c0 = kGroupCh * g, nchg = std::min(kGroupCh, nch - c0);
                    const int amin = dmin[(size_t)b * nch + c0];
                    // One record per presum row: the byte offset of each channel's
                    // first dword from the chunk's first channel row at time tile 0
                    // (0xC0000000, past any window, for a missing channel), the
                    // element phases (a byte each), the row's LDS byte offset and
                    // the group's first channel in the chunk.  A tile adds only its
                    // time offset and the lane's sample offset.
                    for (size_t ri = 0; ri < tup[g].size(); ri++) {
                        const auto& t = tup[g][ri];
                        int rec[kRowInts] = {0, 0, 0, 0, 0, (rowstart + (int)ri) * stride * 4, kGroupCh * q, 0};
                        for (int c = 0; c < kGroupCh; c++) {
                            const int64_t e = (int64_t)amin + (c == 0 ? 0 : t[c - 1]);
                            const int64_t bo = 4 * ((int64_t)(kGroupCh * q + c) * (p->ld / epd) + e / epd);
                            // bases and the lane's offset (time tile + sample group, below
                            // (ld + stride) elements) stay below 2^30 each, so a missing
                            // channel's offset never wraps into the window
                            VFRB_REQUIRE(bo < 0x40000000LL && ((int64_t)p->ld + stride) * esz < 0x40000000LL,
                                         VFRB_ERR_UNSUPPORTED, "channel row offset %lld too large for a chunk",
                                         (long long)bo);
                            rec[c] = c < nchg ? (int)bo : (int)0xC0000000u;
                            rec[4] |= (int)(e % epd) << (8 * c);
                        }
                        recs.insert(recs.end(), rec, rec + kRowInts);
                    }
                    // wide tiles: two DMs' offsets per word (low half the even DM),
                    // each below the 64 KB the data budget stays under
                    for (int d = 0; d < DMB; d++) {
                        uint32_t o = 0;
                        if (d0 + d < d1) {
                            const int a = p->off[(size_t)(d0 + d) * nch + c0] - amin;
                            o = (uint32_t)(((rowstart + rowof[g][d]) * stride + a) * 4);
                        }
                        if (tc.mode == 3) {
                            if (d % 2 == 0) roffs.push_back(o);
                            else roffs.back() |= o << 16;
                        } else {
                            roffs.push_back(o);
                        }
                    }
                    rowstart += (int)tup[g].size();
                }
                max_data = std::max(max_data, (size_t)rows * 4 * stride);
                // accumulate (one add per group) + staging (one presum per row sample)
                cost[b] += (int64_t)count * DMB * T + 4LL * stride * rows;
                g0 += count;
            }
            blk[b].y = (int)chunks.size() - blk[b].x;
            continue;
        }
        if (tc.mode == 1) {
            // Pair p = channels 2p, 2p+1 (the last alone when nch is odd).
            // dmin keeps the first channel's start row; the second channel's
            // slot holds the pair's lowest relative delay dlo.  A chunk's
            // pairs share the row count nd and the stride, a multiple of 4
            // f32 samples covering T + span.
            const int npair = (nch + 1) / 2;
            std::vector<int> dlo(npair, 0), nd(npair, 1);
            for (int pp = 0; pp < npair; pp++) {
                const int c1 = 2 * pp, c2 = c1 + 1;
                if (c2 >= nch) continue;
                int lo = INT32_MAX, hi = INT32_MIN;
                for (int d = d0; d < d1; d++) {
                    const int dl = p->off[(size_t)d * nch + c2] - p->off[(size_t)d * nch + c1];
                    lo = std::min(lo, dl);
                    hi = std::max(hi, dl);
                }
                dlo[pp] = lo;
                nd[pp] = hi - lo + 1;
                VFRB_REQUIRE(nd[pp] <= kPairMaxND, VFRB_ERR_UNSUPPORTED,
                             "pair tile: the relative delay of channels %d,%d spans %d rows across a DM "
                             "block (at most %d)", c1, c2, nd[pp], kPairMaxND);
            }
            for (int pp = 0; pp < npair; pp++)
                if (2 * pp + 1 < nch) dmin[(size_t)b * nch + 2 * pp + 1] = dlo[pp];
            auto pstride = [&](int sp) { return (T + sp + 3) / 4 * 4; };
            int pp = 0;
            while (pp < npair) {
                int count = 0, smax = 0, nmax = 1;
                while (pp + count < npair && count < kPairMax) {
                    const int s2 = std::max(smax, span[2 * (pp + count)]);
                    const int n2 = std::max(nmax, nd[pp + count]);
                    if ((size_t)(count + 1) * n2 * 4 * (size_t)pstride(s2) > kDataBudget) break;
                    smax = s2;
                    nmax = n2;
                    count++;
                }
                VFRB_REQUIRE(count > 0, VFRB_ERR_UNSUPPORTED,
                             "one channel pair's presum rows exceed the LDS budget; use a tile with fewer "
                             "DMs per block");
                const int stride = pstride(smax);
                chunks.push_back(Chunk{2 * pp, count, stride, (int)roffs.size(), nmax, {0, 0, 0}});
                for (int pi = 0; pi < count; pi++) {
                    const int c1 = 2 * (pp + pi), c2 = c1 + 1;
                    for (int d = 0; d < DMB; d++) {
                        uint32_t o = 0;
                        if (d0 + d < d1) {
                            const int a1 = p->off[(size_t)(d0 + d) * nch + c1] - dmin[(size_t)b * nch + c1];
                            const int row = c2 < nch ? p->off[(size_t)(d0 + d) * nch + c2] -
                                                           p->off[(size_t)(d0 + d) * nch + c1] - dlo[pp + pi]
                                                     : 0;
                            o = (uint32_t)(((pi * nmax + row) * stride + a1) * 4);
                        }
                        roffs.push_back(o);
                    }
                }
                max_data = std::max(max_data, (size_t)count * nmax * 4 * stride);
                // accumulate (one add per pair) + staging (nd presum rows per pair)
                cost[b] += (int64_t)count * ((int64_t)DMB * T + 2LL * stride * nmax);
                pp += count;
            }
            blk[b].y = (int)chunks.size() - blk[b].x;
            continue;
        }
        int c = 0;
        while (c < nch) {
            int count = 0, smax = 0;
            while (c + count < nch && count < kChunkMax) {
                const int s2 = std::max(smax, span[c + count]);
                const size_t need = (size_t)(count + 1) * 4 * (size_t)stride_of(s2);
                if (need > kDataBudget) break;
                smax = s2;
                count++;
            }
            VFRB_REQUIRE(count > 0, VFRB_ERR_UNSUPPORTED,
                         "one channel's staged span (%d rows) exceeds the LDS budget; use a tile "
                         "with fewer DMs per block", T + span[c]);
            const int stride = stride_of(smax);
            chunks.push_back(Chunk{c, count, stride, (int)roffs.size(), 1, {0, 0, 0}});
            for (int ci = 0; ci < count; ci++) {
                const int cc = c + ci;
                const int phase = dmin[(size_t)b * nch + cc] & (epd - 1);
                for (int d = 0; d < DMB; d++) {
                    const int r = d0 + d < d1 ? p->off[(size_t)(d0 + d) * nch + cc] - dmin[(size_t)b * nch + cc] : 0;
                    const int e = r + phase;  // element index from the aligned start
                    const int q = e & (epd - 1);
                    roffs.push_back((uint32_t)(ci * 4 * stride + q * stride * esz + (e - q) * esz));
                }
            }
            max_data = std::max(max_data, (size_t)count * 4 * stride);
            // accumulate + staging (one load and EPD LDS writes per element)
            cost[b] += (int64_t)count * ((int64_t)DMB * T + 2LL * stride * epd);
            c += count;
        }
        blk[b].y = (int)chunks.size() - blk[b].x;
    }
    if (tc.mode >= 2) dmin.swap(recs);
    if (dmin.empty()) dmin.push_back(0);
    free_tiles(out);
    hipError_t e = hipMalloc(&out.d_roffs, roffs.size() * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&out.d_dmin, dmin.size() * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&out.d_chunks, chunks.size() * sizeof(Chunk));
    if (e == hipSuccess) e = hipMalloc(&out.d_blk_chunks, blk.size() * sizeof(int2));
    if (e == hipSuccess) e = hipMemcpy(out.d_roffs, roffs.data(), roffs.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(out.d_dmin, dmin.data(), dmin.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(out.d_chunks, chunks.data(), chunks.size() * sizeof(Chunk), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(out.d_blk_chunks, blk.data(), blk.size() * sizeof(int2), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_tiles(out);
        return vfrb::hip_fail(e, "build_tiles", __FILE__, __LINE__);
    }
    out.nblk = nblk;
    out.tile = i;
    out.blk_cost = std::move(cost);
    out.lds_bytes = lds_head_bytes(DMB) + max_data;
    return VFRB_OK;
}

vfrb_status build_tiles(vfrb_search_plan* p, int i) { return build_tile_set(p, kTiles[i], i, p->ts); }

const void* dedisp_kernel_for(const vfrb_search_plan* p) {
    return tile_kernel_for(p->storage, p->ts.tile, false);
}

void destroy(vfrb_search_plan* p) {
    free_tiles(p->ts);
    free_tiles(p->sts);
    (void)hipFree(p->d_off);
    (void)hipFree(p->d_win);
    (void)hipFree(p->d_bp_mean);
    (void)hipFree(p->d_range);
    (void)hipFree(p->d_f16_exp);
    (void)hipFree(p->d_u8_scale);
    (void)hipFree(p->d_u8_off);
    (void)hipFree(p->d_u8_offsum);
    (void)hipFree(p->d_tmp_fb);
    (void)hipFree(p->d_shift);
    delete p;
}

}  // namespace

extern "C" {

int32_t vfrb_search_tile_count(void) { return kNumTiles; }

const char* vfrb_search_tile_name(int32_t i) {
    if (i < 0 || i >= kNumTiles) return "";
    return kTiles[i].name;
}

vfrb_status vfrb_search_plan_create(vfrb_search_plan** plan, const vfrb_search_desc* d) {
    VFRB_REQUIRE(plan && d, VFRB_ERR_INVALID, "NULL argument");
    *plan = nullptr;
    VFRB_REQUIRE(d->nchans > 0 && d->ndm > 0 && d->dms, VFRB_ERR_INVALID, "empty channel or DM set");
    VFRB_REQUIRE(d->max_rows > 0, VFRB_ERR_INVALID, "max_rows=%d", d->max_rows);
    VFRB_REQUIRE(d->storage >= VFRB_STORE_F32 && d->storage <= VFRB_STORE_U8, VFRB_ERR_INVALID,
                 "storage=%d", d->storage);
    VFRB_REQUIRE(d->tile_config >= -1 && d->tile_config < kNumTiles, VFRB_ERR_INVALID,
                 "tile_config=%d", d->tile_config);
    vfrb_search_plan* p = new vfrb_search_plan();
    p->desc = *d;
    p->nch = d->nchans;
    p->ndm = d->ndm;
    p->storage = d->storage;
    p->dms.assign(d->dms, d->dms + d->ndm);
    p->desc.dms = p->dms.data();
    p->off.resize((size_t)p->ndm * p->nch);
    vfrb_status st = vfrb_delay_table(p->dms.data(), p->ndm, p->nch, d->fch1, d->chanbw, d->tsamp,
                                      p->off.data(), &p->max_delay, nullptr);
    if (st != VFRB_OK) {
        delete p;
        return st;
    }
    int32_t maxoff = 0;
    for (int32_t v : p->off) maxoff = std::max(maxoff, v);
    if (maxoff > p->max_delay) {
        delete p;
        vfrb::set_error("largest delay %d exceeds max_delay %d of the last DM (DM list must end "
                        "with its largest value)", maxoff, p->max_delay);
        return VFRB_ERR_INVALID;
    }
    // Window rows padded to a multiple of 64 elements (256-byte rows in f32).
    p->ld = ((int64_t)d->max_rows + 63) / 64 * 64;
    if ((int64_t)kChunkMax * p->ld * (int64_t)storage_bytes(p->storage) >= (1LL << 31)) {
        delete p;
        vfrb::set_error("max_rows %d too long: %d window rows must stay under 2 GiB", d->max_rows, kChunkMax);
        return VFRB_ERR_UNSUPPORTED;
    }
    const size_t win_bytes = (size_t)p->nch * p->ld * storage_bytes(p->storage);
    const int nbpblk = (d->max_rows + BP_ROWS - 1) / BP_ROWS;
    hipError_t e = hipMalloc(&p->d_off, p->off.size() * sizeof(int32_t));
    if (e == hipSuccess) e = hipMemcpy(p->d_off, p->off.data(), p->off.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&p->d_win, win_bytes);
    if (e == hipSuccess) e = hipMemset(p->d_win, 0, win_bytes);
    if (e == hipSuccess) e = hipMalloc(&p->d_bp_mean, sizeof(float) * (size_t)nbpblk * p->nch);
    if (e == hipSuccess) e = hipMalloc(&p->d_range, sizeof(unsigned) * 2 * p->nch);
    if (e == hipSuccess) e = hipMalloc(&p->d_f16_exp, sizeof(int));
    if (e == hipSuccess) e = hipMemset(p->d_f16_exp, 0, sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&p->d_u8_scale, sizeof(float) * p->nch);
    if (e == hipSuccess) e = hipMalloc(&p->d_u8_off, sizeof(float) * p->nch);
    if (e == hipSuccess) e = hipMalloc(&p->d_u8_offsum, sizeof(float));
    if (e != hipSuccess) {
        st = vfrb::hip_fail(e, "vfrb_search_plan_create", __FILE__, __LINE__);
        destroy(p);
        return st;
    }
    if (d->tile_config >= 0) {
        st = build_tiles(p, d->tile_config);
    } else {
        st = VFRB_ERR_UNSUPPORTED;
        for (int t : kTileFallback)
            if (st == VFRB_ERR_UNSUPPORTED) st = build_tiles(p, t);
    }
    if (st != VFRB_OK) {
        destroy(p);
        return st;
    }
    *plan = p;
    return VFRB_OK;
}

vfrb_status vfrb_search_plan_destroy(vfrb_search_plan* p) {
    if (p) destroy(p);
    return VFRB_OK;
}

int32_t vfrb_search_max_delay(const vfrb_search_plan* p) { return p ? p->max_delay : -1; }
int32_t vfrb_search_dm_count(const vfrb_search_plan* p) { return p ? p->ndm : -1; }

vfrb_status vfrb_search_get_offsets(const vfrb_search_plan* p, int32_t* host) {
    VFRB_REQUIRE(p && host, VFRB_ERR_INVALID, "NULL argument");
    std::memcpy(host, p->off.data(), p->off.size() * sizeof(int32_t));
    return VFRB_OK;
}

vfrb_status vfrb_search_set_tile(vfrb_search_plan* p, int32_t i) {
    VFRB_REQUIRE(p && i >= 0 && i < kNumTiles, VFRB_ERR_INVALID, "tile %d", i);
    if (p->ts.tile == i) return VFRB_OK;
    return build_tiles(p, i);
}

int32_t vfrb_search_get_tile(const vfrb_search_plan* p) { return p ? p->ts.tile : -1; }

vfrb_status vfrb_search_load(vfrb_search_plan* p, const float* fb, int64_t ring_rows, int64_t row0,
                             int32_t nrows, int32_t apply_bp, void* stream) {
    VFRB_REQUIRE(p && fb, VFRB_ERR_INVALID, "NULL argument");
    VFRB_REQUIRE(nrows > 0 && nrows <= p->desc.max_rows, VFRB_ERR_INVALID,
                 "nrows=%d outside 1..%d", nrows, p->desc.max_rows);
    VFRB_REQUIRE(ring_rows == 0 || nrows <= ring_rows, VFRB_ERR_INVALID, "window exceeds ring");
    VFRB_REQUIRE(row0 >= 0, VFRB_ERR_INVALID, "row0 < 0");
    hipStream_t s = vfrb::as_stream(stream);
    dim3 grid((unsigned)vfrb::ceil_div(p->nch, BP_CH), (unsigned)vfrb::ceil_div(nrows, BP_ROWS));
    dim3 block(256);
    if (p->storage == VFRB_STORE_F32) {
        hipLaunchKernelGGL(load_kernel<LOAD_F32>, grid, block, 0, s, fb, ring_rows, row0, nrows,
                           p->nch, apply_bp, p->d_win, p->ld, p->d_bp_mean, nullptr, nullptr,
                           nullptr, nullptr);
        VFRB_LAUNCH_CHECK("load_kernel<f32>");
    } else {
        hipLaunchKernelGGL(range_init_kernel, dim3((unsigned)vfrb::ceil_div(2 * p->nch, 256)), dim3(256),
                           0, s, p->d_range, p->nch);
        VFRB_LAUNCH_CHECK("range_init_kernel");
        hipLaunchKernelGGL(load_kernel<LOAD_STATS>, grid, block, 0, s, fb, ring_rows, row0, nrows,
                           p->nch, apply_bp, p->d_win, p->ld, p->d_bp_mean, p->d_range, nullptr,
                           nullptr, nullptr);
        VFRB_LAUNCH_CHECK("load_kernel<stats>");
        hipLaunchKernelGGL(finish_range_kernel, dim3(1), dim3(256), 0, s, p->d_range, p->nch,
                           p->storage, p->d_f16_exp, p->d_u8_scale, p->d_u8_off, p->d_u8_offsum);
        VFRB_LAUNCH_CHECK("finish_range_kernel");
        if (p->storage == VFRB_STORE_F16) {
            hipLaunchKernelGGL(load_kernel<LOAD_APPLY_F16>, grid, block, 0, s, fb, ring_rows, row0,
                               nrows, p->nch, apply_bp, p->d_win, p->ld, p->d_bp_mean, nullptr,
                               p->d_f16_exp, nullptr, nullptr);
        } else {
            hipLaunchKernelGGL(load_kernel<LOAD_APPLY_U8>, grid, block, 0, s, fb, ring_rows, row0,
                               nrows, p->nch, apply_bp, p->d_win, p->ld, p->d_bp_mean, nullptr,
                               nullptr, p->d_u8_scale, p->d_u8_off);
        }
        VFRB_LAUNCH_CHECK("load_kernel<apply>");
    }
    p->loaded_rows = nrows;
    return VFRB_OK;
}

vfrb_status vfrb_search_dedisperse(vfrb_search_plan* p, int32_t nrows, float* out, int64_t ld_out,
                                   void* stream) {
    VFRB_REQUIRE(p && out, VFRB_ERR_INVALID, "NULL argument");
    VFRB_REQUIRE(nrows > 0 && nrows <= p->desc.max_rows, VFRB_ERR_INVALID, "nrows=%d", nrows);
    int nt = nrows - p->max_delay;
    VFRB_REQUIRE(nt > 0, VFRB_ERR_INVALID, "window of %d rows is not longer than max_delay %d",
                 nrows, p->max_delay);
    VFRB_REQUIRE(ld_out >= nt, VFRB_ERR_INVALID, "ld_out=%lld < %d", (long long)ld_out, nt);
    VFRB_REQUIRE(nrows <= p->loaded_rows, VFRB_ERR_INVALID,
                 "window holds %d rows, %d requested", p->loaded_rows, nrows);
    const TileCfg& tc = kTiles[p->ts.tile];
    const int T = tile_samples(tc.sl, tc.mode);
    const int64_t nwg = vfrb::ceil_div(nt, T) * p->ts.nblk;
    VFRB_REQUIRE(nwg < INT32_MAX, VFRB_ERR_UNSUPPORTED, "%lld dedispersion tiles", (long long)nwg);
    dim3 grid((unsigned)nwg), block(64 * tc.w);
    int nblk = p->ts.nblk;
    const void* kern = dedisp_kernel_for(p);
    VFRB_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->ts.lds_bytes));
    const void* win = p->d_win;
    int64_t ld = p->ld;
    int nch = p->nch, ndm = p->ndm;
    void* args[] = {(void*)&win, (void*)&ld, (void*)&nch, (void*)&nrows, (void*)&nt,
                    (void*)&ndm, (void*)&p->ts.d_roffs, (void*)&p->ts.d_dmin, (void*)&p->ts.d_chunks,
                    (void*)&p->ts.d_blk_chunks, (void*)&p->d_u8_scale, (void*)&p->d_f16_exp,
                    (void*)&p->d_u8_offsum, (void*)&out, (void*)&ld_out, (void*)&nblk};
    VFRB_HIP(hipLaunchKernel(kern, grid, block, args, p->ts.lds_bytes, vfrb::as_stream(stream)));
    return VFRB_OK;
}

// Autotuner: times every tile configuration that fits this DM list on the
// window currently loaded and keeps the fastest.
vfrb_status vfrb_search_autotune(vfrb_search_plan* p, int32_t nrows, float* out, int64_t ld_out,
                                 int32_t iters, float* ms, void* stream) {
    VFRB_REQUIRE(p && out && iters > 0, VFRB_ERR_INVALID, "bad autotune arguments");
    hipStream_t s = vfrb::as_stream(stream);
    hipEvent_t e0, e1;
    VFRB_HIP(hipEventCreate(&e0));
    VFRB_HIP(hipEventCreate(&e1));
    const int keep = p->ts.tile;
    int best = -1;
    float best_ms = 0.0f;
    vfrb_status st = VFRB_OK;
    for (int i = 0; i < kNumTiles; i++) {
        if (ms) ms[i] = -1.0f;
        if (build_tiles(p, i) != VFRB_OK) continue;  // does not fit this DM list
        st = vfrb_search_dedisperse(p, nrows, out, ld_out, stream);  // warm-up
        if (st != VFRB_OK) break;
        if (hipEventRecord(e0, s) != hipSuccess) { st = VFRB_ERR_HIP; break; }
        for (int k = 0; k < iters && st == VFRB_OK; k++) st = vfrb_search_dedisperse(p, nrows, out, ld_out, stream);
        if (st != VFRB_OK) break;
        float t = 0.0f;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
            st = VFRB_ERR_HIP;
            break;
        }
        t /= iters;
        if (ms) ms[i] = t;
        if (best < 0 || t < best_ms) {
            best = i;
            best_ms = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const vfrb_status st2 = build_tiles(p, best >= 0 ? best : keep);
    if (st != VFRB_OK) return st;
    return st2;
}

vfrb_status vfrb_search_execute(vfrb_search_plan* p, int32_t nrows, const float* fb_tm, float* out,
                                int64_t ld_out, void* stream) {
    vfrb_status st = vfrb_search_load(p, fb_tm, 0, 0, nrows, 0, stream);
    if (st != VFRB_OK) return st;
    return vfrb_search_dedisperse(p, nrows, out, ld_out, stream);
}

vfrb_status vfrb_search_dedisperse_stats(vfrb_search_plan* p, int32_t nrows, float* out, float* sig,
                                         int64_t ld_out, void* stream) {
    VFRB_REQUIRE(p && out && sig, VFRB_ERR_INVALID, "NULL argument");
    VFRB_REQUIRE(p->storage == VFRB_STORE_F32, VFRB_ERR_UNSUPPORTED, "needs an f32 plan");
    const int nt = nrows - p->max_delay;
    VFRB_REQUIRE(nrows > 0 && nrows <= p->desc.max_rows && nt > 0 && ld_out >= nt, VFRB_ERR_INVALID,
                 "nrows=%d max_delay=%d ld_out=%lld", nrows, p->max_delay, (long long)ld_out);
    VFRB_REQUIRE(nrows <= p->loaded_rows, VFRB_ERR_INVALID,
                 "window holds %d rows, %d requested", p->loaded_rows, nrows);
    if (p->sts.tile < 0) {
        const vfrb_status st = build_tile_set(p, kStatsTile, 0, p->sts);
        if (st != VFRB_OK) return st;
    }
    if (!p->d_shift) VFRB_HIP(hipMalloc(&p->d_shift, sizeof(float)));
    hipLaunchKernelGGL(stats_shift_kernel, dim3(1), dim3(256), 0, vfrb::as_stream(stream), (const float*)p->d_win,
                       p->ld, p->nch, nrows, p->d_shift);
    constexpr TileCfg tc = kStatsTile;
    const int64_t nwg = vfrb::ceil_div(nt, 64 * tc.sl) * p->sts.nblk;
    VFRB_REQUIRE(nwg < INT32_MAX, VFRB_ERR_UNSUPPORTED, "%lld dedispersion tiles", (long long)nwg);
    const void* kern = reinterpret_cast<const void*>(&dedisp_stats_kernel<tc.sl, tc.dw, tc.w>);
    VFRB_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->sts.lds_bytes));
    hipLaunchKernelGGL((dedisp_stats_kernel<tc.sl, tc.dw, tc.w>), dim3((unsigned)nwg), dim3(64 * tc.w),
                       p->sts.lds_bytes, vfrb::as_stream(stream), (const float*)p->d_win, p->ld, p->nch, nrows,
                       nt, p->ndm, p->sts.d_roffs, p->sts.d_dmin, p->sts.d_chunks, p->sts.d_blk_chunks, out, sig,
                       ld_out, p->sts.nblk, (const float*)p->d_shift);
    VFRB_LAUNCH_CHECK("dedisp_stats_kernel");
    return VFRB_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched dedispersion (host side)
// ---------------------------------------------------------------------------
namespace {

void drop_graph(vfrb_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// LDS of a batch launch: the largest any plan can need with tile i, so one
// captured launch serves every job set.
vfrb_status batch_configure(vfrb_batch* b, int i) {
    const TileCfg& tc = kTiles[i];
    const size_t lds = lds_head_bytes(tc.w * tc.dw) + kDataBudget;
    const void* kern = tile_kernel_for(b->storage, i, true);
    VFRB_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int dev = 0, per_cu = 0, cus = 0;
    VFRB_HIP(hipGetDevice(&dev));
    VFRB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    VFRB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * tc.w, lds));
    VFRB_REQUIRE(per_cu > 0, VFRB_ERR_UNSUPPORTED, "tile %s does not fit a compute unit", tc.name);
    drop_graph(b);
    b->tile = i;
    b->lds_bytes = lds;
    b->grid = per_cu * cus;
    return VFRB_OK;
}

// Builds the work list of the stored jobs for the batch's tile configuration
// and uploads it on stream s.  Work items are ordered by the estimated cost
// of their DM block, largest first, so the persistent grid drains evenly.
vfrb_status batch_build(vfrb_batch* b, hipStream_t s) {
    const int njobs = (int)b->plans.size();
    if (b->staged_pending) {
        VFRB_HIP(hipEventSynchronize(b->staged));
        b->staged_pending = false;
    }
    const TileCfg& tc = kTiles[b->tile];
    const int T = tile_samples(tc.sl, tc.mode);
    struct Blk {
        int64_t cost;
        int job, blk;
    };
    std::vector<Blk> blks;
    int64_t total = 0;
    for (int j = 0; j < njobs; j++) {
        vfrb_search_plan* p = b->plans[j];
        if (p->ts.tile != b->tile) {
            const vfrb_status st = build_tiles(p, b->tile);
            if (st != VFRB_OK) return st;
        }
        const int nrows = b->nrows[j];
        const int nt = nrows - p->max_delay;
        BatchJob& jb = b->h_jobs[j];
        jb.win = p->d_win;
        jb.ld = p->ld;
        jb.nch = p->nch;
        jb.nrows = nrows;
        jb.nt = nt;
        jb.ndm = p->ndm;
        jb.roffs = p->ts.d_roffs;
        jb.dmin = p->ts.d_dmin;
        jb.chunks = p->ts.d_chunks;
        jb.blk_chunks = p->ts.d_blk_chunks;
        jb.chan_scale = p->d_u8_scale;
        jb.f16_exp = p->d_f16_exp;
        jb.offsum = p->d_u8_offsum;
        jb.out = b->outs[j];
        jb.ld_out = b->ld_outs[j];
        for (int k = 0; k < p->ts.nblk; k++) blks.push_back(Blk{p->ts.blk_cost[k], j, k});
        total += (int64_t)p->ts.nblk * vfrb::ceil_div(nt, T);
    }
    VFRB_REQUIRE(total <= b->max_tiles, VFRB_ERR_CAPACITY,
                 "%lld work items exceed the batch capacity %d", (long long)total, b->max_tiles);
    std::stable_sort(blks.begin(), blks.end(), [](const Blk& x, const Blk& y) { return x.cost > y.cost; });
    int n = 0;
    for (const Blk& k : blks) {
        const int ntt = (int)vfrb::ceil_div(b->h_jobs[k.job].nt, T);
        for (int x = 0; x < ntt; x++) b->h_tiles[n++] = BatchTile{k.job, k.blk, x, 0};
    }
    *b->h_ntiles = n;
    VFRB_HIP(hipMemcpyAsync(b->d_jobs, b->h_jobs, sizeof(BatchJob) * (size_t)std::max(njobs, 1),
                            hipMemcpyHostToDevice, s));
    VFRB_HIP(hipMemcpyAsync(b->d_tiles, b->h_tiles, sizeof(BatchTile) * (size_t)std::max(n, 1),
                            hipMemcpyHostToDevice, s));
    VFRB_HIP(hipMemcpyAsync(b->d_ntiles, b->h_ntiles, sizeof(int), hipMemcpyHostToDevice, s));
    VFRB_HIP(hipEventRecord(b->staged, s));
    b->staged_pending = true;
    b->ntiles = n;
    return VFRB_OK;
}

vfrb_status batch_launch(vfrb_batch* b, hipStream_t s) {
    const TileCfg& tc = kTiles[b->tile];
    const void* kern = tile_kernel_for(b->storage, b->tile, true);
    void* args[] = {(void*)&b->d_jobs, (void*)&b->d_tiles, (void*)&b->d_ntiles, (void*)&b->d_ctr};
    VFRB_HIP(hipLaunchKernel(kern, dim3((unsigned)b->grid), dim3(64 * tc.w), args, b->lds_bytes, s));
    return VFRB_OK;
}

void batch_free(vfrb_batch* b) {
    drop_graph(b);
    if (b->staged) {
        if (b->staged_pending) (void)hipEventSynchronize(b->staged);
        (void)hipEventDestroy(b->staged);
    }
    (void)hipHostFree(b->h_jobs);
    (void)hipHostFree(b->h_tiles);
    (void)hipHostFree(b->h_ntiles);
    (void)hipFree(b->d_jobs);
    (void)hipFree(b->d_tiles);
    (void)hipFree(b->d_ntiles);
    (void)hipFree(b->d_ctr);
    delete b;
}

}  // namespace

extern "C" {

vfrb_status vfrb_batch_create(vfrb_batch** batch, int32_t storage, int32_t tile_config,
                              int32_t max_jobs, int32_t max_tiles) {
    VFRB_REQUIRE(batch, VFRB_ERR_INVALID, "NULL argument");
    *batch = nullptr;
    VFRB_REQUIRE(storage >= VFRB_STORE_F32 && storage <= VFRB_STORE_U8, VFRB_ERR_INVALID,
                 "storage=%d", storage);
    VFRB_REQUIRE(tile_config >= -1 && tile_config < kNumTiles, VFRB_ERR_INVALID,
                 "tile_config=%d", tile_config);
    VFRB_REQUIRE(max_jobs > 0 && max_tiles > 0, VFRB_ERR_INVALID, "max_jobs=%d max_tiles=%d",
                 max_jobs, max_tiles);
    vfrb_batch* b = new vfrb_batch();
    b->storage = storage;
    b->max_jobs = max_jobs;
    b->max_tiles = max_tiles;
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&b->h_jobs), sizeof(BatchJob) * max_jobs);
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&b->h_tiles), sizeof(BatchTile) * (size_t)max_tiles);
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&b->h_ntiles), sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&b->d_jobs, sizeof(BatchJob) * max_jobs);
    if (e == hipSuccess) e = hipMalloc(&b->d_tiles, sizeof(BatchTile) * (size_t)max_tiles);
    if (e == hipSuccess) e = hipMalloc(&b->d_ntiles, sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&b->d_ctr, 2 * sizeof(unsigned));
    if (e == hipSuccess) e = hipMemset(b->d_ctr, 0, 2 * sizeof(unsigned));
    if (e == hipSuccess) e = hipMemset(b->d_ntiles, 0, sizeof(int));
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->staged, hipEventDisableTiming);
    if (e != hipSuccess) {
        const vfrb_status st = vfrb::hip_fail(e, "vfrb_batch_create", __FILE__, __LINE__);
        batch_free(b);
        return st;
    }
    b->tile_default = tile_config < 0;
    b->pref_tile = tile_config >= 0 ? tile_config : kDefaultTile;
    const vfrb_status st = batch_configure(b, b->pref_tile);
    if (st != VFRB_OK) {
        batch_free(b);
        return st;
    }
    *batch = b;
    return VFRB_OK;
}

vfrb_status vfrb_batch_destroy(vfrb_batch* batch) {
    if (batch) batch_free(batch);
    return VFRB_OK;
}

vfrb_status vfrb_batch_set(vfrb_batch* b, int32_t njobs, vfrb_search_plan* const* plans,
                           const int32_t* nrows, float* const* outs, const int64_t* ld_outs,
                           int32_t* ntiles, void* stream) {
    VFRB_REQUIRE(b && njobs >= 0, VFRB_ERR_INVALID, "bad batch arguments");
    VFRB_REQUIRE(njobs == 0 || (plans && nrows && outs && ld_outs), VFRB_ERR_INVALID, "NULL argument");
    VFRB_REQUIRE(njobs <= b->max_jobs, VFRB_ERR_CAPACITY, "%d jobs exceed the batch capacity %d",
                 njobs, b->max_jobs);
    for (int j = 0; j < njobs; j++) {
        const vfrb_search_plan* p = plans[j];
        VFRB_REQUIRE(p && outs[j], VFRB_ERR_INVALID, "job %d: NULL plan or output", j);
        VFRB_REQUIRE(p->storage == b->storage, VFRB_ERR_INVALID,
                     "job %d: plan storage %d differs from the batch's %d", j, p->storage, b->storage);
        const int nt = nrows[j] - p->max_delay;
        VFRB_REQUIRE(nrows[j] <= p->loaded_rows && nt > 0, VFRB_ERR_INVALID,
                     "job %d: %d rows requested, %d loaded, max_delay %d", j, nrows[j],
                     p->loaded_rows, p->max_delay);
        VFRB_REQUIRE(ld_outs[j] >= nt, VFRB_ERR_INVALID, "job %d: ld_out=%lld < %d", j,
                     (long long)ld_outs[j], nt);
    }
    b->plans.assign(plans, plans + njobs);
    b->nrows.assign(nrows, nrows + njobs);
    b->outs.assign(outs, outs + njobs);
    b->ld_outs.assign(ld_outs, ld_outs + njobs);
    hipStream_t s = vfrb::as_stream(stream);
    vfrb_status st = VFRB_ERR_UNSUPPORTED;
    if (!b->tile_default) {
        st = batch_build(b, s);
    } else {
        // A default-tiled batch starts every job set from its preferred tile
        // (the default, or the autotuner's pick) and then walks the whole
        // fallback list, so a coarse DM list in one set leaves later sets on
        // the tile they would have had anyway.  A tile that is already
        // configured keeps its captured graph.
        int order[1 + sizeof(kTileFallback) / sizeof(int)];
        int n = 0;
        order[n++] = b->pref_tile;
        for (int t : kTileFallback)
            if (t != b->pref_tile) order[n++] = t;
        for (int i = 0; i < n && st == VFRB_ERR_UNSUPPORTED; i++) {
            st = b->tile == order[i] ? VFRB_OK : batch_configure(b, order[i]);
            if (st == VFRB_OK) st = batch_build(b, s);
        }
    }
    if (st != VFRB_OK) {
        b->ntiles = -1;
        return st;
    }
    if (ntiles) *ntiles = b->ntiles;
    return VFRB_OK;
}

vfrb_status vfrb_batch_dedisperse(vfrb_batch* b, void* stream) {
    VFRB_REQUIRE(b && b->ntiles >= 0, VFRB_ERR_INVALID, "no jobs set");
    return batch_launch(b, vfrb::as_stream(stream));
}

vfrb_status vfrb_batch_graph_launch(vfrb_batch* b, void* stream) {
    VFRB_REQUIRE(b && b->ntiles >= 0, VFRB_ERR_INVALID, "no jobs set");
    VFRB_REQUIRE(stream, VFRB_ERR_INVALID, "graph capture needs a created stream");
    hipStream_t s = vfrb::as_stream(stream);
    if (!b->exec) {
        VFRB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const vfrb_status st = batch_launch(b, s);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (st != VFRB_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return vfrb::hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
        b->graph = g;
        const hipError_t e2 = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e2 != hipSuccess) {
            drop_graph(b);
            return vfrb::hip_fail(e2, "hipGraphInstantiate", __FILE__, __LINE__);
        }
    }
    VFRB_HIP(hipGraphLaunch(b->exec, s));
    return VFRB_OK;
}

vfrb_status vfrb_batch_autotune(vfrb_batch* b, int32_t iters, float* ms, void* stream) {
    VFRB_REQUIRE(b && b->ntiles >= 0 && iters > 0, VFRB_ERR_INVALID, "bad autotune arguments");
    hipStream_t s = vfrb::as_stream(stream);
    hipEvent_t e0, e1;
    VFRB_HIP(hipEventCreate(&e0));
    VFRB_HIP(hipEventCreate(&e1));
    const int keep = b->tile;
    int best = -1;
    float best_ms = 0.0f;
    vfrb_status st = VFRB_OK;
    for (int i = 0; i < kNumTiles; i++) {
        if (ms) ms[i] = -1.0f;
        if (batch_configure(b, i) != VFRB_OK || batch_build(b, s) != VFRB_OK) continue;
        st = batch_launch(b, s);  // warm-up
        if (st != VFRB_OK) break;
        if (hipEventRecord(e0, s) != hipSuccess) { st = VFRB_ERR_HIP; break; }
        for (int k = 0; k < iters && st == VFRB_OK; k++) st = batch_launch(b, s);
        if (st != VFRB_OK) break;
        float t = 0.0f;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
            st = VFRB_ERR_HIP;
            break;
        }
        t /= iters;
        if (ms) ms[i] = t;
        if (best < 0 || t < best_ms) {
            best = i;
            best_ms = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    vfrb_status st2 = batch_configure(b, best >= 0 ? best : keep);
    if (st2 == VFRB_OK) st2 = batch_build(b, s);
    if (st2 == VFRB_OK && best >= 0) b->pref_tile = best;
    if (st != VFRB_OK) return st;
    return st2;
}

int32_t vfrb_batch_get_tile(const vfrb_batch* b) { return b ? b->tile : -1; }
int32_t vfrb_batch_grid(const vfrb_batch* b) { return b ? b->grid : -1; }

}  // extern "C"
