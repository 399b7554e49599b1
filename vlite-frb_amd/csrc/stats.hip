// Statistics kernels: in-place bandpass (subtractbpCAL), per-channel block
// statistics (STATS calcstats), per-DM statistics (PIPELINE calcstats) and
// the candidate list.
#include <cmath>

#include "bandpass.hpp"
#include "common.hpp"

using vfrb::BP_CH;
using vfrb::BP_PER;
using vfrb::BP_Q;
using vfrb::BP_ROWS;

namespace {

// subtractbpCAL (PIPELINE:1427-1474) on a time-major f32 filterbank, in place.
__global__ void __launch_bounds__(256)
bandpass_inplace_kernel(float* __restrict__ fb, int nrows, int nch) {
    __shared__ float red[BP_Q * BP_CH];
    __shared__ int ired[BP_Q * BP_CH];
    const int lane = threadIdx.x & 63;
    const int q = threadIdx.x >> 6;
    const int c = blockIdx.x * BP_CH + lane;
    const int rbase = blockIdx.y * BP_ROWS;
    const int npts = min(BP_ROWS, nrows - rbase);
    const int mine0 = q * BP_PER;
    const bool cvalid = c < nch;
    const int n_mine = cvalid ? max(0, min(BP_PER, npts - mine0)) : 0;
    float x[BP_PER];
#pragma unroll
    for (int i = 0; i < BP_PER; i++)
        x[i] = i < n_mine ? fb[(int64_t)(rbase + mine0 + i) * nch + c] : 0.0f;
    const float m = vfrb::bp_clipped_mean(x, n_mine, npts, red, ired, lane, q);
#pragma unroll
    for (int i = 0; i < BP_PER; i++)
        if (i < n_mine) fb[(int64_t)(rbase + mine0 + i) * nch + c] = x[i] - m;
}

// Calls f(r) for the rows r of [lo, hi) that row quarter q owns: the 32-row
// chunks k = r / 32 with k % 4 == q, i.e. exactly the rows quarter q holds of
// every 128-row bandpass block, in increasing order.
// A whole chunk is unrolled, so that its 32 loads are in flight together.
template <class F>
__device__ __forceinline__ void for_owned_rows(int lo, int hi, int q, F f) {
    const int k0 = lo >> 5;
    for (int k = k0 + ((q - k0) & 3); (k << 5) < hi; k += 4) {
        const int kb = k << 5;
        if (kb >= lo && kb + 32 <= hi) {
#pragma unroll
            for (int i = 0; i < 32; i++) f(kb + i);
        } else {
            for (int r = max(kb, lo); r < min(kb + 32, hi); r++) f(r);
        }
    }
}

// One thread's share of a channel's stat block for calcstats (STATS:910-946),
// in one pass: the plain float sum s (the mean, as the reference forms it),
// sums a and b of x - K and (x - K)^2 about a shift K (the block's first
// sample, so that the spread needs no second pass and has no cancellation
// beyond the data's own), and the smallest sample (a NaN counts as -inf).
struct ChanAcc {
    float s = 0.0f, a = 0.0f, b = 0.0f, mn = INFINITY;
    __device__ __forceinline__ void add(float x, float k) {
        s += x;
        const float t = x - k;
        a += t;
        b = fmaf(t, t, b);
        mn = fminf(mn, x != x ? -INFINITY : x);
    }
};

__device__ __forceinline__ float bp_reduce_min(float v, float* red, int lane, int q) {
    __syncthreads();
    red[q * BP_CH + lane] = v;
    __syncthreads();
    return fminf(fminf(red[lane], red[BP_CH + lane]), fminf(red[2 * BP_CH + lane], red[3 * BP_CH + lane]));
}

// calcstats of one channel and stat block of npts rows from the four
// quarters' ChanAcc: ave = s/npts; with d = ave - K, sum = sum(x - ave) =
// a - npts*d and sum2 = sum((x - ave)^2) = b - 2*d*a + npts*d^2 (exact in
// real arithmetic, formed in double); sig = (sum2 - sum*sum/npts)/(npts-1)
// (the variance); the outlier fraction is the reference's count/npts in
// integer division, which is 1 exactly when every sample exceeds
// HI = ave + 5*sig, i.e. when the smallest does (a NaN sample never does).
__device__ __forceinline__ void chan_stats_finish(const ChanAcc& acc, float k, int npts, bool cvalid, int lane,
                                                  int q, float* red, int64_t o, float* __restrict__ ave,
                                                  float* __restrict__ sig, float* __restrict__ outlier) {
    const float m = vfrb::bp_reduce(acc.s, red, lane, q) / (float)npts;
    const float a = vfrb::bp_reduce(acc.a, red, lane, q);
    const float b = vfrb::bp_reduce(acc.b, red, lane, q);
    const float mn = bp_reduce_min(acc.mn, red, lane, q);
    const double n = npts, d = (double)m - (double)k;
    const double sum = (double)a - n * d;
    const double sum2 = (double)b - 2.0 * d * (double)a + n * d * d;
    const float v = (float)((sum2 - (sum * sum) / n) / (n - 1.0));
    const float hi = (float)((double)m + (5.0 * (double)v));
    if (q == 0 && cvalid) {
        ave[o] = m;
        sig[o] = v;
        outlier[o] = mn > hi ? 1.0f : 0.0f;  // (count / npts), STATS:943
    }
}

// STATS calcstats (STATS:910-946) on a bandpassed filterbank.  Block = 64
// channels x 4 row quarters, one stat block; one read of the data.
__global__ void __launch_bounds__(256)
chan_stats_kernel(const float* __restrict__ fb, int nrows, int nch, int ncol, float* __restrict__ ave,
                  float* __restrict__ sig, float* __restrict__ outlier) {
    __shared__ float red[BP_Q * BP_CH];
    const int lane = threadIdx.x & 63;
    const int q = threadIdx.x >> 6;
    const int c = blockIdx.x * BP_CH + lane;
    const int colmin = ncol * blockIdx.y;
    const int colmax = min(colmin + ncol, nrows);
    const bool cvalid = c < nch;
    auto at = [&](int r) { return fb[(int64_t)r * nch + c]; };
    const float k = cvalid ? at(colmin) : 0.0f;
    ChanAcc acc;
    if (cvalid) for_owned_rows(colmin, colmax, q, [&](int r) { acc.add(at(r), k); });
    chan_stats_finish(acc, k, colmax - colmin, cvalid, lane, q, red, (int64_t)blockIdx.y * nch + c, ave, sig,
                      outlier);
}

// subtractbpCAL (PIPELINE:1427-1474) and STATS calcstats (STATS:910-946) in
// one read of the raw filterbank, which is left unchanged: the workgroup of
// (64 channels, stat block) computes the clipped mean of every 128-row
// bandpass block the stat block touches, with the layout and arithmetic of
// bandpass_inplace_kernel (so the same bandpassed floats), and feeds each
// bandpassed sample of the stat block to the statistics as it goes.  Each
// quarter sees its rows in the order chan_stats_kernel's quarter does, so
// the results equal bandpass_inplace_kernel followed by chan_stats_kernel
// bit for bit.
__global__ void __launch_bounds__(256)
bandpass_chan_stats_kernel(const float* __restrict__ fb, int nrows, int nch, int ncol,
                           float* __restrict__ ave, float* __restrict__ sig, float* __restrict__ outlier) {
    __shared__ float red[BP_Q * BP_CH];
    __shared__ int ired[BP_Q * BP_CH];
    const int lane = threadIdx.x & 63;
    const int q = threadIdx.x >> 6;
    const int c = blockIdx.x * BP_CH + lane;
    const int colmin = ncol * blockIdx.y;
    const int colmax = min(colmin + ncol, nrows);
    const bool cvalid = c < nch;
    const int b0 = colmin / BP_ROWS, b1 = (colmax - 1) / BP_ROWS;
    const int mine0 = q * BP_PER;
    float k = 0.0f;
    ChanAcc acc;
    for (int b = b0; b <= b1; b++) {
        const int rbase = b * BP_ROWS;
        const int npts = min(BP_ROWS, nrows - rbase);
        const int n_mine = cvalid ? max(0, min(BP_PER, npts - mine0)) : 0;
        float x[BP_PER];
#pragma unroll
        for (int i = 0; i < BP_PER; i++)
            x[i] = i < n_mine ? fb[(int64_t)(rbase + mine0 + i) * nch + c] : 0.0f;
        const float m = vfrb::bp_clipped_mean(x, n_mine, npts, red, ired, lane, q);
        // the shift: the first bandpassed sample of the stat block
        if (b == b0 && cvalid) k = fb[(int64_t)colmin * nch + c] - m;
#pragma unroll
        for (int i = 0; i < BP_PER; i++) {
            const int r = rbase + mine0 + i;
            if (i < n_mine && r >= colmin && r < colmax) acc.add(x[i] - m, k);
        }
    }
    chan_stats_finish(acc, k, colmax - colmin, cvalid, lane, q, red, (int64_t)blockIdx.y * nch + c, ave, sig,
                      outlier);
}

// Neumaier-compensated float sum.
struct CSum {
    float s = 0.0f, c = 0.0f;
    __device__ void add(float x) {
        const float t = s + x;
        if (fabsf(s) >= fabsf(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    __device__ float value() const { return s + c; }
};

template <typename V>
__device__ V block_sum(V v, V* sh) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s) sh[tid] += sh[tid + s];
        __syncthreads();
    }
    const V r = sh[0];
    __syncthreads();
    return r;
}

__device__ CSum block_csum(CSum v, float* sh_s, float* sh_c) {
    const int tid = threadIdx.x;
    sh_s[tid] = v.s;
    sh_c[tid] = v.c;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s) {
            CSum a;
            a.s = sh_s[tid];
            a.c = sh_c[tid];
            a.add(sh_s[tid + s]);
            a.c += sh_c[tid + s];
            sh_s[tid] = a.s;
            sh_c[tid] = a.c;
        }
        __syncthreads();
    }
    CSum r;
    r.s = sh_s[0];
    r.c = sh_c[0];
    __syncthreads();
    return r;
}

// calcstats (PIPELINE:1337-1370), one workgroup of 256 per DM.
template <int ACC>
__global__ void __launch_bounds__(256)
dm_stats_kernel(const float* __restrict__ out, int64_t ld, int nt, int* __restrict__ frb,
                float* __restrict__ ave, float* __restrict__ sigma, float* __restrict__ peak,
                int* __restrict__ peak_t) {
    __shared__ double shd[256];
    __shared__ float shs[256], shc[256];
    __shared__ int shi[256];
    const int dm = blockIdx.x;
    const int tid = threadIdx.x;
    const float* x = out + (int64_t)dm * ld;
    float ave_f, sig_f;
    if (ACC == VFRB_ACC_F64) {
        double s = 0.0;
        for (int t = tid; t < nt; t += 256) s += (double)x[t];
        const double tave = block_sum(s, shd) / nt;
        double s1 = 0.0, s2 = 0.0;
        for (int t = tid; t < nt; t += 256) {
            const double d = (double)x[t] - tave;
            s2 += d * d;
            s1 += d;
        }
        const double sum = block_sum(s1, shd);
        const double sum2 = block_sum(s2, shd);
        const double tsig = (sum2 - (sum * sum) / nt) / (nt - 1);
        ave_f = (float)tave;
        sig_f = (float)sqrt(tsig);
    } else {
        CSum s;
        for (int t = tid; t < nt; t += 256) s.add(x[t]);
        const float tave = block_csum(s, shs, shc).value() / (float)nt;
        CSum s1, s2;
        for (int t = tid; t < nt; t += 256) {
            const float d = x[t] - tave;
            s2.add(d * d);
            s1.add(d);
        }
        const float sum = block_csum(s1, shs, shc).value();
        const float sum2 = block_csum(s2, shs, shc).value();
        const float tsig = (sum2 - (sum * sum) / (float)nt) / (float)(nt - 1);
        ave_f = tave;
        sig_f = sqrtf(tsig);
    }
    // SNR in float from the stored float mean/sigma (PIPELINE:1359-1366).
    float best = 0.0f;
    int best_t = -1, flag = 0;
    for (int t = tid; t < nt; t += 256) {
        const float snr = (x[t] - ave_f) / sig_f;
        if (snr > 3.0f) flag = 1;
        if (snr > best) {
            best = snr;
            best_t = t;
        }
    }
    // max value, smallest index among equal maxima (the first occurrence).
    shs[tid] = best;
    shi[tid] = best_t;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            const float a = shs[tid], b = shs[tid + s];
            const int ia = shi[tid], ib = shi[tid + s];
            if (b > a || (b == a && ib >= 0 && (ia < 0 || ib < ia))) {
                shs[tid] = b;
                shi[tid] = ib;
            }
        }
        __syncthreads();
    }
    const int anyflag = __syncthreads_or(flag);
    if (tid == 0) {
        ave[dm] = ave_f;
        sigma[dm] = sig_f;
        if (frb) frb[dm] = anyflag ? 1 : 0;
        if (peak) peak[dm] = shs[0];
        if (peak_t) peak_t[dm] = shi[0];
    }
}

__global__ void zero_int_kernel(int* p) { *p = 0; }

// Candidate list: (x - ave)/sigma > threshold, float arithmetic as calcstats.
__global__ void candidates_kernel(const float* __restrict__ out, int64_t ld, int ndm, int nt,
                                  const float* __restrict__ ave, const float* __restrict__ sigma,
                                  float thr, int* __restrict__ cdm, int* __restrict__ ct,
                                  float* __restrict__ csnr, int max_cand, int* __restrict__ count) {
    const int dm = blockIdx.y;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (dm >= ndm || t >= nt) return;
    const float snr = (out[(int64_t)dm * ld + t] - ave[dm]) / sigma[dm];
    if (snr > thr) {
        const int slot = atomicAdd(count, 1);
        if (slot < max_cand) {
            cdm[slot] = dm;
            ct[slot] = t;
            csnr[slot] = snr;
        }
    }
}

}  // namespace

extern "C" {

vfrb_status vfrb_bandpass_inplace(float* fb, int32_t nrows, int32_t nch, void* stream) {
    VFRB_REQUIRE(fb && nrows > 0 && nch > 0, VFRB_ERR_INVALID, "bad bandpass arguments");
    dim3 grid((unsigned)vfrb::ceil_div(nch, BP_CH), (unsigned)vfrb::ceil_div(nrows, BP_ROWS));
    hipLaunchKernelGGL(bandpass_inplace_kernel, grid, dim3(256), 0, vfrb::as_stream(stream), fb, nrows, nch);
    VFRB_LAUNCH_CHECK("bandpass_inplace_kernel");
    return VFRB_OK;
}

vfrb_status vfrb_chan_stats(const float* fb, int32_t nrows, int32_t nch, int32_t ncol, int32_t nblk,
                            float* ave, float* sig, float* outlier, void* stream) {
    VFRB_REQUIRE(fb && ave && sig && outlier && nch > 0 && ncol > 0 && nblk > 0, VFRB_ERR_INVALID,
                 "bad chan_stats arguments");
    VFRB_REQUIRE((int64_t)ncol * (nblk - 1) < nrows, VFRB_ERR_INVALID,
                 "stat block %d starts beyond the %d rows", nblk - 1, nrows);
    dim3 grid((unsigned)vfrb::ceil_div(nch, BP_CH), (unsigned)nblk);
    hipLaunchKernelGGL(chan_stats_kernel, grid, dim3(256), 0, vfrb::as_stream(stream), fb, nrows, nch,
                       ncol, ave, sig, outlier);
    VFRB_LAUNCH_CHECK("chan_stats_kernel");
    return VFRB_OK;
}

vfrb_status vfrb_bandpass_chan_stats(const float* fb, int32_t nrows, int32_t nch, int32_t ncol, int32_t nblk,
                                     float* ave, float* sig, float* outlier, void* stream) {
    VFRB_REQUIRE(fb && ave && sig && outlier && nch > 0 && ncol > 0 && nblk > 0, VFRB_ERR_INVALID,
                 "bad chan_stats arguments");
    VFRB_REQUIRE((int64_t)ncol * (nblk - 1) < nrows, VFRB_ERR_INVALID,
                 "stat block %d starts beyond the %d rows", nblk - 1, nrows);
    dim3 grid((unsigned)vfrb::ceil_div(nch, BP_CH), (unsigned)nblk);
    hipLaunchKernelGGL(bandpass_chan_stats_kernel, grid, dim3(256), 0, vfrb::as_stream(stream), fb, nrows,
                       nch, ncol, ave, sig, outlier);
    VFRB_LAUNCH_CHECK("bandpass_chan_stats_kernel");
    return VFRB_OK;
}

vfrb_status vfrb_dm_stats(const float* out, int64_t ld, int32_t ndm, int32_t nt, int32_t acc,
                          int32_t* frb, float* ave, float* sigma, float* peak, int32_t* peak_t,
                          void* stream) {
    VFRB_REQUIRE(out && ave && sigma && ndm > 0 && nt > 1 && ld >= nt, VFRB_ERR_INVALID,
                 "bad dm_stats arguments");
    VFRB_REQUIRE(acc == VFRB_ACC_F64 || acc == VFRB_ACC_F32C, VFRB_ERR_INVALID, "acc=%d", acc);
    hipStream_t s = vfrb::as_stream(stream);
    if (acc == VFRB_ACC_F64)
        hipLaunchKernelGGL(dm_stats_kernel<VFRB_ACC_F64>, dim3(ndm), dim3(256), 0, s, out, ld, nt, frb,
                           ave, sigma, peak, peak_t);
    else
        hipLaunchKernelGGL(dm_stats_kernel<VFRB_ACC_F32C>, dim3(ndm), dim3(256), 0, s, out, ld, nt, frb,
                           ave, sigma, peak, peak_t);
    VFRB_LAUNCH_CHECK("dm_stats_kernel");
    return VFRB_OK;
}

vfrb_status vfrb_candidates(const float* out, int64_t ld, int32_t ndm, int32_t nt, const float* ave,
                            const float* sigma, float threshold, int32_t* cand_dm, int32_t* cand_t,
                            float* cand_snr, int32_t max_cand, int32_t* count, void* stream) {
    VFRB_REQUIRE(out && ave && sigma && count && ndm > 0 && nt > 0 && ld >= nt && max_cand >= 0,
                 VFRB_ERR_INVALID, "bad candidates arguments");
    VFRB_REQUIRE(max_cand == 0 || (cand_dm && cand_t && cand_snr), VFRB_ERR_INVALID,
                 "candidate arrays missing");
    hipStream_t s = vfrb::as_stream(stream);
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, s, count);
    VFRB_LAUNCH_CHECK("zero_int_kernel");
    dim3 grid((unsigned)vfrb::ceil_div(nt, 256), (unsigned)ndm);
    hipLaunchKernelGGL(candidates_kernel, grid, dim3(256), 0, s, out, ld, ndm, nt, ave, sigma, threshold,
                       cand_dm, cand_t, cand_snr, max_cand, count);
    VFRB_LAUNCH_CHECK("candidates_kernel");
    return VFRB_OK;
}

}  // extern "C"
