// Device helpers for the block bandpass statistics of subtractbpCAL
// (PIPELINE:1427-1474): mean, variance-as-sigma, clipped mean.
#pragma once

#include <hip/hip_runtime.h>

namespace vfrb {

// Workgroup layout used by every bandpass-style kernel: 256 threads cover 64
// channels (lane = channel) x 4 row quarters; each thread owns up to 32 rows
// of its channel inside one block of up to 128 rows.
constexpr int BP_CH = 64;
constexpr int BP_ROWS = 128;
constexpr int BP_Q = 4;
constexpr int BP_PER = BP_ROWS / BP_Q;  // 32

// Sum over the 4 row quarters of one channel; red is [BP_Q][BP_CH] in LDS.
__device__ __forceinline__ float bp_reduce(float v, float* red, int lane, int q) {
    __syncthreads();
    red[q * BP_CH + lane] = v;
    __syncthreads();
    return (red[lane] + red[BP_CH + lane]) + (red[2 * BP_CH + lane] + red[3 * BP_CH + lane]);
}

__device__ __forceinline__ int bp_reduce_int(int v, int* red, int lane, int q) {
    __syncthreads();
    red[q * BP_CH + lane] = v;
    __syncthreads();
    return red[lane] + red[BP_CH + lane] + red[2 * BP_CH + lane] + red[3 * BP_CH + lane];
}

// Clipped block mean of subtractbpCAL for the values x[0..n_mine) of this
// thread; npts = rows in the block (all quarters).  Arithmetic follows the
// reference: float sums, "sig" = (sum2 - sum*sum/npts)/(npts-1) (the
// variance), HI/LO = ave +- 3.0*sig evaluated in double and stored as float,
// strict inequalities, count 0 -> 1.  With npts == 1 the variance is 0/0,
// every comparison is false and the block mean becomes 0.
__device__ __forceinline__ float bp_clipped_mean(const float* x, int n_mine, int npts, float* red,
                                                 int* ired, int lane, int q) {
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < BP_PER; i++)
        if (i < n_mine) s += x[i];
    float ave = bp_reduce(s, red, lane, q) / (float)npts;
    float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
    for (int i = 0; i < BP_PER; i++) {
        if (i < n_mine) {
            const float d = x[i] - ave;
            s2 += d * d;
            s1 += d;
        }
    }
    const float sum = bp_reduce(s1, red, lane, q);
    const float sum2 = bp_reduce(s2, red, lane, q);
    const float sig = (sum2 - (sum * sum) / (float)npts) / (float)(npts - 1);
    const float hi = (float)((double)ave + (3.0 * (double)sig));
    const float lo = (float)((double)ave - (3.0 * (double)sig));
    float cs = 0.0f;
    int cnt = 0;
#pragma unroll
    for (int i = 0; i < BP_PER; i++) {
        if (i < n_mine && x[i] < hi && x[i] > lo) {
            cs += x[i];
            cnt++;
        }
    }
    const float csum = bp_reduce(cs, red, lane, q);
    int ctot = bp_reduce_int(cnt, ired, lane, q);
    if (ctot == 0) ctot = 1;
    return csum / (float)ctot;
}

}  // namespace vfrb
