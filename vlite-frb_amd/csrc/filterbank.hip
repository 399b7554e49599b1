// Fused front end: 8-bit unpack -> R2C FFT -> |X|^2 of both pols -> NAVE
// average -> channel select -> one filterbank row per workgroup.
//
// Replaces convertarray, cufftExecR2C (x2), sumpower, avepower and the
// channel-select copies of PIPELINE:582-632,1292-1322.
//
// Both polarisations go through ONE complex FFT: z[n] = (x0[n]-128) +
// i*(x1[n]-128).  With Z = FFT(z), the spectra of the two real inputs are
// X0[k] = (Z[k] + conj Z[N-k])/2 and X1[k] = (Z[k] - conj Z[N-k])/(2i), so
//     |X0[k]|^2 + |X1[k]|^2 = (|Z[k]|^2 + |Z[N-k]|^2) / 2
// exactly; no post-processing twiddles are needed.  Subtracting 128 only
// moves the DC bin (k = 0), which is never selected (nskip >= 1); it keeps
// the FP32 transform well conditioned.
//
// The FFT is a radix-8 Stockham transform held in LDS (real and imaginary
// parts in separate arrays, one pad word every 32 to spread banks), with
// twiddles from an FP64-computed table.
#include <vector>
#include <cmath>

#include "common.hpp"

namespace {

struct cf {
    float x, y;
};
__device__ __forceinline__ cf cadd(cf a, cf b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cf csub(cf a, cf b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cf cmul(cf a, cf b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cf mul_negi(cf a) { return {a.y, -a.x}; }   // -i * a
__device__ __forceinline__ cf cscale(cf a, float s) { return {a.x * s, a.y * s}; }

__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 5); }

// Forward DFTs (kernel exp(-2*pi*i*n*k/R)) on registers.
__device__ __forceinline__ void dft2(cf* v) {
    cf a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}
__device__ __forceinline__ void dft4(cf* v) {
    cf a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    cf a2 = cadd(v[1], v[3]), a3 = csub(v[1], v[3]);
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, mul_negi(a3));
    v[3] = csub(a1, mul_negi(a3));
}
__device__ __forceinline__ void dft8(cf* v) {
    const float r2 = 0.70710678118654752440f;
    cf a0 = cadd(v[0], v[4]), a1 = csub(v[0], v[4]);
    cf a2 = cadd(v[2], v[6]), a3 = csub(v[2], v[6]);
    cf a4 = cadd(v[1], v[5]), a5 = csub(v[1], v[5]);
    cf a6 = cadd(v[3], v[7]), a7 = csub(v[3], v[7]);
    cf e0 = cadd(a0, a2), e2 = csub(a0, a2);
    cf e1 = cadd(a1, mul_negi(a3)), e3 = csub(a1, mul_negi(a3));
    cf o0 = cadd(a4, a6), o2 = csub(a4, a6);
    cf o1 = cadd(a5, mul_negi(a7)), o3 = csub(a5, mul_negi(a7));
    // w8^1 = (1-i)/sqrt2, w8^2 = -i, w8^3 = -(1+i)/sqrt2
    cf t1 = {(o1.x + o1.y) * r2, (o1.y - o1.x) * r2};
    cf t2 = mul_negi(o2);
    cf t3 = {(o3.y - o3.x) * r2, -(o3.x + o3.y) * r2};
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, t1);
    v[5] = csub(e1, t1);
    v[2] = cadd(e2, t2);
    v[6] = csub(e2, t2);
    v[3] = cadd(e3, t3);
    v[7] = csub(e3, t3);
}
__device__ __forceinline__ void dft3(cf* v) {
    const float s = 0.86602540378443864676f;
    cf t1 = cadd(v[1], v[2]);
    cf t2 = {v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y};
    cf t3 = cscale(csub(v[1], v[2]), s);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, mul_negi(t3));
    v[2] = csub(t2, mul_negi(t3));
}
__device__ __forceinline__ void dft5(cf* v) {
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    cf x0 = v[0];
    cf a1 = cadd(v[1], v[4]), a2 = cadd(v[2], v[3]);
    cf b1 = csub(v[1], v[4]), b2 = csub(v[2], v[3]);
    cf p1 = {x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y};
    cf p2 = {x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y};
    cf q1 = {s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
    cf q2 = {s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
    v[0] = cadd(x0, cadd(a1, a2));
    v[1] = cadd(p1, mul_negi(q1));
    v[4] = csub(p1, mul_negi(q1));
    v[2] = cadd(p2, mul_negi(q2));
    v[3] = csub(p2, mul_negi(q2));
}
template <int R>
__device__ __forceinline__ void dft(cf* v) {
    if constexpr (R == 2) dft2(v);
    else if constexpr (R == 3) dft3(v);
    else if constexpr (R == 4) dft4(v);
    else if constexpr (R == 5) dft5(v);
    else dft8(v);
}

// ---------------------------------------------------------------------------
// Power-of-two kernel: 32 complex elements per thread, N/32 threads per FFT,
// BLOCK/(N/32) FFTs of the same output row in flight per workgroup.
// ---------------------------------------------------------------------------
template <int N>
struct P2 {
    static constexpr int TPF = N / 32;
    static constexpr int BLOCK = TPF > 256 ? TPF : 256;
    static constexpr int FPB = BLOCK / TPF;
    static constexpr int TOT = FPB * N;
    static constexpr int LDS_WORDS = TOT + TOT / 32;      // per component
    static constexpr int CPT = (N / 2 + BLOCK - 1) / BLOCK;  // output channels per thread
    static constexpr int LOG2 = __builtin_ctz(N);
};

// One radix-R Stockham pass of size-N FFTs; t = thread index within its FFT,
// fb = element offset of that FFT in LDS.
template <int N, int R>
__device__ __forceinline__ void p2_pass(float* __restrict__ re, float* __restrict__ im, int Ns, int t,
                                        int fb, const float2* __restrict__ tw) {
    constexpr int TPF = P2<N>::TPF;
    constexpr int Q = 32 / R;
    constexpr int STRIDE = N / R;
    cf v[Q][R];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int j = t + TPF * q;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int idx = lds_pad(fb + j + r * STRIDE);
            v[q][r] = cf{re[idx], im[idx]};
        }
    }
    __syncthreads();
    const int tw_shift = P2<N>::LOG2 - __builtin_ctz(Ns * R);  // N/(Ns*R) = 1 << tw_shift
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int j = t + TPF * q;
        const int k = j & (Ns - 1);
        if (Ns > 1) {
            const float2 w = tw[k << tw_shift];
            const cf w1 = {w.x, w.y};
            if constexpr (R == 2) {
                v[q][1] = cmul(v[q][1], w1);
            } else if constexpr (R == 4) {
                const cf w2 = cmul(w1, w1), w3 = cmul(w2, w1);
                v[q][1] = cmul(v[q][1], w1);
                v[q][2] = cmul(v[q][2], w2);
                v[q][3] = cmul(v[q][3], w3);
            } else {
                const cf w2 = cmul(w1, w1), w3 = cmul(w2, w1), w4 = cmul(w2, w2);
                const cf w5 = cmul(w4, w1), w6 = cmul(w3, w3), w7 = cmul(w4, w3);
                v[q][1] = cmul(v[q][1], w1);
                v[q][2] = cmul(v[q][2], w2);
                v[q][3] = cmul(v[q][3], w3);
                v[q][4] = cmul(v[q][4], w4);
                v[q][5] = cmul(v[q][5], w5);
                v[q][6] = cmul(v[q][6], w6);
                v[q][7] = cmul(v[q][7], w7);
            }
        }
        dft<R>(v[q]);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int idx = lds_pad(fb + base + r * Ns);
            re[idx] = v[q][r].x;
            im[idx] = v[q][r].y;
        }
    }
    __syncthreads();
}

template <int N>
__global__ void __launch_bounds__(P2<N>::BLOCK)
fb_pow2_kernel(const uint8_t* __restrict__ pol0, const uint8_t* __restrict__ pol1,
               const float2* __restrict__ tw, float* __restrict__ fb, int64_t nrows, int nave,
               int nskip, int nchanout, int64_t ring_rows, int64_t row0, float scale) {
    using C = P2<N>;
    extern __shared__ float lds[];
    float* re = lds;
    float* im = lds + C::LDS_WORDS;
    const int tid = threadIdx.x;
    const int64_t row = blockIdx.x;
    if (row >= nrows) return;
    const int fslot = tid / C::TPF;  // FFT slot of this thread
    const int t = tid % C::TPF;
    const int fbase = fslot * N;

    float acc[C::CPT];
#pragma unroll
    for (int i = 0; i < C::CPT; i++) acc[i] = 0.0f;

    const int e0 = tid * 32;  // first element this thread unpacks
    const int lframe = e0 / N;
    for (int it = 0; it < nave; it += C::FPB) {
        // ---- unpack 32 samples per pol: z = (x0-128) + i(x1-128)
        const int frame = it + lframe;
        uint4 a0 = {0, 0, 0, 0}, a1 = {0, 0, 0, 0}, b0 = {0, 0, 0, 0}, b1 = {0, 0, 0, 0};
        if (frame < nave) {
            const size_t off = ((size_t)row * nave + frame) * N + (e0 % N);
            const uint4* p0 = reinterpret_cast<const uint4*>(pol0 + off);
            const uint4* p1 = reinterpret_cast<const uint4*>(pol1 + off);
            a0 = p0[0];
            a1 = p0[1];
            b0 = p1[0];
            b1 = p1[1];
        }
        const uint32_t w0[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
        const uint32_t w1[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
        const bool live = frame < nave;
#pragma unroll
        for (int w = 0; w < 8; w++) {
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int idx = lds_pad(e0 + w * 4 + b);
                re[idx] = live ? (float)((w0[w] >> (8 * b)) & 0xffu) - 128.0f : 0.0f;
                im[idx] = live ? (float)((w1[w] >> (8 * b)) & 0xffu) - 128.0f : 0.0f;
            }
        }
        __syncthreads();
        // ---- FFT: radix-8 passes, then a radix-2 or radix-4 pass
        int Ns = 1;
#pragma unroll
        for (int p = 0; p < C::LOG2 / 3; p++) {
            p2_pass<N, 8>(re, im, Ns, t, fbase, tw);
            Ns *= 8;
        }
        if constexpr (C::LOG2 % 3 == 1) p2_pass<N, 2>(re, im, Ns, t, fbase, tw);
        if constexpr (C::LOG2 % 3 == 2) p2_pass<N, 4>(re, im, Ns, t, fbase, tw);
        // ---- detect: |Z[k]|^2 + |Z[N-k]|^2 summed over the FFTs in flight
#pragma unroll
        for (int i = 0; i < C::CPT; i++) {
            const int c = tid + C::BLOCK * i;
            if (c < nchanout) {
                const int k = nskip + c;
                float s = 0.0f;
#pragma unroll
                for (int f = 0; f < C::FPB; f++) {
                    const int ia = lds_pad(f * N + k), ib = lds_pad(f * N + ((N - k) & (N - 1)));
                    s += re[ia] * re[ia] + im[ia] * im[ia] + re[ib] * re[ib] + im[ib] * im[ib];
                }
                acc[i] += s;
            }
        }
        __syncthreads();
    }
    int64_t orow = row0 + row;
    if (ring_rows > 0) orow %= ring_rows;
    float* dst = fb + orow * (int64_t)nchanout;
#pragma unroll
    for (int i = 0; i < C::CPT; i++) {
        const int c = tid + C::BLOCK * i;
        if (c < nchanout) dst[c] = acc[i] * scale;
    }
}

// ---------------------------------------------------------------------------
// Mixed-radix kernel for other even lengths (factors 2,3,4,5,8), one FFT per
// workgroup of 256 threads, N <= 16384.
// ---------------------------------------------------------------------------
constexpr int GEN_BLOCK = 256;
constexpr int GEN_NMAX = 16384;
constexpr int GEN_MAXPASS = 16;

struct GenPlan {
    int npass;
    int radix[GEN_MAXPASS];
};

template <int R>
__device__ __forceinline__ void gen_pass(float* __restrict__ re, float* __restrict__ im, int N, int Ns,
                                         const float2* __restrict__ tw) {
    constexpr int MAXQ = (GEN_NMAX / R + GEN_BLOCK - 1) / GEN_BLOCK;
    const int nb = N / R;
    const int tid = threadIdx.x;
    cf v[MAXQ][R];
#pragma unroll
    for (int q = 0; q < MAXQ; q++) {
        const int j = tid + GEN_BLOCK * q;
        if (j < nb) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int idx = lds_pad(j + r * nb);
                v[q][r] = cf{re[idx], im[idx]};
            }
        }
    }
    __syncthreads();
    const int span = Ns * R;
    const int twstep = N / span;
#pragma unroll
    for (int q = 0; q < MAXQ; q++) {
        const int j = tid + GEN_BLOCK * q;
        if (j < nb) {
            const int k = j % Ns;
#pragma unroll
            for (int r = 1; r < R; r++) {
                const float2 w = tw[r * k * twstep];
                v[q][r] = cmul(v[q][r], cf{w.x, w.y});
            }
            dft<R>(v[q]);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int idx = lds_pad(base + r * Ns);
                re[idx] = v[q][r].x;
                im[idx] = v[q][r].y;
            }
        }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(GEN_BLOCK)
fb_generic_kernel(const uint8_t* __restrict__ pol0, const uint8_t* __restrict__ pol1,
                  const float2* __restrict__ tw, GenPlan plan, int N, float* __restrict__ fb,
                  int64_t nrows, int nave, int nskip, int nchanout, int64_t ring_rows, int64_t row0,
                  float scale) {
    extern __shared__ float lds[];
    const int words = N + N / 32;
    float* re = lds;
    float* im = lds + words;
    const int tid = threadIdx.x;
    const int64_t row = blockIdx.x;
    if (row >= nrows) return;
    constexpr int CPT = (GEN_NMAX / 2 + GEN_BLOCK - 1) / GEN_BLOCK;
    float acc[CPT];
#pragma unroll
    for (int i = 0; i < CPT; i++) acc[i] = 0.0f;
    for (int frame = 0; frame < nave; frame++) {
        const size_t off = ((size_t)row * nave + frame) * N;
        for (int e = tid; e < N; e += GEN_BLOCK) {
            const int idx = lds_pad(e);
            re[idx] = (float)pol0[off + e] - 128.0f;
            im[idx] = (float)pol1[off + e] - 128.0f;
        }
        __syncthreads();
        int Ns = 1;
        for (int p = 0; p < plan.npass; p++) {
            switch (plan.radix[p]) {
                case 2: gen_pass<2>(re, im, N, Ns, tw); break;
                case 3: gen_pass<3>(re, im, N, Ns, tw); break;
                case 4: gen_pass<4>(re, im, N, Ns, tw); break;
                case 5: gen_pass<5>(re, im, N, Ns, tw); break;
                default: gen_pass<8>(re, im, N, Ns, tw); break;
            }
            Ns *= plan.radix[p];
        }
#pragma unroll
        for (int i = 0; i < CPT; i++) {
            const int c = tid + GEN_BLOCK * i;
            if (c < nchanout) {
                const int k = nskip + c;
                const int ia = lds_pad(k), ib = lds_pad(k == 0 ? 0 : N - k);
                acc[i] += re[ia] * re[ia] + im[ia] * im[ia] + re[ib] * re[ib] + im[ib] * im[ib];
            }
        }
        __syncthreads();
    }
    int64_t orow = row0 + row;
    if (ring_rows > 0) orow %= ring_rows;
    float* dst = fb + orow * (int64_t)nchanout;
#pragma unroll
    for (int i = 0; i < CPT; i++) {
        const int c = tid + GEN_BLOCK * i;
        if (c < nchanout) dst[c] = acc[i] * scale;
    }
}

}  // namespace

struct vfrb_fb_plan {
    int N = 0, nave = 0, nskip = 0, nchanout = 0;
    bool pow2 = false;
    GenPlan gen{};
    float2* d_tw = nullptr;
    size_t lds_bytes = 0;
    int block = 256;
    const void* kernel = nullptr;
};

namespace {

template <int N>
void select_pow2(vfrb_fb_plan* p) {
    p->kernel = reinterpret_cast<const void*>(&fb_pow2_kernel<N>);
    p->lds_bytes = 2 * sizeof(float) * P2<N>::LDS_WORDS;
    p->block = P2<N>::BLOCK;
}

bool factorise(int n, GenPlan* g) {
    g->npass = 0;
    const int radices[] = {8, 5, 4, 3, 2};
    for (int r : radices) {
        while (n % r == 0) {
            if (g->npass >= GEN_MAXPASS) return false;
            g->radix[g->npass++] = r;
            n /= r;
        }
    }
    return n == 1;
}

}  // namespace

extern "C" {

vfrb_status vfrb_fb_plan_create(vfrb_fb_plan** plan, int32_t twonchan, int32_t nave, int32_t nskip,
                                int32_t nchanout) {
    VFRB_REQUIRE(plan, VFRB_ERR_INVALID, "plan is NULL");
    *plan = nullptr;
    VFRB_REQUIRE(twonchan >= 16 && twonchan % 2 == 0, VFRB_ERR_INVALID, "twonchan=%d", twonchan);
    VFRB_REQUIRE(nave >= 1, VFRB_ERR_INVALID, "nave=%d", nave);
    VFRB_REQUIRE(nskip >= 1 && nchanout >= 1 && nskip + nchanout - 1 <= twonchan / 2,
                 VFRB_ERR_INVALID, "bins %d..%d outside 1..%d", nskip, nskip + nchanout - 1,
                 twonchan / 2);
    vfrb_fb_plan* p = new vfrb_fb_plan();
    p->N = twonchan;
    p->nave = nave;
    p->nskip = nskip;
    p->nchanout = nchanout;
    const int N = twonchan;
    p->pow2 = (N & (N - 1)) == 0 && N >= 256 && N <= 16384;
    if (p->pow2) {
        switch (N) {
            case 256: select_pow2<256>(p); break;
            case 512: select_pow2<512>(p); break;
            case 1024: select_pow2<1024>(p); break;
            case 2048: select_pow2<2048>(p); break;
            case 4096: select_pow2<4096>(p); break;
            case 8192: select_pow2<8192>(p); break;
            default: select_pow2<16384>(p); break;
        }
    } else {
        if (N > GEN_NMAX || !factorise(N, &p->gen)) {
            delete p;
            vfrb::set_error("FFT length %d: need a power of two <= 16384 or factors {2,3,4,5,8} and <= %d",
                            N, GEN_NMAX);
            return VFRB_ERR_UNSUPPORTED;
        }
        p->kernel = reinterpret_cast<const void*>(&fb_generic_kernel);
        p->lds_bytes = 2 * sizeof(float) * (size_t)(N + N / 32);
        p->block = GEN_BLOCK;
    }
    // Twiddles exp(-2*pi*i*m/N), m < N, computed in FP64 and rounded once.
    std::vector<float2> tw(N);
    for (int m = 0; m < N; m++) {
        const double a = -2.0 * M_PI * (double)m / (double)N;
        tw[m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    hipError_t e = hipMalloc(&p->d_tw, sizeof(float2) * N);
    if (e == hipSuccess) e = hipMemcpy(p->d_tw, tw.data(), sizeof(float2) * N, hipMemcpyHostToDevice);
    if (e == hipSuccess && p->lds_bytes > 65536)
        e = hipFuncSetAttribute(p->kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_bytes);
    if (e != hipSuccess) {
        if (p->d_tw) (void)hipFree(p->d_tw);
        delete p;
        return vfrb::hip_fail(e, "vfrb_fb_plan_create", __FILE__, __LINE__);
    }
    *plan = p;
    return VFRB_OK;
}

vfrb_status vfrb_fb_plan_destroy(vfrb_fb_plan* p) {
    if (!p) return VFRB_OK;
    if (p->d_tw) VFRB_HIP(hipFree(p->d_tw));
    delete p;
    return VFRB_OK;
}

vfrb_status vfrb_fb_execute(vfrb_fb_plan* p, const uint8_t* pol0, const uint8_t* pol1, int64_t nrows,
                            float* fb, int64_t ring_rows, int64_t row0, void* stream) {
    VFRB_REQUIRE(p && pol0 && pol1 && fb, VFRB_ERR_INVALID, "NULL argument");
    VFRB_REQUIRE(nrows >= 0 && row0 >= 0 && ring_rows >= 0, VFRB_ERR_INVALID, "negative size");
    VFRB_REQUIRE(ring_rows == 0 || nrows <= ring_rows, VFRB_ERR_INVALID,
                 "nrows=%lld exceeds ring of %lld rows", (long long)nrows, (long long)ring_rows);
    VFRB_REQUIRE(nrows <= INT32_MAX, VFRB_ERR_INVALID, "nrows too large for one launch");
    if (nrows == 0) return VFRB_OK;
    if (p->pow2) {
        VFRB_REQUIRE(((uintptr_t)pol0 % 16) == 0 && ((uintptr_t)pol1 % 16) == 0, VFRB_ERR_INVALID,
                     "pol buffers must be 16-byte aligned");
    }
    const float scale = (float)(0.5 / ((double)(p->N / 2 + 1) * (double)p->nave));
    hipStream_t s = vfrb::as_stream(stream);
    dim3 grid((unsigned)nrows), block(p->block);
    if (p->pow2) {
        void* args[] = {(void*)&pol0, (void*)&pol1, (void*)&p->d_tw, (void*)&fb,  (void*)&nrows,
                        (void*)&p->nave, (void*)&p->nskip, (void*)&p->nchanout, (void*)&ring_rows,
                        (void*)&row0, (void*)&scale};
        VFRB_HIP(hipLaunchKernel(p->kernel, grid, block, args, p->lds_bytes, s));
    } else {
        hipLaunchKernelGGL(fb_generic_kernel, grid, block, p->lds_bytes, s, pol0, pol1,
                           (const float2*)p->d_tw, p->gen, p->N, fb, nrows, p->nave, p->nskip,
                           p->nchanout, ring_rows, row0, scale);
        VFRB_LAUNCH_CHECK("fb_generic_kernel");
    }
    return VFRB_OK;
}

}  // extern "C"
