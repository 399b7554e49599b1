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

// Additive LDS padding i + M*(i >> S): the constant strides of a pass stay
// constant offsets, so the padded address folds into ds_read/ds_write
// immediates.
template <int S, int M>
struct Pad {
    static constexpr int SPAN = 1 << S;
    __device__ static __forceinline__ int idx(int i) { return i + M * (i >> S); }
    // idx(x + c) == idx(x) + off(c) whenever (x mod SPAN) + (c mod SPAN) <
    // SPAN: a thread's padded base plus a constant, i.e. an immediate offset.
    static constexpr int off(int c) { return c + M * (c >> S); }
    static constexpr int words(int n) { return n + M * ((n - 1) >> S) + 1; }
};

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
// Twiddle constants w_R^m = exp(-2*pi*i*m/R) for the in-register radix-16
// and radix-32 DFTs, from quarter-wave cosine tables (exact 0 and 1).
// cos(2*pi*i/32) for i = 0..8 (a quarter wave); R = 16 uses every other entry.
__device__ __forceinline__ constexpr float qcos32(int i) {
    switch (i) {
        case 0: return 1.0f;
        case 1: return 0.98078528040323044913f;
        case 2: return 0.92387953251128675613f;
        case 3: return 0.83146961230254523708f;
        case 4: return 0.70710678118654752440f;
        case 5: return 0.55557023301960222474f;
        case 6: return 0.38268343236508977173f;
        case 7: return 0.19509032201612826785f;
        default: return 0.0f;
    }
}
template <int R>
__device__ __forceinline__ constexpr float cos_r(int m) {  // cos(2*pi*m/R), R | 32
    constexpr int Q = R / 4;
    constexpr int S = 32 / R;
    m = ((m % R) + R) % R;
    if (m <= Q) return qcos32(S * m);
    if (m <= 2 * Q) return -qcos32(S * (2 * Q - m));
    if (m <= 3 * Q) return -qcos32(S * (m - 2 * Q));
    return qcos32(S * (4 * Q - m));
}
template <int R>
__device__ __forceinline__ cf wconst(int m) {  // exp(-2*pi*i*m/R)
    return cf{cos_r<R>(m), -cos_r<R>(m - R / 4)};
}

template <int R>
__device__ __forceinline__ void dft(cf* v);

// R = R1*R2 DFT: R1 DFTs of length R2 over n2, twiddles w_R^(n1*k2), R2 DFTs
// of length R1 over n1; output k = k2 + R2*k1 (natural order).
template <int R1, int R2>
__device__ __forceinline__ void dft_split(cf* v) {
    constexpr int R = R1 * R2;
    cf a[R1][R2];
#pragma unroll
    for (int n1 = 0; n1 < R1; n1++) {
#pragma unroll
        for (int n2 = 0; n2 < R2; n2++) a[n1][n2] = v[n1 + R1 * n2];
        dft<R2>(a[n1]);
    }
#pragma unroll
    for (int n1 = 1; n1 < R1; n1++)
#pragma unroll
        for (int k2 = 1; k2 < R2; k2++) a[n1][k2] = cmul(a[n1][k2], wconst<R>(n1 * k2));
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++) {
        cf b[R1];
#pragma unroll
        for (int n1 = 0; n1 < R1; n1++) b[n1] = a[n1][k2];
        dft<R1>(b);
#pragma unroll
        for (int k1 = 0; k1 < R1; k1++) v[k2 + R2 * k1] = b[k1];
    }
}

template <int R>
__device__ __forceinline__ void dft(cf* v) {
    if constexpr (R == 2) dft2(v);
    else if constexpr (R == 3) dft3(v);
    else if constexpr (R == 4) dft4(v);
    else if constexpr (R == 5) dft5(v);
    else if constexpr (R == 8) dft8(v);
    else if constexpr (R == 16) dft_split<4, 4>(v);
    else dft_split<4, 8>(v);
}

// Multiplies v[r] (r = 1..R-1) by w^r, where w^(2^b) = tw[(2^b)*idx] come
// from the FP64-computed table and the other powers are products of those
// (at most log2(R)-1 multiplications deep).
template <int R>
struct TwSet {
    float2 w[__builtin_ctz(R)];  // w^(2^b), b < log2(R)
};
template <int R>
__device__ __forceinline__ TwSet<R> load_twiddles(const float2* __restrict__ tw, int idx) {
    TwSet<R> s;
#pragma unroll
    for (int b = 0; (1 << b) < R; b++) s.w[b] = tw[(1 << b) * idx];
    return s;
}
// Hides a frame-invariant set's values from the optimiser once per frame, so
// that the products built from it are not hoisted out of the frame loop into
// registers (the asm emits no instruction).
template <int R>
__device__ __forceinline__ void refresh(TwSet<R>& s) {
#pragma unroll
    for (int b = 0; (1 << b) < R; b++) asm volatile("" : "+v"(s.w[b].x), "+v"(s.w[b].y));
}
// Every power w^r, r = 1..R-1, of a set, formed once before the frame loop.
template <int R>
struct TwFull {
    cf w[R];
};
template <int R>
__device__ __forceinline__ TwFull<R> expand(const TwSet<R>& s) {
    TwFull<R> f;
    f.w[0] = cf{1.0f, 0.0f};
#pragma unroll
    for (int b = 0; (1 << b) < R; b++) f.w[1 << b] = cf{s.w[b].x, s.w[b].y};
#pragma unroll
    for (int r = 3; r < R; r++) {
        if ((r & (r - 1)) == 0) continue;
        const int low = r & -r;
        f.w[r] = cmul(f.w[r - low], f.w[low]);
    }
    return f;
}
// The full set when H (the lengths whose registers hold it without spilling),
// else the w^(2^b) table values, whose products are formed in every frame.
template <bool H, int R>
__device__ __forceinline__ auto pick_twiddles(const TwSet<R>& s) {
    if constexpr (H) return expand(s);
    else return s;
}
template <int R>
__device__ __forceinline__ void refresh(TwFull<R>& f) {
#pragma unroll
    for (int r = 1; r < R; r++) asm volatile("" : "+v"(f.w[r].x), "+v"(f.w[r].y));
}
template <int R>
__device__ __forceinline__ void apply_twiddle_set(cf* v, const TwFull<R>& f) {
#pragma unroll
    for (int r = 1; r < R; r++) v[r] = cmul(v[r], f.w[r]);
}
// apply_twiddles below with the table values already in registers.
template <int R>
__device__ __forceinline__ void apply_twiddle_set(cf* v, const TwSet<R>& s) {
    cf w[R];
#pragma unroll
    for (int b = 0; (1 << b) < R; b++) w[1 << b] = cf{s.w[b].x, s.w[b].y};
#pragma unroll
    for (int r = 3; r < R; r++) {
        if ((r & (r - 1)) == 0) continue;
        const int low = r & -r;
        w[r] = cmul(w[r - low], w[low]);
    }
#pragma unroll
    for (int r = 1; r < R; r++) v[r] = cmul(v[r], w[r]);
}
template <int R>
__device__ __forceinline__ void apply_twiddles(cf* v, const float2* __restrict__ tw, int idx) {
    cf w[R];
#pragma unroll
    for (int b = 1; b < R; b <<= 1) {
        const float2 t = tw[b * idx];
        w[b] = cf{t.x, t.y};
    }
#pragma unroll
    for (int r = 3; r < R; r++) {
        if ((r & (r - 1)) == 0) continue;
        const int low = r & -r;
        w[r] = cmul(w[r - low], w[low]);
    }
#pragma unroll
    for (int r = 1; r < R; r++) v[r] = cmul(v[r], w[r]);
}

// Twiddled DFT of a power-of-two length R in one pass of FMA butterflies:
// y_k = sum_r w^r x_r exp(-2 pi i r k / R) as a radix-2 decimation in time
// (inputs in bit-reversed order, compile-time indices).  Its stage of size S
// combines two half-size results of twiddle base w^(R/S) with
// z_k = w^(R/S) exp(-2 pi i k / S), k < S/2:
//     y_k = e_k + z_k o_k,   y_(k+S/2) = 2 e_k - y_k,
// 4 + 2 FMAs per butterfly, against a complex product (4) and two complex
// additions (4) when the twiddles are applied before a plain DFT.  For
// k >= S/4, z_k = z_(k-S/4) (-i), which only swaps operands and signs, so
// a stage needs S/4 stored values (one for S = 2): R/2 per set.
template <int R>
struct TwFused {
    cf t[R / 2];  // stage S: entries [S/4, S/2) for S >= 4, entry 0 for S = 2
};
// From the table values w^(2^b): stage S's z_j = w^(R/S) * exp(-2 pi i j/S).
template <int R>
__device__ __forceinline__ TwFused<R> fuse(const TwSet<R>& s) {
    TwFused<R> f;
    f.t[0] = cf{s.w[__builtin_ctz(R) - 1].x, s.w[__builtin_ctz(R) - 1].y};  // S = 2: w^(R/2)
#pragma unroll
    for (int S = 4; S <= R; S *= 2) {
        const int b = __builtin_ctz(R / S);
        const cf base = cf{s.w[b].x, s.w[b].y};
#pragma unroll
        for (int j = 0; j < S / 4; j++) f.t[S / 4 + j] = j == 0 ? base : cmul(base, wconst<32>(j * (32 / S)));
    }
    return f;
}
template <int R>
__device__ __forceinline__ void refresh(TwFused<R>& f) {
#pragma unroll
    for (int i = 0; i < R / 2; i++) asm volatile("" : "+v"(f.t[i].x), "+v"(f.t[i].y));
}
__host__ __device__ constexpr int bitrev(int x, int bits) {
    int y = 0;
    for (int i = 0; i < bits; i++) y |= ((x >> i) & 1) << (bits - 1 - i);
    return y;
}
template <int R>
__device__ __forceinline__ void tdft(cf* v, const TwFused<R>& f) {
    constexpr int LB = __builtin_ctz(R);
    cf a[R];
#pragma unroll
    for (int r = 0; r < R; r++) a[r] = v[bitrev(r, LB)];
#pragma unroll
    for (int S = 2; S <= R; S *= 2) {
#pragma unroll
        for (int blk = 0; blk < R; blk += S) {
#pragma unroll
            for (int k = 0; k < S / 2; k++) {
                const int q = S == 2 ? 0 : S / 4 + (k % (S / 4));
                const bool rot = S >= 4 && k >= S / 4;  // z_k = z_(k-S/4) (-i)
                const cf z = f.t[q];
                const cf e = a[blk + k], o = a[blk + k + S / 2];
                cf y;
                if (!rot) {  // e + z o
                    y.x = fmaf(z.x, o.x, fmaf(-z.y, o.y, e.x));
                    y.y = fmaf(z.x, o.y, fmaf(z.y, o.x, e.y));
                } else {     // e - i z o
                    y.x = fmaf(z.x, o.y, fmaf(z.y, o.x, e.x));
                    y.y = fmaf(-z.x, o.x, fmaf(z.y, o.y, e.y));
                }
                a[blk + k] = y;
                a[blk + k + S / 2] = cf{fmaf(2.0f, e.x, -y.x), fmaf(2.0f, e.y, -y.y)};
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = a[r];
}
// A pass's twiddles as the fused set: formed once (H) or every frame.
template <bool H, int R>
__device__ __forceinline__ auto pick_fused(const TwSet<R>& s) {
    if constexpr (H) return fuse(s);
    else return s;
}
template <int R>
__device__ __forceinline__ void twiddled_dft(cf* v, const TwFused<R>& f) {
    tdft<R>(v, f);
}
template <int R>
__device__ __forceinline__ void twiddled_dft(cf* v, const TwSet<R>& s) {
    tdft<R>(v, fuse(s));
}

// ---------------------------------------------------------------------------
// Power-of-two kernel.  32 complex elements per thread, TPF = N/32 threads per
// FFT, FPB = BLOCK/TPF FFTs of the same output row in flight per workgroup.
// Pass plan: radix 8 straight from the byte payload (4 consecutive
// butterflies per thread, dword loads), radix-R middle passes through LDS,
// and a final radix-16 pass in which thread t owns the butterflies t and
// N/16 - t, so Z[k] and Z[N-k] meet in registers and the detected power
// never goes back to LDS.
// ---------------------------------------------------------------------------
template <int N>
struct P2 {
    static constexpr int TPF = N / 32;
    static constexpr int BLOCK = TPF > 256 ? TPF : 256;
    static constexpr int FPB = BLOCK / TPF;
    static constexpr int TOT = FPB * N;
    static constexpr int LOG2 = __builtin_ctz(N);
    static constexpr int MID_BITS = LOG2 - 7;             // between radix-8 and radix-16
    // middle radices (0 = unused): 8192 -> 8,8; 16384 -> 16,8; else one pass
    static constexpr int MID0 = MID_BITS == 6 ? 8 : MID_BITS == 7 ? 16 : (1 << MID_BITS);
    static constexpr int MID1 = MID_BITS == 6 ? 8 : MID_BITS == 7 ? 8 : 0;
    // One padding per LDS generation (written by one pass, read by the
    // next), each chosen by enumerating the wave access patterns of its
    // writer and reader against 32 four-byte banks: A = pass 1 -> first
    // middle pass, B = between the two middle passes, C = -> last pass.
    using PadA = Pad<5, 1>;
    using PadB = Pad<5, 2>;
    using PadC = Pad<(N >= 8192 ? 8 : 5), (N >= 8192 ? 1 : N == 2048 ? 2 : N == 1024 ? 3 : 1)>;
    static_assert(MID0 > 1, "the power-of-two kernel needs at least one middle pass");
    // All twiddle powers held in registers across the frame loop (products
    // formed once) where that fits the 256-VGPR budget without spilling:
    // 8192 points 3.84 -> 3.72 ms per 20 s segment, 2048 points -3 %.
    static constexpr bool HOIST = N == 1024 || N == 2048 || N == 8192;
    // Separate |Z|^2 sums of the last pass's two butterflies (16 more VGPRs)
    static constexpr bool SPLIT = N >= 1024;
    static constexpr int cmax(int a, int b) { return a > b ? a : b; }
    static constexpr int LDS_WORDS =                      // per component
        cmax(PadA::words(TOT), cmax(MID1 > 1 ? PadB::words(TOT) : 0, PadC::words(TOT)));
};

// One radix-R Stockham pass through LDS (NS = product of the radices
// before it); t = thread index within its FFT, fb = element offset of that
// FFT in LDS (a multiple of N), ws = the twiddles of butterfly t mod NS
// (the same for all of the thread's butterflies); reads in padding PR,
// writes in PW.  Where
// the constant part of an index cannot carry into the padding's shift
// (checked at compile time), the address is the thread's padded base plus
// an immediate.
template <int N, int R, int NS, class PR, class PW, class TW>
__device__ __forceinline__ void p2_pass(float* __restrict__ re, float* __restrict__ im, int t,
                                        int fb, const TW& ws) {
    constexpr int TPF = P2<N>::TPF;
    constexpr int Q = 32 / R;
    constexpr int STRIDE = N / R;
    // reads fb + t + (TPF*q + r*STRIDE): the constant is a multiple of SPAN
    constexpr bool FOLD_RD = TPF % PR::SPAN == 0 && STRIDE % PR::SPAN == 0;
    // writes fb + (t-k)*R + k + (TPF*q*R + r*NS), k = t mod NS < NS: the
    // base is k mod SPAN, the constant's residue at most SPAN - NS
    constexpr bool FOLD_WR = TPF % NS == 0 && (NS * R) % PW::SPAN == 0 && (TPF * R) % PW::SPAN == 0;
    cf v[Q][R];
    const int xr = PR::idx(fb + t);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int j = t + TPF * q;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int idx = FOLD_RD ? xr + PR::off(TPF * q + r * STRIDE) : PR::idx(fb + j + r * STRIDE);
            v[q][r] = cf{re[idx], im[idx]};
        }
    }
    __syncthreads();
    static_assert(TPF % NS == 0, "every butterfly group of a thread has the same twiddle index");
    const int k0 = t & (NS - 1);
    const int xw = PW::idx(fb + (t - k0) * R + k0);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int j = t + TPF * q;
        const int k = j & (NS - 1);
        twiddled_dft<R>(v[q], ws);  // k == t mod NS for every q
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int idx = FOLD_WR ? xw + PW::off(TPF * q * R + r * NS) : PW::idx(fb + base + r * NS);
            re[idx] = v[q][r].x;
            im[idx] = v[q][r].y;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ float norm2(cf a) { return a.x * a.x + a.y * a.y; }

// LDS (70 KB for N <= 8192) allows two workgroups per CU, i.e. 2 waves per
// SIMD: hold the register allocator to that budget.
// Raw VDIF frames (FRAMED): pol0 points at frame pairs (thread 0 then 1,
// 5032 bytes each), table[s] is the pair feeding payload slot s (5000 bytes
// per pol) or -1 for a zero-filled slot, byte0 is the payload-stream offset
// of row 0, and pol1 points at one frame pair of zero bytes.  Otherwise
// pol0/pol1 are the de-framed payload streams.
constexpr int kFrameBytes = 5032, kHeaderBytes = 32, kPayloadBytes = 5000;
constexpr int kFbBufferFlags = 0x00020000;  // word 3 of a gfx9 raw buffer resource

template <int N, bool FRAMED>
__global__ void __launch_bounds__(P2<N>::BLOCK) __attribute__((amdgpu_waves_per_eu(2, 2)))
fb_pow2_kernel(const uint8_t* __restrict__ pol0, const uint8_t* __restrict__ pol1,
               const float2* __restrict__ tw, float* __restrict__ fb, int64_t nrows, int nave,
               int nskip, int nchanout, int64_t ring_rows, int64_t row0, float scale,
               const int32_t* __restrict__ table, int64_t byte0) {
    using C = P2<N>;
    constexpr int TPF = C::TPF;
    constexpr int NS = N / 16;      // butterflies of the last pass
    constexpr int EIGHTH = N / 8;   // input stride of the first pass
    extern __shared__ float lds[];
    float* re = lds;
    float* im = lds + C::LDS_WORDS;
    const int tid = threadIdx.x;
    const int64_t row = blockIdx.x;
    if (row >= nrows) return;
    // the FFT slot of the workgroup; wave-uniform where an FFT spans whole waves
    const int slot = TPF >= 64 ? __builtin_amdgcn_readfirstlane(tid / TPF) : tid / TPF;
    const int t = tid % TPF;
    const int fbase = slot * N;
    const uint8_t* zero = pol1;  // FRAMED: a zeroed frame pair, read for missing slots

    // Last-pass butterflies of this thread and their twiddles (frame invariant).
    const int ja = t == 0 ? 0 : t;
    const int jb = t == 0 ? N / 32 : NS - t;

    // Where registers allow (C::SPLIT), |Z|^2 of the two last-pass butterflies
    // is summed over the frames separately (sa, sb: two FMAs per bin and
    // frame) and the conjugate bins are paired once after the frame loop;
    // otherwise acc sums the pairs every frame.
    float acc[16], sa[C::SPLIT ? 16 : 1], sb[C::SPLIT ? 16 : 1];
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = 0.0f;
#pragma unroll
    for (int i = 0; i < (C::SPLIT ? 16 : 1); i++) sa[i] = sb[i] = 0.0f;

    // prefetch registers: 8 dwords per pol = samples 4t..4t+3 + r*N/8
    uint32_t x0[8], x1[8];
    // FRAMED, TPF >= 64: payload slot fs and byte fw in it where the wave's
    // next frame starts (wave-uniform; one FP64 division per row, then steps)
    int64_t fs = 0;
    int fw = 0;
    // and the table entries of the slots that frame spans, loaded one frame
    // ahead so that no wave waits for them
    constexpr int FMAXS = (N + kPayloadBytes - 1) / kPayloadBytes + 1;
    int tk[FMAXS];
    auto load_tk = [&]() {
        const int ns = (fw + N - 1) / kPayloadBytes;
#pragma unroll
        for (int j = 0; j < FMAXS; j++) tk[j] = j <= ns ? table[fs + j] : -1;
    };
    if constexpr (FRAMED && TPF >= 64) {
        const int64_t of = byte0 + ((int64_t)row * nave + slot) * N;
        int64_t s = (int64_t)((double)of * (1.0 / kPayloadBytes));
        int w = (int)(of - s * kPayloadBytes);
        if (w < 0) {
            s--;
            w += kPayloadBytes;
        } else if (w >= kPayloadBytes) {
            s++;
            w -= kPayloadBytes;
        }
        fs = ((int64_t)__builtin_amdgcn_readfirstlane((int)(s >> 32)) << 32) |
             (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)s);
        fw = __builtin_amdgcn_readfirstlane(w);
        if (slot < nave) load_tk();
    }
    auto fetch = [&](int it) {
        const int frame = it + slot;
        if (frame >= nave) {
#pragma unroll
            for (int r = 0; r < 8; r++) {  // 128 -> 0
                x0[r] = 0x80808080u;
                x1[r] = 0x80808080u;
            }
            return;
        }
        if constexpr (FRAMED && TPF >= 64) {
            // The FFT frame's N bytes span at most MAXS payload slots, from
            // slot fs on (byte fw of it), whose frame pairs tk[] holds.  The
            // slots' addresses are wave-uniform, formed once per frame; each
            // lane then picks, per dword, one of them by comparing its byte
            // position with the slot edges.
            constexpr int MAXS = (N + kPayloadBytes - 1) / kPayloadBytes + 1;
            static_assert(MAXS == FMAXS, "one slot count");
            const int w0 = fw;
            const int ns = (w0 + N - 1) / kPayloadBytes;  // last slot of the frame, < MAXS
            const uint32_t p0 = (uint32_t)w0 + 4u * (uint32_t)t;
            // Lowest and highest frame pair of the frame's slots: where no
            // slot is missing and they lie within 400000 pairs (4 GB), a
            // lane's address is one scalar base plus a 32-bit offset.
            int kmin = tk[0], kmax = tk[0];
#pragma unroll
            for (int j = 1; j < MAXS; j++)
                if (j <= ns) {
                    kmin = min(kmin, tk[j]);
                    kmax = max(kmax, tk[j]);
                }
            if (kmin >= 0 && kmax - kmin < 400000) {
                // slot j's payload byte p (p counted from the frame's slot 0)
                // is at lo + d[j] + p; d[j] < 2^32 - 2^22
                const uint8_t* lo = pol0 + (int64_t)kmin * (2 * kFrameBytes) + kHeaderBytes -
                                    (MAXS - 1) * kPayloadBytes;
                uint32_t d[MAXS];
#pragma unroll
                for (int j = 0; j < MAXS; j++)
                    d[j] = j <= ns ? (uint32_t)(tk[j] - kmin) * (uint32_t)(2 * kFrameBytes) +
                                         (uint32_t)((MAXS - 1 - j) * kPayloadBytes)
                                   : 0u;
                // one raw buffer per pol based at lo: the lane's 32-bit offset,
                // r*N/8 as scalar and instruction offsets
                const __amdgpu_buffer_rsrc_t rs0 =
                    __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(lo), 0, -1, kFbBufferFlags);
                const __amdgpu_buffer_rsrc_t rs1 = __builtin_amdgcn_make_buffer_rsrc(
                    const_cast<uint8_t*>(lo + kFrameBytes), 0, -1, kFbBufferFlags);
#pragma unroll
                for (int r = 0; r < 8; r++) {
                    uint32_t o = d[0];
#pragma unroll
                    for (int m = 1; m < MAXS; m++) {
                        const int th = m * kPayloadBytes - r * EIGHTH;  // p0 >= th: slot m or later
                        if (th <= 0 || p0 >= (uint32_t)th) o = d[m];
                    }
                    const int v = (int)(p0 + o + (uint32_t)((r & 3) * EIGHTH));
                    const int so = (r >> 2) * 4 * EIGHTH;
                    x0[r] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rs0, v, so, 0);
                    x1[r] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rs1, v, so, 0);
                }
            } else {
                // missing slots (the zero pair) or pairs far apart: 64-bit
                // addresses, slot j's byte p at base[j] + p
                const uint8_t* base[MAXS];
#pragma unroll
                for (int j = 0; j < MAXS; j++)
                    base[j] = (tk[j] >= 0 ? pol0 + (int64_t)tk[j] * (2 * kFrameBytes) : zero) + kHeaderBytes -
                              j * kPayloadBytes;
#pragma unroll
                for (int r = 0; r < 8; r++) {
                    const uint32_t p = p0 + (uint32_t)(r * EIGHTH);
                    const uint8_t* b = base[0];
#pragma unroll
                    for (int m = 1; m < MAXS; m++) b = p >= (uint32_t)(m * kPayloadBytes) ? base[m] : b;
                    x0[r] = *reinterpret_cast<const uint32_t*>(b + p);
                    x1[r] = *reinterpret_cast<const uint32_t*>(b + p + kFrameBytes);
                }
            }
            // the wave's next frame, and its table entries (read now, used in
            // the next call)
            fw += C::FPB * N;
            const int q = fw / kPayloadBytes;
            fs += q;
            fw -= q * kPayloadBytes;
            if (frame + C::FPB < nave) load_tk();
        } else if constexpr (FRAMED) {
            // payload offset -> (slot, byte): one FP64 division, then steps of
            // N/8 < 5000 bytes carry into the next slot at most once
            const int64_t o = byte0 + ((int64_t)row * nave + frame) * N + 4 * t;
            int64_t sl = (int64_t)((double)o * (1.0 / kPayloadBytes));
            int w = (int)(o - sl * kPayloadBytes);
            if (w < 0) {
                sl--;
                w += kPayloadBytes;
            } else if (w >= kPayloadBytes) {
                sl++;
                w -= kPayloadBytes;
            }
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const int k = table[sl];
                uint32_t a0 = 0u, a1 = 0u;  // zero-filled payload of a missing frame
                if (k >= 0) {
                    const uint8_t* f = pol0 + (int64_t)k * (2 * kFrameBytes) + kHeaderBytes + w;
                    a0 = *reinterpret_cast<const uint32_t*>(f);
                    a1 = *reinterpret_cast<const uint32_t*>(f + kFrameBytes);
                }
                x0[r] = a0;
                x1[r] = a1;
                w += EIGHTH;
                if (w >= kPayloadBytes) {
                    w -= kPayloadBytes;
                    sl++;
                }
            }
        } else {
            const uint8_t* b0 = pol0 + ((size_t)row * nave + frame) * N + 4 * t;
            const uint8_t* b1 = pol1 + ((size_t)row * nave + frame) * N + 4 * t;
#pragma unroll
            for (int r = 0; r < 8; r++) {
                x0[r] = *reinterpret_cast<const uint32_t*>(b0 + r * EIGHTH);
                x1[r] = *reinterpret_cast<const uint32_t*>(b1 + r * EIGHTH);
            }
        }
    };
    // The middle and last passes' twiddles depend only on the thread index:
    // loaded once (14 complex values at 8192 points), so that inside the
    // frame loop the next frame's bytes are the only vector loads, and no
    // wait for a twiddle also waits for them.  Pass p's butterfly
    // k = t mod NS_p uses w^(2^b) = tw[(2^b) * (k << s_p)], N = NS_p * R_p << s_p.
    constexpr int NS2 = 8 * C::MID0;
    TwSet<(C::MID1 > 1 ? C::MID1 : 2)> tw_m10{};
    if constexpr (C::MID1 > 1)
        tw_m10 = load_twiddles<C::MID1>(tw, (t & (NS2 - 1)) << (C::LOG2 - __builtin_ctz(NS2 * C::MID1)));
    auto tw_m0 = pick_fused<C::HOIST>(
        load_twiddles<C::MID0>(tw, (t & 7) << (C::LOG2 - __builtin_ctz(8 * C::MID0))));
    auto tw_m1 = pick_fused<C::HOIST>(tw_m10);
    auto tw_a = pick_fused<C::HOIST>(load_twiddles<16>(tw, ja));
    auto tw_b = pick_fused<C::HOIST>(load_twiddles<16>(tw, jb));
    fetch(0);
    for (int it = 0; it < nave; it += C::FPB) {
        refresh(tw_m0);
        if constexpr (C::MID1 > 1) refresh(tw_m1);
        refresh(tw_a);
        refresh(tw_b);
        // ---- pass 1: radix 8, butterflies 4t+q, inputs z[4t+q + r*N/8].
        // The offset-binary bytes enter unshifted: the radix-8 DFT reaches
        // every output but the first through differences of its inputs, which
        // cancel the +128 exactly (integers below 2^24), so only output 0
        // carries 8*128*(1+i), removed after the DFT: bit-identical to
        // subtracting 128 from every sample.
        cf v[4][8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                v[q][r] = cf{(float)((x0[r] >> (8 * q)) & 0xffu), (float)((x1[r] >> (8 * q)) & 0xffu)};
            }
        }
        if (it + C::FPB < nave) fetch(it + C::FPB);  // next frame's bytes in flight
#pragma unroll
        for (int q = 0; q < 4; q++) {
            dft8(v[q]);
            v[q][0].x -= 1024.0f;
            v[q][0].y -= 1024.0f;
        }
        __syncthreads();  // previous frame's last pass has read LDS
        const int xa = C::PadA::idx(fbase + 32 * t);  // 32t + (8q + r < 32): no carry
#pragma unroll
        for (int q = 0; q < 4; q++) {
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const int idx = xa + C::PadA::off(8 * q + r);
                re[idx] = v[q][r].x;
                im[idx] = v[q][r].y;
            }
        }
        __syncthreads();
        // ---- middle passes
        if constexpr (C::MID1 > 1) {
            p2_pass<N, C::MID0, 8, typename C::PadA, typename C::PadB>(re, im, t, fbase, tw_m0);
            p2_pass<N, C::MID1, 8 * C::MID0, typename C::PadB, typename C::PadC>(re, im, t, fbase, tw_m1);
        } else {
            p2_pass<N, C::MID0, 8, typename C::PadA, typename C::PadC>(re, im, t, fbase, tw_m0);
        }
        // last pass reads base + r*NS: NS a multiple of the padding span folds
        constexpr bool FOLD_LAST = NS % C::PadC::SPAN == 0;
        // ---- last pass: radix 16 on butterflies ja, jb (Ns == N/16); only
        // |Z|^2 is needed: na[r] = |Z[ja + r*NS]|^2, nb[r] = |Z[jb + r*NS]|^2
        float na[16], nb[16];
        {
            cf a[16];
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int ia = FOLD_LAST ? C::PadC::idx(fbase + ja) + C::PadC::off(r * NS)
                                         : C::PadC::idx(fbase + ja + r * NS);
                a[r] = cf{re[ia], im[ia]};
            }
            twiddled_dft<16>(a, tw_a);
#pragma unroll
            for (int r = 0; r < 16; r++) {
                if constexpr (C::SPLIT) sa[r] = fmaf(a[r].x, a[r].x, fmaf(a[r].y, a[r].y, sa[r]));
                else na[r] = norm2(a[r]);
            }
        }
        {
            cf b[16];
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int ib = FOLD_LAST ? C::PadC::idx(fbase + jb) + C::PadC::off(r * NS)
                                         : C::PadC::idx(fbase + jb + r * NS);
                b[r] = cf{re[ib], im[ib]};
            }
            twiddled_dft<16>(b, tw_b);
#pragma unroll
            for (int r = 0; r < 16; r++) {
                if constexpr (C::SPLIT) sb[r] = fmaf(b[r].x, b[r].x, fmaf(b[r].y, b[r].y, sb[r]));
                else nb[r] = norm2(b[r]);
            }
        }
        if constexpr (C::SPLIT) continue;
        // ---- detect: |Z[k]|^2 + |Z[N-k]|^2
        if (t != 0) {
#pragma unroll
            for (int r = 0; r < 8; r++) {
                acc[r] += na[r] + nb[15 - r];      // k = ja + r*NS
                acc[8 + r] += nb[r] + na[15 - r];  // k = jb + r*NS
            }
        } else {
#pragma unroll
            for (int r = 0; r < 8; r++) {
                acc[r] += nb[r] + nb[15 - r];          // k = N/32 + r*NS
                acc[8 + r] += na[r + 1] + na[15 - r];  // k = (r+1)*NS
            }
        }
    }
    if constexpr (C::SPLIT) {  // pair the conjugate bins: |Z[k]|^2 + |Z[N-k]|^2
        if (t != 0) {
#pragma unroll
            for (int r = 0; r < 8; r++) {
                acc[r] = sa[r] + sb[15 - r];      // k = ja + r*NS
                acc[8 + r] = sb[r] + sa[15 - r];  // k = jb + r*NS
            }
        } else {
#pragma unroll
            for (int r = 0; r < 8; r++) {
                acc[r] = sb[r] + sb[15 - r];          // k = N/32 + r*NS
                acc[8 + r] = sa[r + 1] + sa[15 - r];  // k = (r+1)*NS
            }
        }
    }
    // ---- sum the FFT slots of this row, then write the selected channels:
    // accumulator i of butterfly thread u holds bin k(i, u)
    int64_t orow = row0 + row;
    if (ring_rows > 0) orow %= ring_rows;
    float* dst = fb + orow * (int64_t)nchanout;
    auto bin = [&](int i, int u) {
        const int r = i & 7;
        if (u != 0) return (i < 8 ? u : NS - u) + r * NS;
        return i < 8 ? N / 32 + r * NS : (r + 1) * NS;
    };
    if constexpr (C::FPB > 1) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; i++) re[i * C::BLOCK + tid] = acc[i];
        __syncthreads();
        // one (accumulator, butterfly thread) pair per thread, summed over the
        // FFT slots in slot order
        for (int o = tid; o < 16 * TPF; o += C::BLOCK) {
            const int i = o / TPF, u = o % TPF;
            float s = 0.0f;
#pragma unroll
            for (int f = 0; f < C::FPB; f++) s += re[i * C::BLOCK + f * TPF + u];
            const int c = bin(i, u) - nskip;
            if (c >= 0 && c < nchanout) dst[c] = s * scale;
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int c = bin(i, t) - nskip;
        if (c >= 0 && c < nchanout) dst[c] = acc[i] * scale;
    }
}

// ---------------------------------------------------------------------------
// Mixed-radix kernel for other even lengths (factors 2,3,4,5,8), one FFT per
// workgroup of 256 threads, N <= 16384.
// ---------------------------------------------------------------------------
constexpr int GEN_BLOCK = 1024;
// Threads of a fixed-length workgroup.  At 12500 points, 640 and 768 threads
// (fewer idle lanes in the 2500- and 3125-butterfly passes) ran 0.97 and
// 0.95 ms per data-second against 0.92 for 1024.
constexpr int FIXED_BLOCK = 1024;
constexpr int GEN_NMAX = 16384;
constexpr int GEN_MAXPASS = 16;

struct GenPlan {
    int npass;
    int radix[GEN_MAXPASS];
};

template <int NMAX, int R>
__device__ __forceinline__ void gen_pass(float* __restrict__ re, float* __restrict__ im, int N, int Ns,
                                         const float2* __restrict__ tw, int tid) {
    constexpr int MAXQ = (NMAX / R + GEN_BLOCK - 1) / GEN_BLOCK;
    const int nb = N / R;
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

// The same pass with every size a constant: a fixed length N, the product NS
// of the radices before it, radix R.  Butterfly counts, trip counts, the
// division by NS and the twiddle stride fold at compile time, so a pass holds
// only its samples in registers.  The operations are gen_pass's, in the same
// order (the compiler may fuse a different set of them into FMAs).
template <int N, int BLOCK, int NS, int R>
__device__ __forceinline__ void fixed_pass(float* __restrict__ re, float* __restrict__ im,
                                           const float2* __restrict__ tw, int tid) {
    constexpr int NB = N / R;
    constexpr int Q = (NB + BLOCK - 1) / BLOCK;
    constexpr int TWSTEP = N / (NS * R);
    cf v[Q][R];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int j = tid + BLOCK * q;
        if (NB % BLOCK == 0 || j < NB) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int idx = lds_pad(j + r * NB);
                v[q][r] = cf{re[idx], im[idx]};
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int j = tid + BLOCK * q;
        if (NB % BLOCK == 0 || j < NB) {
            const int k = j % NS;
#pragma unroll
            for (int r = 1; r < R; r++) {
                const float2 w = tw[r * k * TWSTEP];
                v[q][r] = cmul(v[q][r], cf{w.x, w.y});
            }
            dft<R>(v[q]);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int idx = lds_pad(base + r * NS);
                re[idx] = v[q][r].x;
                im[idx] = v[q][r].y;
            }
        }
    }
    __syncthreads();
}

// The first pass of a fixed length reads its radix-R inputs straight from
// the frame's bytes (z = (x0-128) + i(x1-128)), with no staging pass through
// LDS.  With NS = 1 every twiddle is 1, so none is applied.
template <int N, int BLOCK, int R>
__device__ __forceinline__ void fixed_first_pass(const uint8_t* __restrict__ x0, const uint8_t* __restrict__ x1,
                                                 float* __restrict__ re, float* __restrict__ im, int tid) {
    constexpr int NB = N / R;
    constexpr int Q = (NB + BLOCK - 1) / BLOCK;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int j = tid + BLOCK * q;
        if (NB % BLOCK == 0 || j < NB) {
            cf v[R];
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = cf{(float)x0[j + r * NB] - 128.0f, (float)x1[j + r * NB] - 128.0f};
            dft<R>(v);
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int idx = lds_pad(j * R + r);
                re[idx] = v[r].x;
                im[idx] = v[r].y;
            }
        }
    }
    __syncthreads();
}

template <int N, int BLOCK, int NS, int R, int... Rest>
__device__ __forceinline__ void fixed_passes(float* __restrict__ re, float* __restrict__ im,
                                             const float2* __restrict__ tw, int tid) {
    fixed_pass<N, BLOCK, NS, R>(re, im, tw, tid);
    if constexpr (sizeof...(Rest) > 0) fixed_passes<N, BLOCK, NS * R, Rest...>(re, im, tw, tid);
}

template <int N, int BLOCK, int R0, int... Rest>
__device__ __forceinline__ void fixed_rest(const uint8_t* __restrict__ x0, const uint8_t* __restrict__ x1,
                                           float* __restrict__ re, float* __restrict__ im,
                                           const float2* __restrict__ tw, int tid) {
    fixed_first_pass<N, BLOCK, R0>(x0, x1, re, im, tid);
    if constexpr (sizeof...(Rest) > 0) fixed_passes<N, BLOCK, R0, Rest...>(re, im, tw, tid);
}

// One output row per workgroup: for each of the row's nave frames, both pols
// as z = (x0-128) + i(x1-128), the complex FFT by `passes` (in LDS), then
// |Z[k]|^2 + |Z[N-k]|^2 = |X0[k]|^2 + |X1[k]|^2 (times 2) accumulated per
// selected channel.
template <int BLOCK, int NMAX, class Passes>
__device__ __forceinline__ void mixed_rows(const uint8_t* __restrict__ pol0, const uint8_t* __restrict__ pol1,
                                           int N, float* __restrict__ fb, int64_t nrows, int nave, int nskip,
                                           int nchanout, int64_t ring_rows, int64_t row0, float scale,
                                           Passes passes) {
    extern __shared__ float lds[];
    const int words = N + N / 32;
    float* re = lds;
    float* im = lds + words;
    const int tid = threadIdx.x;
    const int64_t row = blockIdx.x;
    if (row >= nrows) return;
    constexpr int CPT = (NMAX / 2 + BLOCK - 1) / BLOCK;  // channels per thread (N/2 at most)
    float acc[CPT];
#pragma unroll
    for (int i = 0; i < CPT; i++) acc[i] = 0.0f;
    for (int frame = 0; frame < nave; frame++) {
        const size_t off = ((size_t)row * nave + frame) * N;
        // The passes' per-thread LDS addresses depend only on the thread
        // index; an opaque copy of it per frame keeps the compiler from
        // hoisting all of them out of the frame loop into (spilled) registers.
        int t = tid;
        asm volatile("" : "+v"(t));
        passes(pol0 + off, pol1 + off, re, im, t);
#pragma unroll
        for (int i = 0; i < CPT; i++) {
            const int c = tid + BLOCK * i;
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
        const int c = tid + BLOCK * i;
        if (c < nchanout) dst[c] = acc[i] * scale;
    }
}

// Any length the plan factorises: the radix sequence is read at run time.
// NMAX >= N bounds the butterflies per thread (register arrays are sized
// for it), so shorter lengths use a smaller bucket.
template <int NMAX>
__global__ void __launch_bounds__(GEN_BLOCK)
fb_generic_kernel(const uint8_t* __restrict__ pol0, const uint8_t* __restrict__ pol1,
                  const float2* __restrict__ tw, GenPlan plan, int N, float* __restrict__ fb,
                  int64_t nrows, int nave, int nskip, int nchanout, int64_t ring_rows, int64_t row0,
                  float scale) {
    mixed_rows<GEN_BLOCK, NMAX>(pol0, pol1, N, fb, nrows, nave, nskip, nchanout, ring_rows, row0, scale,
               [&](const uint8_t* x0, const uint8_t* x1, float* re, float* im, int t) {
                   for (int e = t; e < N; e += GEN_BLOCK) {
                       const int idx = lds_pad(e);
                       re[idx] = (float)x0[e] - 128.0f;
                       im[idx] = (float)x1[e] - 128.0f;
                   }
                   __syncthreads();
                   int Ns = 1;
                   for (int p = 0; p < plan.npass; p++) {
                       switch (plan.radix[p]) {
                           case 2: gen_pass<NMAX, 2>(re, im, N, Ns, tw, t); break;
                           case 3: gen_pass<NMAX, 3>(re, im, N, Ns, tw, t); break;
                           case 4: gen_pass<NMAX, 4>(re, im, N, Ns, tw, t); break;
                           case 5: gen_pass<NMAX, 5>(re, im, N, Ns, tw, t); break;
                           default: gen_pass<NMAX, 8>(re, im, N, Ns, tw, t); break;
                       }
                       Ns *= plan.radix[p];
                   }
               });
}

// A length with its radix sequence fixed at compile time (the order
// factorise() gives); same arguments as fb_generic_kernel, whose plan and N
// it ignores in favour of its own.
template <int N, int BLOCK, int... R>
__global__ void __launch_bounds__(BLOCK)
fb_fixed_kernel(const uint8_t* __restrict__ pol0, const uint8_t* __restrict__ pol1,
                const float2* __restrict__ tw, GenPlan, int, float* __restrict__ fb, int64_t nrows, int nave,
                int nskip, int nchanout, int64_t ring_rows, int64_t row0, float scale) {
    static_assert((R * ...) == N, "radices must multiply to N");
    mixed_rows<BLOCK, N>(pol0, pol1, N, fb, nrows, nave, nskip, nchanout, ring_rows, row0, scale,
                      [&](const uint8_t* x0, const uint8_t* x1, float* re, float* im, int t) {
                          fixed_rest<N, BLOCK, R...>(x0, x1, re, im, tw, t);
                      });
}

}  // namespace

struct vfrb_fb_plan {
    int N = 0, nave = 0, nskip = 0, nchanout = 0;
    bool pow2 = false;
    GenPlan gen{};
    float2* d_tw = nullptr;
    uint8_t* d_zero = nullptr;  // one frame pair of zeros: the framed kernel's missing slots
    size_t lds_bytes = 0;
    int block = 256;
    const void* kernel = nullptr;
    const void* kernel_framed = nullptr;  // power-of-two lengths only
};

namespace {

template <int N>
void select_pow2(vfrb_fb_plan* p) {
    p->kernel = reinterpret_cast<const void*>(&fb_pow2_kernel<N, false>);
    p->kernel_framed = reinterpret_cast<const void*>(&fb_pow2_kernel<N, true>);
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
        // 12500 points (NCHAN 6250, the reference programs' default) and
        // its halvings down to 1250 have kernels of their own (radices in
        // factorise()'s order); other lengths read their radices at run time
        const void* fixed = nullptr;
        switch (N) {
            case 12500: fixed = reinterpret_cast<const void*>(&fb_fixed_kernel<12500, FIXED_BLOCK, 5, 5, 5, 5, 5, 4>); break;
            case 6250: fixed = reinterpret_cast<const void*>(&fb_fixed_kernel<6250, FIXED_BLOCK, 5, 5, 5, 5, 5, 2>); break;
            case 2500: fixed = reinterpret_cast<const void*>(&fb_fixed_kernel<2500, FIXED_BLOCK, 5, 5, 5, 5, 4>); break;
            case 1250: fixed = reinterpret_cast<const void*>(&fb_fixed_kernel<1250, FIXED_BLOCK, 5, 5, 5, 5, 2>); break;
            default: break;
        }
        const void* generic = N <= 2048   ? reinterpret_cast<const void*>(&fb_generic_kernel<2048>)
                              : N <= 4096 ? reinterpret_cast<const void*>(&fb_generic_kernel<4096>)
                              : N <= 8192 ? reinterpret_cast<const void*>(&fb_generic_kernel<8192>)
                                          : reinterpret_cast<const void*>(&fb_generic_kernel<GEN_NMAX>);
        p->kernel = fixed ? fixed : generic;
        p->lds_bytes = 2 * sizeof(float) * (size_t)(N + N / 32);
        p->block = fixed ? FIXED_BLOCK : GEN_BLOCK;
    }
    // Twiddles exp(-2*pi*i*m/N), m < N, computed in FP64 and rounded once.
    std::vector<float2> tw(N);
    for (int m = 0; m < N; m++) {
        const double a = -2.0 * M_PI * (double)m / (double)N;
        tw[m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    hipError_t e = hipMalloc(&p->d_tw, sizeof(float2) * N);
    if (e == hipSuccess) e = hipMemcpy(p->d_tw, tw.data(), sizeof(float2) * N, hipMemcpyHostToDevice);
    if (e == hipSuccess && p->kernel_framed) e = hipMalloc(&p->d_zero, 2 * kFrameBytes);
    if (e == hipSuccess && p->kernel_framed) e = hipMemset(p->d_zero, 0, 2 * kFrameBytes);
    if (e == hipSuccess && p->lds_bytes > 65536)
        e = hipFuncSetAttribute(p->kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_bytes);
    if (e == hipSuccess && p->kernel_framed && p->lds_bytes > 65536)
        e = hipFuncSetAttribute(p->kernel_framed, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_bytes);
    if (e != hipSuccess) {
        if (p->d_tw) (void)hipFree(p->d_tw);
        if (p->d_zero) (void)hipFree(p->d_zero);
        delete p;
        return vfrb::hip_fail(e, "vfrb_fb_plan_create", __FILE__, __LINE__);
    }
    *plan = p;
    return VFRB_OK;
}

vfrb_status vfrb_fb_plan_destroy(vfrb_fb_plan* p) {
    if (!p) return VFRB_OK;
    if (p->d_tw) VFRB_HIP(hipFree(p->d_tw));
    if (p->d_zero) VFRB_HIP(hipFree(p->d_zero));
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
        VFRB_REQUIRE(((uintptr_t)pol0 % 4) == 0 && ((uintptr_t)pol1 % 4) == 0, VFRB_ERR_INVALID,
                     "pol buffers must be 4-byte aligned");
    }
    const float scale = (float)(0.5 / ((double)(p->N / 2 + 1) * (double)p->nave));
    hipStream_t s = vfrb::as_stream(stream);
    dim3 grid((unsigned)nrows), block(p->block);
    if (p->pow2) {
        const int32_t* no_table = nullptr;
        int64_t no_byte0 = 0;
        void* args[] = {(void*)&pol0, (void*)&pol1, (void*)&p->d_tw, (void*)&fb,  (void*)&nrows,
                        (void*)&p->nave, (void*)&p->nskip, (void*)&p->nchanout, (void*)&ring_rows,
                        (void*)&row0, (void*)&scale, (void*)&no_table, (void*)&no_byte0};
        VFRB_HIP(hipLaunchKernel(p->kernel, grid, block, args, p->lds_bytes, s));
    } else {
        const float2* tw = p->d_tw;
        void* args[] = {(void*)&pol0, (void*)&pol1, (void*)&tw, (void*)&p->gen, (void*)&p->N, (void*)&fb,
                        (void*)&nrows, (void*)&p->nave, (void*)&p->nskip, (void*)&p->nchanout,
                        (void*)&ring_rows, (void*)&row0, (void*)&scale};
        VFRB_HIP(hipLaunchKernel(p->kernel, grid, block, args, p->lds_bytes, s));
    }
    return VFRB_OK;
}

vfrb_status vfrb_fb_execute_frames(vfrb_fb_plan* p, const uint8_t* frames, const int32_t* table,
                                   int64_t nslots, int64_t byte0, int64_t nrows, float* fb,
                                   int64_t ring_rows, int64_t row0, void* stream) {
    VFRB_REQUIRE(p && frames && table && fb, VFRB_ERR_INVALID, "NULL argument");
    VFRB_REQUIRE(p->pow2, VFRB_ERR_UNSUPPORTED, "frame input needs a power-of-two FFT length");
    VFRB_REQUIRE(nrows >= 0 && row0 >= 0 && ring_rows >= 0 && byte0 >= 0 && nslots >= 0,
                 VFRB_ERR_INVALID, "negative size");
    VFRB_REQUIRE(byte0 % 4 == 0 && ((uintptr_t)frames % 4) == 0, VFRB_ERR_INVALID,
                 "frames and byte0 must be 4-byte aligned");
    VFRB_REQUIRE(ring_rows == 0 || nrows <= ring_rows, VFRB_ERR_INVALID, "nrows exceeds ring");
    VFRB_REQUIRE(nrows <= INT32_MAX, VFRB_ERR_INVALID, "nrows too large for one launch");
    VFRB_REQUIRE(byte0 + nrows * (int64_t)p->nave * p->N <= nslots * (int64_t)kPayloadBytes,
                 VFRB_ERR_INVALID, "rows reach past the %lld table slots", (long long)nslots);
    if (nrows == 0) return VFRB_OK;
    const float scale = (float)(0.5 / ((double)(p->N / 2 + 1) * (double)p->nave));
    const uint8_t* zero = p->d_zero;
    void* args[] = {(void*)&frames, (void*)&zero, (void*)&p->d_tw, (void*)&fb,  (void*)&nrows,
                    (void*)&p->nave, (void*)&p->nskip, (void*)&p->nchanout, (void*)&ring_rows,
                    (void*)&row0, (void*)&scale, (void*)&table, (void*)&byte0};
    VFRB_HIP(hipLaunchKernel(p->kernel_framed, dim3((unsigned)nrows), dim3(p->block), args, p->lds_bytes,
                             vfrb::as_stream(stream)));
    return VFRB_OK;
}

}  // extern "C"
