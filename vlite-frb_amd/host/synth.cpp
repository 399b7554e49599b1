// Synthetic VLITE voltages: 8-bit offset-binary samples of Gaussian noise
// N(mean, sigma^2) with dispersed pulses and narrowband tones.
//
// Every sample depends only on (seed, pol, absolute sample index), so any
// block range can be generated independently and in parallel.  Pulses are
// synthesised in the frequency domain per nfft-sample block: FFT bin k of
// block b carries extra complex-Gaussian power A*g(t_b - t0 - d(f_k)) on top
// of the white noise, with g a Gaussian of the given FWHM and d the cold
// plasma delay 4.148808e3*DM*(f^-2 - f_top^-2) s (f in MHz).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <thread>
#include <vector>

#include "../csrc/common.hpp"

namespace {

inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

inline uint64_t stream_key(uint64_t seed, uint64_t pol, uint64_t purpose) {
    return mix64(mix64(seed * 0x2545f4914f6cdd1dull + 0x1234567ull) ^ mix64(pol * 31 + purpose * 1009 + 7));
}

// Two standard normals from one 64-bit counter (Box-Muller on 32-bit halves).
inline void normal_pair(uint64_t key, uint64_t counter, float* z0, float* z1) {
    const uint64_t r = mix64(key ^ (counter * 0xd1342543de82ef95ull));
    const float u1 = ((float)(uint32_t)(r >> 32) + 0.5f) * (1.0f / 4294967296.0f);
    const float u2 = ((float)(uint32_t)r + 0.5f) * (1.0f / 4294967296.0f);
    const float rad = std::sqrt(-2.0f * std::log(u1));
    float s, c;
    sincosf(6.283185307179586f * u2, &s, &c);
    *z0 = rad * c;
    *z1 = rad * s;
}

// In-place iterative radix-2 complex FFT; tw[k] = exp(+2*pi*i*k/n), k < n/2
// (the inverse kernel, unnormalised).
void ifft_pow2(std::vector<std::complex<double>>& a, const std::vector<std::complex<double>>& tw) {
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t step = n / len;
        for (size_t i = 0; i < n; i += len) {
            for (size_t k = 0; k < len / 2; k++) {
                const std::complex<double> w = tw[k * step];
                const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
        }
    }
}

template <class F>
void parallel_for(int64_t n, int nthreads, F&& body) {
    if (nthreads <= 0) nthreads = (int)std::thread::hardware_concurrency();
    if (nthreads > 64) nthreads = 64;
    if (nthreads <= 1 || n < 2) {
        for (int64_t i = 0; i < n; i++) body(i);
        return;
    }
    std::vector<std::thread> pool;
    const int64_t chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; t++) {
        const int64_t lo = t * chunk, hi = std::min<int64_t>(n, lo + chunk);
        if (lo >= hi) break;
        pool.emplace_back([lo, hi, &body] {
            for (int64_t i = lo; i < hi; i++) body(i);
        });
    }
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" {

double vfrb_synth_amp_for_snr(double snr, int32_t nave, int32_t nchan) {
    if (nave <= 0 || nchan <= 0) return 0.0;
    return snr / std::sqrt(2.0 * nave * nchan);
}

vfrb_status vfrb_synth_payload(const vfrb_synth_spec* spec, int32_t pol, int64_t block0,
                               int64_t nblocks, uint8_t* out) {
    VFRB_REQUIRE(spec && out, VFRB_ERR_INVALID, "NULL argument");
    VFRB_REQUIRE(pol == 0 || pol == 1, VFRB_ERR_INVALID, "pol must be 0 or 1");
    const int64_t n = spec->nfft;
    VFRB_REQUIRE(n >= 16 && n % 2 == 0 && block0 >= 0 && nblocks >= 0, VFRB_ERR_INVALID,
                 "bad nfft/block range");
    VFRB_REQUIRE(spec->npulse >= 0 && spec->npulse <= VFRB_SYNTH_MAX_PULSES && spec->ntone >= 0 &&
                     spec->ntone <= VFRB_SYNTH_MAX_TONES,
                 VFRB_ERR_INVALID, "too many pulses/tones");
    VFRB_REQUIRE(spec->npulse == 0 || (n & (n - 1)) == 0, VFRB_ERR_UNSUPPORTED,
                 "pulse injection needs a power-of-two nfft");
    const uint64_t knoise = stream_key(spec->seed, pol, 1);
    const uint64_t kpulse = stream_key(spec->seed, pol, 2);
    const double fs = VFRB_SAMPLERATE;
    const double bin_mhz = spec->bw_mhz / (double)(n / 2);
    const double ftop2 = 1.0 / (spec->f_top_mhz * spec->f_top_mhz);
    std::vector<double> tone_phase(spec->ntone);
    for (int i = 0; i < spec->ntone; i++) {
        const uint64_t r = mix64(stream_key(spec->seed, pol, 3) + (uint64_t)i);
        tone_phase[i] = 2.0 * M_PI * (double)(r >> 11) * (1.0 / 9007199254740992.0);
    }
    const float mean = (float)spec->mean, sig = (float)spec->sigma;
    // Tones on integer bins repeat every block: their sum is tabulated once.
    bool periodic = true;
    for (int t = 0; t < spec->ntone; t++)
        periodic = periodic && spec->tone_bin[t] == std::floor(spec->tone_bin[t]);
    std::vector<float> tone_table;
    if (spec->ntone > 0 && periodic) {
        tone_table.assign(n, 0.0f);
        std::vector<double> acc(n, 0.0);
        for (int t = 0; t < spec->ntone; t++)
            for (int64_t i = 0; i < n; i++) {
                const double ph = 2.0 * M_PI * std::fmod(spec->tone_bin[t] * (double)i / (double)n, 1.0) +
                                  tone_phase[t];
                acc[i] += spec->tone_amp[t] * std::cos(ph);
            }
        for (int64_t i = 0; i < n; i++) tone_table[i] = (float)acc[i];
    }
    std::vector<std::complex<double>> tw;
    if (spec->npulse > 0) {
        tw.resize(n / 2);
        for (int64_t k = 0; k < n / 2; k++)
            tw[k] = std::complex<double>(std::cos(2.0 * M_PI * k / n), std::sin(2.0 * M_PI * k / n));
    }

    parallel_for(nblocks, spec->nthreads, [&](int64_t ib) {
        const int64_t b = block0 + ib;
        const int64_t s0 = b * n;
        std::vector<float> x(n);
        for (int64_t i = 0; i < n; i += 2) {
            float z0, z1;
            normal_pair(knoise, (uint64_t)(s0 + i) >> 1, &z0, &z1);
            x[i] = mean + sig * z0;
            x[i + 1] = mean + sig * z1;
        }
        if (!tone_table.empty())
            for (int64_t i = 0; i < n; i++) x[i] += tone_table[i];
        // Other tones: a unit phasor rotated by exp(2*pi*i*f) per sample,
        // re-anchored from the exact phase every 256 samples (drift ~1e-14).
        for (int t = 0; t < spec->ntone && tone_table.empty(); t++) {
            const double f = spec->tone_bin[t] / (double)n;  // cycles per sample
            const double amp = spec->tone_amp[t];
            const std::complex<double> step(std::cos(2.0 * M_PI * f), std::sin(2.0 * M_PI * f));
            for (int64_t i0 = 0; i0 < n; i0 += 256) {
                const double ph = 2.0 * M_PI * std::fmod(f * (double)(s0 + i0), 1.0) + tone_phase[t];
                std::complex<double> z(std::cos(ph), std::sin(ph));
                const int64_t i1 = std::min<int64_t>(n, i0 + 256);
                for (int64_t i = i0; i < i1; i++) {
                    x[i] += (float)(amp * z.real());
                    z *= step;
                }
            }
        }
        if (spec->npulse > 0) {
            const double tb = ((double)b + 0.5) * (double)n / fs;
            std::vector<std::complex<double>> spec_c;
            bool any = false;
            // Bins whose delayed pulse centre lies within 4 FWHM of this
            // block: the delay is monotone in k, so invert it at both ends.
            int64_t klo = n, khi = -1;
            for (int p = 0; p < spec->npulse; p++) {
                const double w4 = 4.0 * spec->pulse_fwhm[p];
                const double dlo = tb - spec->pulse_t0[p] - w4, dhi = tb - spec->pulse_t0[p] + w4;
                if (dhi < 0.0) continue;
                auto k_of = [&](double d) {
                    if (spec->pulse_dm[p] <= 0.0) return d <= 0.0 ? 0.0 : (double)(n / 2);
                    const double f = 1.0 / std::sqrt(std::max(d, 0.0) / (4.148808e3 * spec->pulse_dm[p]) + ftop2);
                    return (spec->f_top_mhz - f) / bin_mhz;
                };
                const int64_t a = (int64_t)std::floor(k_of(dlo)) - 1;
                const int64_t c = (int64_t)std::ceil(k_of(dhi)) + 1;
                klo = std::min(klo, std::max<int64_t>(a, 0));
                khi = std::max(khi, std::min<int64_t>(c, n / 2));
            }
            for (int64_t k = klo; k <= khi; k++) {
                const double fk = spec->f_top_mhz - (double)k * bin_mhz;
                double excess = 0.0;
                for (int p = 0; p < spec->npulse; p++) {
                    const double d = 4.148808e3 * spec->pulse_dm[p] * (1.0 / (fk * fk) - ftop2);
                    const double u = (tb - spec->pulse_t0[p] - d) / spec->pulse_fwhm[p];
                    if (std::fabs(u) > 4.0) continue;
                    excess += spec->pulse_amp[p] * std::exp(-4.0 * M_LN2 * u * u);
                }
                if (excess <= 1e-7) continue;
                if (!any) {
                    spec_c.assign(n, std::complex<double>(0.0, 0.0));
                    any = true;
                }
                float z0, z1;
                normal_pair(kpulse, (uint64_t)(b * (n / 2 + 1) + k), &z0, &z1);
                const double var = (double)sig * sig * (double)n * excess;
                if (k == 0 || k == n / 2) {
                    spec_c[k] = std::complex<double>(std::sqrt(var) * z0, 0.0);
                } else {
                    const double a = std::sqrt(var * 0.5);
                    spec_c[k] = std::complex<double>(a * z0, a * z1);
                    spec_c[n - k] = std::conj(spec_c[k]);
                }
            }
            if (any) {
                ifft_pow2(spec_c, tw);
                const double inv = 1.0 / (double)n;
                for (int64_t i = 0; i < n; i++) x[i] += (float)(spec_c[i].real() * inv);
            }
        }
        uint8_t* o = out + ib * n;
        for (int64_t i = 0; i < n; i++) {
            float v = std::floor(x[i] + 0.5f);
            v = v < 0.0f ? 0.0f : (v > 255.0f ? 255.0f : v);
            o[i] = (uint8_t)v;
        }
    });
    return VFRB_OK;
}

}  // extern "C"
