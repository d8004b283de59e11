// inputs.cpp -- input generators of the bfly_check program.  Each formula
// follows the reference's fft_check generator named beside it, so a row of
// a given (mode, n, batch, seed) holds the same floats in both programs.
#include "inputs.hpp"

#include <cmath>
#include <cstring>

namespace bfly_check {
namespace {

constexpr double kPi = 3.14159265358979323846;

// xorshift32 hash of (row, index), two rounds of (13, 17, 5)
// (src/host/fft_check.cpp:63-68)
uint32_t mix(uint32_t b, uint32_t k) {
    uint32_t x = b * 2654435761u + k * 40503u;
    for (int r = 0; r < 2; r++) {
        x ^= x << 13;
        x ^= x >> 17;
        x ^= x << 5;
    }
    return x;
}

// low 16 bits of a hash as a value in [-1, 1)
float unit16(uint32_t x) { return (float)((int32_t)(x & 0xffffu) - 32768) * (1.0f / 32768.0f); }

bool is(const char* mode, const char* a, const char* b = nullptr) {
    return std::strcmp(mode, a) == 0 || (b && std::strcmp(mode, b) == 0);
}

// numerical regression vectors (src/host/fft_check.cpp:83-116); unknown
// modes and "zero" give zeros
void numerical(float* h, uint32_t n, uint32_t batch, const char* mode, uint32_t seed) {
    static const float kScale[4] = {1.0f, 1e-2f, 1e-4f, 1e-6f};
    for (uint32_t b = 0; b < batch; b++) {
        const uint32_t tone = 1u + b % 7u;
        for (uint32_t k = 0; k < n; k++) {
            float re = 0.0f, im = 0.0f;
            if (is(mode, "impulse")) {
                re = k == b % n ? 1.0f : 0.0f;
            } else if (is(mode, "constant")) {
                re = 0.25f;
                im = -0.5f;
            } else if (is(mode, "tone", "single-tone")) {
                const double ph = 2.0 * kPi * (double)tone * (double)k / (double)n;
                re = (float)std::cos(ph);
                im = (float)std::sin(ph);
            } else if (is(mode, "alternating", "nyquist-alternating")) {
                re = (k & 1u) ? -1.0f : 1.0f;
            } else if (is(mode, "random", "random-seeded")) {
                re = unit16(mix(b ^ seed, k));
                im = unit16(mix(b ^ seed, k ^ 0x9e37u));
            } else if (is(mode, "dynamic", "high-dynamic-range")) {
                const float s = kScale[k & 3u];
                re = (k & 1u) ? -s : s;
                im = (k & 2u) ? 0.5f * s : -0.5f * s;
            } else if (is(mode, "near-cancellation")) {
                re = (k & 1u) ? -1.0f + 1e-6f : 1.0f;
                im = (k & 1u) ? 0.5f - 1e-6f : -0.5f;
            }
            h[2 * ((size_t)b * n + k)] = re;
            h[2 * ((size_t)b * n + k) + 1] = im;
        }
    }
}

// application-shaped rows (src/host/fft_check.cpp:118-159): 16-QAM OFDM
// symbols with guard bands and pilots, a four-target radar echo, and
// 8-point blocks of complex activations
void application(float* h, uint32_t n, uint32_t batch, const char* mode) {
    static const float kQam[4] = {-3.0f, -1.0f, 1.0f, 3.0f};
    static const float kFreq[4] = {1.0f / 16, 1.0f / 5, 1.0f / 3, 0.62f};
    static const float kAmp[4] = {1.0f, 0.5f, 0.25f, 0.125f};
    static const float kDop[4] = {0.0f, 1.0f / 64, -1.0f / 128, 3.0f / 256};
    static const float kCos[4] = {1.0f, 0.0f, -1.0f, 0.0f};
    static const float kSin[4] = {0.0f, 1.0f, 0.0f, -1.0f};
    const bool ofdm = is(mode, "ofdm"), radar = is(mode, "radar");
    for (uint32_t b = 0; b < batch; b++) {
        for (uint32_t k = 0; k < n; k++) {
            const uint32_t x = mix(b, k);
            float re = 0.0f, im = 0.0f;
            if (ofdm) {
                const bool guard = k == 0u || k < n / 16u || k >= n - n / 16u;
                if (!guard && k % 12u == 0u) {
                    re = (x & 1u) ? 1.0f : -1.0f;
                    im = (x & 2u) ? 1.0f : -1.0f;
                } else if (!guard) {
                    re = kQam[x & 3u] * 0.31622776f;
                    im = kQam[(x >> 2) & 3u] * 0.31622776f;
                }
            } else if (radar) {
                double sr = 0.0, si = 0.0;
                for (int t = 0; t < 4; t++) {
                    const double ph = 2.0 * kPi * (kFreq[t] * (double)k + kDop[t] * (double)b) +
                                      (t * 0.7 + ((x >> 8) & 255u) / 255.0 * 0.3);
                    sr += kAmp[t] * std::cos(ph);
                    si += kAmp[t] * std::sin(ph);
                }
                re = 0.25f * (float)sr;
                im = 0.25f * (float)si;
            } else {
                const uint32_t xb = mix(b, k / 8u);
                const float mag = (float)((xb >> 8) & 255u) / 255.0f;
                const uint32_t q = (xb >> 16) & 3u;
                re = mag * kCos[q];
                im = mag * kSin[q];
            }
            h[2 * ((size_t)b * n + k)] = re;
            h[2 * ((size_t)b * n + k) + 1] = im;
        }
    }
}

void sin_cos(float* h, size_t count) {
    for (size_t i = 0; i < count; i++) h[i] = (float)(std::sin(0.011 * i) + 0.25 * std::cos(0.037 * i));
}

}  // namespace

bool is_input_mode(const char* mode) {
    static const char* const kModes[] = {"sin",         "deterministic-sin-cos", "zero",
                                         "impulse",     "constant",              "tone",
                                         "single-tone", "alternating",           "nyquist-alternating",
                                         "random",      "random-seeded",         "dynamic",
                                         "high-dynamic-range", "near-cancellation", "ofdm",
                                         "radar",       "dl"};
    if (!mode) return false;
    for (const char* m : kModes)
        if (std::strcmp(mode, m) == 0) return true;
    return false;
}

void c2c_input(float* h, uint32_t n, uint32_t batch, const char* mode, uint32_t seed) {
    if (is(mode, "sin", "deterministic-sin-cos")) sin_cos(h, 2 * (size_t)n * batch);
    else if (is(mode, "ofdm") || is(mode, "radar", "dl")) application(h, n, batch, mode);
    else numerical(h, n, batch, mode, seed);
}

void r2c_input(float* h, uint32_t n, uint32_t batch, const char* mode, uint32_t seed) {
    const size_t count = (size_t)n * batch;
    if (is(mode, "sin", "deterministic-sin-cos")) {
        sin_cos(h, count);
        return;
    }
    float* cx = new float[2 * count];
    c2c_input(cx, n, batch, mode, seed);
    for (size_t i = 0; i < count; i++) h[i] = cx[2 * i];
    delete[] cx;
}

void half_input(float* h, uint32_t n, uint32_t batch, size_t stride) {
    const uint32_t m = n / 2;
    for (uint32_t b = 0; b < batch; b++) {
        float* row = h + (size_t)b * stride;
        for (uint32_t k = 0; k <= m; k++) {
            row[2 * k] = unit16(mix(b, k));
            row[2 * k + 1] = k == m ? 0.0f : unit16(mix(b, k ^ 0x5bd1u));
        }
    }
}

}  // namespace bfly_check
