/*
 * bfly_oracle.c -- CPU restatement of the reference's FFT path, used ONLY as a
 * checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * Nothing in the shipped library links, loads or calls this file.
 *
 * Every function names the reference file:line it restates (paths relative to
 * the reference repository root).  Pinning (what checks each function against
 * something other than its own formula) lives in tests/test_oracle.py:
 *   - the reference's own code compiled from its sources into oracle/_ref
 *     (reference.cpp, fft_k.hpp, long_fft_ub.h, descriptors.hpp, butterfly.cpp
 *     Generator tables),
 *   - golden values the reference's tests hold (tests/test_limits.cpp,
 *     tests/test_runtime_guards.cpp, tests/test_descriptors.py),
 *   - mathematics: an O(N^2) direct DFT, numpy/scipy pocketfft in FP64,
 *     closed forms (impulse, constant, tone), Parseval and linearity.
 *
 * Plain C99; build with oracle/Makefile.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------ helpers */
static uint32_t ilog2u(uint32_t n) {
    uint32_t l = 0;
    while ((1u << l) < n) l++;
    return l;
}

static uint32_t bitrev(uint32_t v, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t b = 0; b < bits; b++) { r = (r << 1) | (v & 1u); v >>= 1; }
    return r;
}

/* ---------------------------------------------------------- FP64 transforms */
/* FP64 radix-2 DIT FFT in place on interleaved doubles, sign = -1 forward,
 * +1 for the unnormalised inverse.  Algorithm of
 * src/framework/reference.cpp:14-50 (refFftF32): bit-reversal permutation,
 * twiddle table w[i] = exp(sign*2*pi*i*i/n) evaluated in FP64, then the
 * len = 2,4,..,n butterfly sweep with w[j*n/len]. */
ORC_API void orc_fft_f64_inplace(double* a, uint32_t n, int sign) {
    if (n < 2) return;
    const uint32_t logn = ilog2u(n);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t r = bitrev(i, logn);
        if (r > i) {
            double t0 = a[2 * i], t1 = a[2 * i + 1];
            a[2 * i] = a[2 * r]; a[2 * i + 1] = a[2 * r + 1];
            a[2 * r] = t0; a[2 * r + 1] = t1;
        }
    }
    double* w = (double*)malloc(sizeof(double) * n);  /* n/2 complex */
    for (uint32_t i = 0; i < n / 2; i++) {
        const double ang = (double)sign * 2.0 * M_PI * (double)i / (double)n;
        w[2 * i] = cos(ang);
        w[2 * i + 1] = sin(ang);
    }
    for (uint32_t len = 2; len <= n; len <<= 1) {
        const uint32_t half = len >> 1, step = n / len;
        for (uint32_t base = 0; base < n; base += len) {
            for (uint32_t j = 0; j < half; j++) {
                const double wr = w[2 * (j * step)], wi = w[2 * (j * step) + 1];
                double* u = a + 2 * (base + j);
                double* v = a + 2 * (base + j + half);
                const double vr = v[0] * wr - v[1] * wi;
                const double vi = v[0] * wi + v[1] * wr;
                v[0] = u[0] - vr; v[1] = u[1] - vi;
                u[0] += vr; u[1] += vi;
            }
        }
    }
    free(w);
}

/* Batched FP64 C2C of fp32 interleaved rows [batch][2n] -> FP64 [batch][2n]. */
ORC_API void orc_c2c_f64(const float* in, double* out, uint32_t n, uint32_t batch, int sign) {
    for (uint32_t b = 0; b < batch; b++) {
        double* o = out + (size_t)b * 2 * n;
        const float* x = in + (size_t)b * 2 * n;
        for (size_t i = 0; i < 2 * (size_t)n; i++) o[i] = (double)x[i];
        orc_fft_f64_inplace(o, n, sign);
    }
}

/* 2-D C2C of `batch` row-major [n1][n2] interleaved fp32 arrays in FP64:
 * the 1-D FP64 FFT along every row, then along every column.  No reference
 * counterpart (2-D transforms are future work there,
 * config/test_matrix.json:94-99); the definition is the separable 2-D DFT. */
ORC_API void orc_fft2_f64(const float* in, double* out, uint32_t n1, uint32_t n2, uint32_t batch, int sign) {
    orc_c2c_f64(in, out, n2, batch * n1, sign);
    double* col = (double*)malloc(sizeof(double) * 2 * n1);
    for (uint32_t b = 0; b < batch; b++) {
        double* m = out + (size_t)b * 2 * n1 * n2;
        for (uint32_t c = 0; c < n2; c++) {
            for (uint32_t r = 0; r < n1; r++) {
                col[2 * r] = m[2 * ((size_t)r * n2 + c)];
                col[2 * r + 1] = m[2 * ((size_t)r * n2 + c) + 1];
            }
            orc_fft_f64_inplace(col, n1, sign);
            for (uint32_t r = 0; r < n1; r++) {
                m[2 * ((size_t)r * n2 + c)] = col[2 * r];
                m[2 * ((size_t)r * n2 + c) + 1] = col[2 * r + 1];
            }
        }
    }
    free(col);
}

/* refFftF32 (src/framework/reference.cpp:14-50): FP64 compute, fp32 result. */
ORC_API void orc_ref_fft_f32(const float* in, float* out, uint32_t n) {
    if (!n) return;
    double* a = (double*)malloc(sizeof(double) * 2 * n);
    for (size_t i = 0; i < 2 * (size_t)n; i++) a[i] = (double)in[i];
    orc_fft_f64_inplace(a, n, -1);
    for (size_t i = 0; i < 2 * (size_t)n; i++) out[i] = (float)a[i];
    free(a);
}

/* O(N^2) direct DFT in FP64 (definition, docs/user-guide/transforms.md:15-22);
 * long-double accumulation and exact-integer angle reduction (j*k mod n) so
 * it is an independent check of the fast oracle at small n. */
ORC_API void orc_dft_direct_f64(const float* in, double* out, uint32_t n, int sign) {
    for (uint32_t k = 0; k < n; k++) {
        long double sr = 0, si = 0;
        for (uint32_t j = 0; j < n; j++) {
            const uint64_t m = ((uint64_t)j * k) % n;
            const long double ang = (long double)sign * 2.0L * (long double)M_PI *
                                    (long double)m / (long double)n;
            const long double c = cosl(ang), s = sinl(ang);
            const long double xr = in[2 * j], xi = in[2 * j + 1];
            sr += xr * c - xi * s;
            si += xr * s + xi * c;
        }
        out[2 * k] = (double)sr;
        out[2 * k + 1] = (double)si;
    }
}

/* refR2CF32 (src/framework/reference.cpp:55-62): n reals -> n/2+1 interleaved
 * bins (numpy.fft.rfft layout, row of n+2 floats), FP64 compute. */
ORC_API void orc_ref_r2c_f32(const float* in, float* out, uint32_t n) {
    if (n < 2) return;
    double* a = (double*)malloc(sizeof(double) * 2 * n);
    for (uint32_t j = 0; j < n; j++) { a[2 * j] = in[j]; a[2 * j + 1] = 0.0; }
    orc_fft_f64_inplace(a, n, -1);
    for (uint32_t j = 0; j <= n / 2; j++) {
        out[2 * j] = (float)a[2 * j];
        out[2 * j + 1] = (float)a[2 * j + 1];
    }
    free(a);
}

/* FP64-output R2C for rel-L2 gating: out[n+2] doubles. */
ORC_API void orc_r2c_f64(const float* in, double* out, uint32_t n, uint32_t batch,
                         uint32_t in_stride, uint32_t out_stride) {
    double* a = (double*)malloc(sizeof(double) * 2 * n);
    for (uint32_t b = 0; b < batch; b++) {
        const float* x = in + (size_t)b * in_stride;
        for (uint32_t j = 0; j < n; j++) { a[2 * j] = x[j]; a[2 * j + 1] = 0.0; }
        orc_fft_f64_inplace(a, n, -1);
        memcpy(out + (size_t)b * out_stride, a, sizeof(double) * (n + 2));
    }
    free(a);
}

/* refC2RF32 (src/framework/reference.cpp:64-79): bins k=0..n/2 are taken from
 * the input, X[n-k] = conj(X[k]) for 0<k<n/2, then out = Re(IDFT)/n computed
 * as Re(conj(F(conj X)))/n.  Only the real parts of the DC and Nyquist bins
 * reach the output (their imaginary parts land in Im, which is dropped). */
static void c2r_expand(const float* in, double* X, uint32_t n) {
    const uint32_t m = n >> 1;
    for (uint32_t k = 0; k <= m; k++) { X[2 * k] = in[2 * k]; X[2 * k + 1] = in[2 * k + 1]; }
    for (uint32_t k = 1; k < m; k++) { X[2 * (n - k)] = in[2 * k]; X[2 * (n - k) + 1] = -in[2 * k + 1]; }
}

ORC_API void orc_ref_c2r_f32(const float* in, float* out, uint32_t n) {
    if (n < 2) return;
    double* X = (double*)malloc(sizeof(double) * 2 * n);
    c2r_expand(in, X, n);
    /* reference.cpp:74-78 routes the conjugated spectrum through fp32 (cin)
     * before the FP64 FFT; the conjugation and the float cast are exact. */
    for (uint32_t j = 0; j < n; j++) X[2 * j + 1] = -(double)(float)X[2 * j + 1];
    orc_fft_f64_inplace(X, n, -1);
    for (uint32_t j = 0; j < n; j++) out[j] = (float)X[2 * j] / (float)n;
    free(X);
}

ORC_API void orc_c2r_f64(const float* in, double* out, uint32_t n, uint32_t batch,
                         uint32_t in_stride, uint32_t out_stride) {
    double* X = (double*)malloc(sizeof(double) * 2 * n);
    for (uint32_t b = 0; b < batch; b++) {
        c2r_expand(in + (size_t)b * in_stride, X, n);
        orc_fft_f64_inplace(X, n, +1);
        double* o = out + (size_t)b * out_stride;
        for (uint32_t j = 0; j < n; j++) o[j] = X[2 * j] / (double)n;
    }
    free(X);
}

/* ------------------------------------------------------------- error gates */
/* maxRelScaled (include/butterfly/reference.hpp:30-41): max |got-ref| over
 * float components divided by max(1, max|ref|); non-finite -> +inf. */
ORC_API double orc_max_rel_scaled(const float* got, const float* ref, size_t len) {
    if (len && (!got || !ref)) return INFINITY;
    double scale = 0, maxErr = 0;
    for (size_t i = 0; i < len; i++) {
        if (!isfinite(got[i]) || !isfinite(ref[i])) return INFINITY;
        const double r = fabs((double)ref[i]);
        const double e = fabs((double)got[i] - (double)ref[i]);
        if (r > scale) scale = r;
        if (e > maxErr) maxErr = e;
    }
    return maxErr / (scale > 1.0 ? scale : 1.0);
}

/* Relative L2 error ||got - ref|| / ||ref|| of one transform against FP64
 * (the 1e-5 per-transform gate).  ||ref|| == 0 falls back to ||got||. */
ORC_API double orc_rel_l2(const float* got, const double* ref, size_t len) {
    long double num = 0, den = 0;
    for (size_t i = 0; i < len; i++) {
        if (!isfinite(got[i])) return INFINITY;
        const long double d = (long double)got[i] - (long double)ref[i];
        num += d * d;
        den += (long double)ref[i] * (long double)ref[i];
    }
    if (den == 0) return (double)sqrtl(num);
    return (double)sqrtl(num / den);
}

/* Worst per-row rel-L2 over a batch of rows (row length len, strides in
 * elements). */
ORC_API double orc_rel_l2_rows(const float* got, const double* ref, uint32_t rows, size_t len,
                               size_t got_stride, size_t ref_stride) {
    double worst = 0;
    for (uint32_t b = 0; b < rows; b++) {
        const double e = orc_rel_l2(got + (size_t)b * got_stride, ref + (size_t)b * ref_stride, len);
        if (!(e <= worst)) worst = e;  /* NaN-propagating max */
    }
    return worst;
}

/* ----------------------------------------------------------------- inputs */
/* patternAt (src/framework/reference.cpp:10-12). */
ORC_API float orc_pattern_at(size_t i) {
    return (float)((i * 2654435761u) % 1000) / 1000.f - 0.5f;
}

ORC_API void orc_fill_pattern(float* dst, size_t count) {
    for (size_t i = 0; i < count; i++) dst[i] = orc_pattern_at(i);
}

/* sampleBatches (include/butterfly/reference.hpp:48-60). */
ORC_API uint32_t orc_sample_batches(uint32_t batch, uint32_t want, uint32_t* out, uint32_t maxOut) {
    if (want > maxOut) want = maxOut;
    if (batch == 0 || want == 0) return 0;
    if (batch <= want) {
        for (uint32_t i = 0; i < batch; i++) out[i] = i;
        return batch;
    }
    if (want == 1) { out[0] = 0; return 1; }
    for (uint32_t i = 0; i < want; i++)
        out[i] = (uint32_t)((uint64_t)i * (uint64_t)(batch - 1) / (uint64_t)(want - 1));
    return want;
}

/* appHash (src/host/fft_check.cpp:63-68): xorshift32, two rounds of three. */
ORC_API uint32_t orc_app_hash(uint32_t b, uint32_t k) {
    uint32_t x = b * 2654435761u + k * 40503u;
    for (int r = 0; r < 2; r++) { x ^= x << 13; x ^= x >> 17; x ^= x << 5; }
    return x;
}

static float hash_unit(uint32_t x) {
    return (float)((int32_t)(x & 0xffffu) - 32768) * (1.f / 32768.f);
}

/* Numerical-regression inputs (src/host/fft_check.cpp:83-116), interleaved
 * [batch][2n].  mode: 0 zero, 1 impulse, 2 constant, 3 tone, 4 alternating,
 * 5 random-seeded, 6 high-dynamic-range, 7 near-cancellation. */
ORC_API void orc_gen_numerical(float* h, uint32_t n, uint32_t batch, int mode, uint32_t seed) {
    static const float scale[4] = {1.f, 1e-2f, 1e-4f, 1e-6f};
    for (uint32_t b = 0; b < batch; b++) {
        const uint32_t toneBin = 1u + b % 7u;
        for (uint32_t k = 0; k < n; k++) {
            float re = 0.f, im = 0.f;
            switch (mode) {
                case 1: re = (k == b % n) ? 1.f : 0.f; break;
                case 2: re = 0.25f; im = -0.5f; break;
                case 3: {
                    const double ph = 2.0 * M_PI * (double)toneBin * (double)k / (double)n;
                    re = (float)cos(ph); im = (float)sin(ph);
                } break;
                case 4: re = (k & 1u) ? -1.f : 1.f; break;
                case 5:
                    re = hash_unit(orc_app_hash(b ^ seed, k));
                    im = hash_unit(orc_app_hash(b ^ seed, k ^ 0x9e37u));
                    break;
                case 6: {
                    const float s = scale[k & 3u];
                    re = (k & 1u) ? -s : s;
                    im = (k & 2u) ? 0.5f * s : -0.5f * s;
                } break;
                case 7:
                    re = (k & 1u) ? -1.f + 1e-6f : 1.f;
                    im = (k & 1u) ? 0.5f - 1e-6f : -0.5f;
                    break;
                default: break;
            }
            h[2u * ((size_t)b * n + k)] = re;
            h[2u * ((size_t)b * n + k) + 1] = im;
        }
    }
}

/* Application-shaped inputs (src/host/fft_check.cpp:118-159):
 * app 0 ofdm, 1 radar, 2 dl. */
ORC_API void orc_gen_app(float* h, uint32_t n, uint32_t batch, int app) {
    static const float q[4] = {-3.f, -1.f, 1.f, 3.f};
    static const float fr[4] = {1.f / 16, 1.f / 5, 1.f / 3, 0.62f};
    static const float am[4] = {1.f, 0.5f, 0.25f, 0.125f};
    static const float dop[4] = {0.f, 1.f / 64, -1.f / 128, 3.f / 256};
    static const float c4[4] = {1.f, 0.f, -1.f, 0.f};
    static const float s4[4] = {0.f, 1.f, 0.f, -1.f};
    for (uint32_t b = 0; b < batch; b++) {
        for (uint32_t k = 0; k < n; k++) {
            const uint32_t x = orc_app_hash(b, k);
            float re = 0.f, im = 0.f;
            if (app == 0) {
                const int guard = (k == 0u) || (k < n / 16u) || (k >= n - n / 16u);
                if (guard) { re = 0.f; im = 0.f; }
                else if ((k % 12u) == 0u) { re = (x & 1u) ? 1.f : -1.f; im = (x & 2u) ? 1.f : -1.f; }
                else { re = q[(x >> 0) & 3u] * 0.31622776f; im = q[(x >> 2) & 3u] * 0.31622776f; }
            } else if (app == 1) {
                double sr = 0, si = 0;
                for (int t = 0; t < 4; t++) {
                    const double ph = 2.0 * M_PI * (fr[t] * (double)k + dop[t] * (double)b) +
                                      (t * 0.7 + ((x >> 8) & 255u) / 255.0 * 0.3);
                    sr += am[t] * cos(ph);
                    si += am[t] * sin(ph);
                }
                re = 0.25f * (float)sr; im = 0.25f * (float)si;
            } else {
                const uint32_t xb = orc_app_hash(b, k / 8u);
                const float mag = (float)((xb >> 8) & 255u) / 255.0f;
                const uint32_t phs = (xb >> 16) & 3u;
                re = mag * c4[phs]; im = mag * s4[phs];
            }
            h[2u * ((size_t)b * n + k)] = re;
            h[2u * ((size_t)b * n + k) + 1] = im;
        }
    }
}

/* Default "sin" input (src/host/fft_check.cpp:167-171). */
ORC_API void orc_gen_sin(float* h, size_t elements) {
    for (size_t i = 0; i < elements; i++)
        h[i] = (float)(sin(0.011 * i) + 0.25 * cos(0.037 * i));
}

/* C2R half-spectrum input (src/host/fft_check.cpp:179-192), written at row
 * stride `stride` floats (the reference uses n+16 internally; the public
 * layout is n+2). */
ORC_API void orc_gen_half(float* h, uint32_t n, uint32_t batch, uint32_t stride) {
    const uint32_t m = n >> 1;
    for (uint32_t b = 0; b < batch; b++) {
        for (uint32_t k = 0; k <= m; k++) {
            const uint32_t x1 = orc_app_hash(b, k);
            const uint32_t x2 = orc_app_hash(b, k ^ 0x5bd1u);
            h[(size_t)b * stride + 2u * k] = hash_unit(x1);
            h[(size_t)b * stride + 2u * k + 1u] = (k == m) ? 0.f : hash_unit(x2);
        }
    }
}

/* ------------------------------------------------------- integer rules */
/* planeKFor (include/butterfly/fft_k.hpp:45-85): K-plane factor in {8,16,32}
 * minimising the Ascend vector-op count; ties go to the larger K. */
ORC_API uint32_t orc_plane_k_for(uint32_t n) {
    const uint32_t kNeed = (n + 63u) >> 6;
    const uint32_t kMin = (n <= 1024u && kNeed > 8u) ? kNeed : 8u;
    uint32_t best = kMin;
    uint64_t bestKey = ~0ull;
    for (uint32_t K = kMin; K <= 32; K <<= 1) {
        if (K * 8u > n) break;
        const uint32_t logK = ilog2u(K), logn = ilog2u(n);
        uint64_t ops = 0;
        uint32_t si = 0;
        for (; si + 1u < logK; si += 2u) {
            const uint32_t h = 1u << si;
            const uint64_t blocks = (uint64_t)K / (4u * h);
            ops += blocks * (16ull + (uint64_t)(h - 1u) * 28ull);
        }
        if (si < logK) {
            const uint32_t h = 1u << si;
            const uint64_t pairs = (uint64_t)K / 2u;
            const uint64_t triv = (uint64_t)K / (2u * h);
            ops += triv * 6ull + (pairs - triv) * 8ull;
        }
        for (uint32_t s = logK; s < logn; s++) {
            const uint32_t h = 1u << s;
            const uint32_t groups = n / (2u * h);
            const uint32_t nSlice = (h + 63u) >> 6;
            const uint32_t merge = (h >= 8u) && (((2u * h) % 8u) == 0u) &&
                                   (((2u * h) / 8u) <= 255u) && ((h / 8u) <= 255u) &&
                                   (groups <= 255u) && (nSlice <= groups);
            ops += merge ? (uint64_t)8u * nSlice : (uint64_t)8u * groups;
            ops += 1ull;
        }
        ops += 7ull;
        if (ops <= bestKey) { bestKey = ops; best = K; }
    }
    return best;
}

/* foldDFor (include/butterfly/fft_k.hpp:116-144): batch fold D in 1..4. */
ORC_API uint32_t orc_fold_d_for(uint32_t n, uint32_t batch, uint32_t nblk) {
    uint32_t D = 4;
    const uint32_t K = orc_plane_k_for(n);
    if (K == 0u || (n % K) != 0u) return 1u;
    if (n / K == 0u) return 1u;
    if ((n >> 3) > 255u) return 1u;
    const uint32_t groupsMax = n >> (ilog2u(K) + 1u);
    if (groupsMax == 0u) return 1u;
    while (D > 1u && D * groupsMax > 255u) D--;
    if (nblk > 1u)
        while (D > 1u && ((batch - 1u) / D + 1u) < nblk) D--;
    return D;
}

/* AB_ROW_FFT_UB_BYTES (include/butterfly/long_fft_ub.h:106-118). */
ORC_API uint64_t orc_row_fft_ub_bytes(uint64_t n, uint64_t D, uint64_t K) {
    const uint64_t half = n >> 1, nDK = n * D / K;
    return 24ull * n * D + 8ull * n + 8ull * (n + 16ull) +
           4ull * (n * D + (half > nDK ? half : nDK) + 3ull * n * D / K);
}

/* resolve_row_fft (include/butterfly/descriptors.hpp:231-251); returns 1 when
 * legal, 0 when an override violates K in {8,16,32}, K | len, len/K >= 8. */
ORC_API int orc_resolve_row_fft(uint32_t len, uint32_t rows, uint32_t dOv, uint32_t kOv,
                                uint32_t* foldD, uint32_t* planeK) {
    uint32_t D = dOv ? dOv : orc_fold_d_for(len, rows, 48u);
    if (D < 1u) D = 1u;
    const uint32_t K = kOv ? kOv : orc_plane_k_for(len);
    *foldD = D;
    *planeK = K;
    const int kSet = (K == 8u || K == 16u || K == 32u);
    if (!kSet || K * 8u > len || (len % K) != 0u) return 0;
    return 1;
}

/* -------------------------------------------- Generator tables (byte offsets) */
/* Generator::genTwiddles (src/framework/butterfly.cpp:96-109): per-stage slot
 * for h = 1..n/2 of align8(h) entries, entry j = exp(-i*pi*j/h); n+16 total. */
ORC_API void orc_gen_twiddles(uint32_t n, float* twr, float* twi) {
    const uint32_t pad = n + 16u;
    for (uint32_t i = 0; i < pad; i++) { twr[i] = 0.f; twi[i] = 0.f; }
    uint32_t off = 0;
    for (uint32_t h = 1; h <= n / 2; h <<= 1) {
        for (uint32_t j = 0; j < h; j++) {
            const double ang = -M_PI * (double)j / (double)h;
            twr[off + j] = (float)cos(ang);
            twi[off + j] = (float)sin(ang);
        }
        off += (h + 7u) & ~7u;
    }
}

/* Generator::genBitReverse (butterfly.cpp:111-123): dst[i] = 2*bitrev(i). */
ORC_API void orc_gen_bit_reverse(uint32_t n, uint32_t* dst) {
    const uint32_t logn = ilog2u(n);
    for (uint32_t i = 0; i < n; i++) dst[i] = 2u * bitrev(i, logn);
}

/* Generator::genBitReversePlane (butterfly.cpp:126-139). */
ORC_API void orc_gen_bit_reverse_plane(uint32_t n, uint32_t* dst) {
    const uint32_t logn = ilog2u(n), K = orc_plane_k_for(n), rows = n / K, logK = ilog2u(K);
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t j = ((k % rows) << logK) | (k / rows);
        dst[k] = 8u * bitrev(j, logn);
    }
}

/* Generator::genTransposePlane (butterfly.cpp:142-149). */
ORC_API void orc_gen_transpose_plane(uint32_t n, uint32_t* dst) {
    const uint32_t K = orc_plane_k_for(n), rows = n / K, logK = ilog2u(K);
    for (uint32_t k = 0; k < n; k++) dst[k] = (((k & (K - 1u)) * rows) + (k >> logK)) * 4u;
}

/* Generator::genInterleave (butterfly.cpp:154-166), length 2nD. */
ORC_API void orc_gen_interleave(uint32_t n, uint32_t D, uint32_t* dst) {
    if (D == 0u) D = 1u;
    for (uint32_t d = 0; d < D; d++)
        for (uint32_t j = 0; j < n; j++) {
            dst[d * 2u * n + 2u * j] = 4u * (d * n + j);
            dst[d * 2u * n + 2u * j + 1u] = 4u * (D * n + d * n + j);
        }
}

/* genR2CTwiddles (butterfly.cpp:692-703 / fft_check.cpp:512-521):
 * W_n^k = exp(-2*pi*i*k/n), k = 0..n/4, padded to a multiple of 8.
 * Returns the padded length. */
ORC_API uint32_t orc_gen_r2c_twiddles(uint32_t n, float* pwr, float* pwi) {
    const uint32_t pk1 = (n >> 2) + 1u, pad = (pk1 + 7u) & ~7u;
    for (uint32_t k = 0; k < pad; k++) { pwr[k] = 0.f; pwi[k] = 0.f; }
    for (uint32_t k = 0; k < pk1; k++) {
        const double ang = -2.0 * M_PI * (double)k / (double)n;
        pwr[k] = (float)cos(ang);
        pwi[k] = (float)sin(ang);
    }
    return pad;
}

/* ------------------------------------------------------------- long FFT */
/* Two-factor split (src/host/fft_check.cpp:295-308): from the largest power
 * of two a with a*a <= n, first scan upward then downward for 64 <= a,
 * n/a <= 4096.  Returns 0 when no legal split exists. */
ORC_API int orc_long_split(uint32_t n, uint32_t* n1, uint32_t* n2) {
    uint32_t target = 1;
    *n1 = 0; *n2 = 0;
    while ((uint64_t)(target << 1) * (uint64_t)(target << 1) <= n) target <<= 1;
    for (int side = 0; side < 2 && !*n1; side++) {
        for (uint32_t a = side ? (target >> 1) : target; a >= 64 && a <= 4096;
             a = side ? (a >> 1) : (a << 1)) {
            if (n % a) continue;
            const uint32_t b = n / a;
            if (b >= 64 && b <= 4096) { *n1 = a; *n2 = b; break; }
        }
    }
    return *n1 != 0;
}

/* Host four-step chain of src/host/fft_check.cpp:904-963 in FP64:
 * transpose-in hT[b][j][i] = x[b][i*N2+j]; N1-point FFTs over i; twiddle
 * wT[j][k1] = W_N^{j*k1} with transpose to [b][k1][j]; N2-point FFTs over j;
 * natural-order X[k1 + N1*k2].  out is FP64 [batch][2n]. */
ORC_API void orc_four_step_f64(const float* in, double* out, uint32_t n, uint32_t batch,
                               uint32_t n1, uint32_t n2) {
    double* col = (double*)malloc(sizeof(double) * 2 * n1);
    double* row = (double*)malloc(sizeof(double) * 2 * n2);
    double* mid = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    for (uint32_t b = 0; b < batch; b++) {
        const float* x = in + (size_t)b * 2 * n;
        for (uint32_t j = 0; j < n2; j++) {
            for (uint32_t i = 0; i < n1; i++) {
                col[2 * i] = x[2 * ((size_t)i * n2 + j)];
                col[2 * i + 1] = x[2 * ((size_t)i * n2 + j) + 1];
            }
            orc_fft_f64_inplace(col, n1, -1);
            for (uint32_t k1 = 0; k1 < n1; k1++) {
                const double ang = -2.0 * M_PI * (double)(((uint64_t)j * k1) % n) / (double)n;
                const double wr = cos(ang), wi = sin(ang);
                const double ar = col[2 * k1], ai = col[2 * k1 + 1];
                mid[2 * ((size_t)k1 * n2 + j)] = ar * wr - ai * wi;
                mid[2 * ((size_t)k1 * n2 + j) + 1] = ar * wi + ai * wr;
            }
        }
        double* o = out + (size_t)b * 2 * n;
        for (uint32_t k1 = 0; k1 < n1; k1++) {
            memcpy(row, mid + 2 * (size_t)k1 * n2, sizeof(double) * 2 * n2);
            orc_fft_f64_inplace(row, n2, -1);
            for (uint32_t k2 = 0; k2 < n2; k2++) {
                o[2 * ((size_t)k1 + (size_t)n1 * k2)] = row[2 * k2];
                o[2 * ((size_t)k1 + (size_t)n1 * k2) + 1] = row[2 * k2 + 1];
            }
        }
    }
    free(col); free(row); free(mid);
}
