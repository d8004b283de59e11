// FP64 reference transforms, the shared error metric and the deterministic
// inputs used by Plan::measure and by callers' own acceptance checks.
// Same declarations and contracts as the reference's
// include/butterfly/reference.hpp:1-62 so code written against it compiles
// unchanged; the implementation (ascend-fft_amd/csrc/framework/reference.cpp)
// is this library's own.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>

namespace bfly {

// Forward C2C of one interleaved row (2n floats), FP64 arithmetic, fp32 out.
void refFftF32(const float* in, float* out, uint32_t n);

// R2C: n reals -> n/2+1 interleaved bins (numpy.fft.rfft layout, n+2 floats).
void refR2CF32(const float* in, float* out, uint32_t n);

// C2R: n+2-float half spectrum -> n reals scaled 1/n (numpy.fft.irfft); only
// the real parts of the DC and Nyquist bins contribute.
void refC2RF32(const float* in, float* out, uint32_t n);

// Mixed absolute/relative error: max |got-ref| / max(1, max |ref|) over
// float components; non-finite values give +inf (reference.hpp:30-41).
inline double maxRelScaled(const float* got, const float* ref, size_t len) {
    if (len && (!got || !ref)) return std::numeric_limits<double>::infinity();
    double peak = 0.0, worst = 0.0;
    for (size_t i = 0; i < len; i++) {
        const double g = got[i], r = ref[i];
        if (!std::isfinite(g) || !std::isfinite(r)) return std::numeric_limits<double>::infinity();
        peak = std::max(peak, std::fabs(r));
        worst = std::max(worst, std::fabs(g - r));
    }
    return worst / std::max(1.0, peak);
}

// Deterministic input pattern (reference.cpp:10-12).
float patternAt(size_t i);

// Up to `want` batch indices spread evenly over [0, batch-1], both ends
// included (reference.hpp:48-60).
inline uint32_t sampleBatches(uint32_t batch, uint32_t want, uint32_t* out, uint32_t maxOut) {
    want = std::min(want, maxOut);
    if (batch == 0 || want == 0) return 0;
    if (batch <= want) {
        for (uint32_t i = 0; i < batch; i++) out[i] = i;
        return batch;
    }
    if (want == 1) {
        out[0] = 0;
        return 1;
    }
    const uint64_t span = (uint64_t)batch - 1;
    for (uint32_t i = 0; i < want; i++) out[i] = (uint32_t)(span * i / (want - 1));
    return want;
}

}  // namespace bfly
