// inputs.hpp -- the deterministic test inputs of the bfly_check program:
// the input modes of the reference's fft_check (src/host/fft_check.cpp:55-192,
// AB_INPUT), generated on the host.
#pragma once
#include <cstddef>
#include <cstdint>

namespace bfly_check {

// True for every AB_INPUT mode name the reference accepts
// (src/host/fft_check.cpp:70-80).
bool is_input_mode(const char* mode);

// C2C input, `batch` interleaved rows of n complex values (2n floats each):
// "sin" / "deterministic-sin-cos" over the flat float index, the three
// application shapes ("ofdm", "radar", "dl"), or a numerical mode; `seed`
// feeds "random" / "random-seeded" (src/host/fft_check.cpp:170-179).
void c2c_input(float* h, uint32_t n, uint32_t batch, const char* mode, uint32_t seed);

// R2C input, `batch` rows of n reals: the sin/cos sequence over the flat
// index, else the real parts of the C2C input of the same mode
// (src/host/fft_check.cpp:395-410).
void r2c_input(float* h, uint32_t n, uint32_t batch, const char* mode, uint32_t seed);

// C2R input: a seeded half spectrum of n/2+1 bins per row at a row stride
// of `stride` floats, imaginary part 0 at the Nyquist bin; the floats past
// the bins are left as they are (src/host/fft_check.cpp:179-192).
void half_input(float* h, uint32_t n, uint32_t batch, size_t stride);

}  // namespace bfly_check
