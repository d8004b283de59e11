// bfly_kernels.hpp -- internal launchers of bfly_kernels.hip (not exported).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace bfly_kern {

hipError_t c2c(const float* in, float* out, uint32_t n, uint32_t batch, int sign, int variant,
               const float* table, void* work, hipStream_t s);
// variant: BFLY_VARIANT_PLAIN (one-shot grid) or BFLY_VARIANT_STREAM.
hipError_t real(int kind, const float* in, float* out, uint32_t n, uint32_t batch, uint32_t bins,
                uint32_t stride, int variant, const float* table, hipStream_t s);
hipError_t copy(const void* src, void* dst, size_t bytes, hipStream_t s);
// Split-FP16 matrix-core C2C (bfly_mfma.hip), n = 1024 / 2048 / 4096;
// `stream` selects the resident-grid form.
hipError_t c2c_mfma(const float* in, float* out, uint32_t n, uint32_t batch, int sign, bool stream,
                    const float* table, hipStream_t s);

}  // namespace bfly_kern
