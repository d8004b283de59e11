// bfly_kernels.hpp -- internal launchers of bfly_kernels.hip (not exported).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <atomic>

namespace bfly_kern {

// Resident workgroups of a kernel on the current device (CUs x the
// occupancy query), cached per device: the launchers size their persistent
// grids with it.  Safe to share between host threads.
struct ResidentCache {
    static constexpr int kDevices = 64;
    std::atomic<int> per_device[kDevices] = {};
};
// 0 when the runtime cannot answer.
int resident(ResidentCache& cache, const void* kernel, int block);

// `mid`, when given, is recorded on `s` between the two launches of the
// four-step chain (unused by the one-launch paths).  `scale` multiplies
// every output in the final store; the LDS-twiddle streaming kernels
// (BFLY_VARIANT_STREAM_LT, and 8192 / 16384 under AUTO) and the four-step
// chain apply it, the other variants require 1.
// Batched C2C over rows `idist` / `odist` complex values apart (>= n):
// the LDS-twiddle streaming kernel for n 16..16384, the four-step chain
// (contiguous workspace `work`) for 32768 / 65536; `table` as for c2c.
hipError_t c2c_rows(const float* in, uint64_t idist, float* out, uint64_t odist, uint32_t n, uint32_t batch,
                    int sign, const float* table, void* work, hipStream_t s);
hipError_t c2c(const float* in, float* out, uint32_t n, uint32_t batch, int sign, int variant,
               const float* table, void* work, hipStream_t s, hipEvent_t mid = nullptr, float scale = 1.0f);
// variant: BFLY_VARIANT_PLAIN (one-shot grid), BFLY_VARIANT_STREAM or
// BFLY_VARIANT_STREAM_LT; `adjoint` runs the transform's adjoint (R2C: the
// C2R's adjoint weighting of the bins; C2R: the R2C's).
hipError_t real(int kind, const float* in, float* out, uint32_t n, uint32_t batch, uint32_t bins,
                uint32_t stride, int variant, bool adjoint, const float* table, hipStream_t s);
hipError_t copy(const void* src, void* dst, size_t bytes, hipStream_t s);
// C2C along the first axis of [batch][n1][ncols] (bfly_fft2d.hip); `work`
// holds n1*ncols*batch complex values when n1 > 256 (unused otherwise).
hipError_t columns(const float* in, float* out, uint32_t n1, uint32_t ncols, uint32_t batch, int sign,
                   const float* table_cols, void* work, hipStream_t s);
// 2-D C2C of [batch][n1][n2] (bfly_fft2d.hip); `work` holds n1*n2*batch
// complex values.
hipError_t fft2d(const float* in, float* out, uint32_t n1, uint32_t n2, uint32_t batch, int sign,
                 const float* table_rows, const float* table_cols, void* work, hipStream_t s);
// Per-mode channel mixing of the spectral convolution (bfly_mix.hip);
// op 0 forward, 1 input gradient, 2 weight gradient (see bfly_mode_mix).
hipError_t mode_mix(int op, const float* a, const float* b, float* c, uint32_t batch, uint32_t c_in,
                    uint32_t c_out, uint32_t modes, hipStream_t s);
// Split-FP16 matrix-core C2C (bfly_mfma.hip), n = 1024 / 2048 / 4096;
// `stream` selects the resident-grid form.
hipError_t c2c_mfma(const float* in, float* out, uint32_t n, uint32_t batch, int sign, bool stream,
                    const float* table, hipStream_t s);

}  // namespace bfly_kern
