// bfly_mix.hip -- per-mode complex channel mixing of the 1-D spectral
// convolution, in the layouts the mode-truncated R2C writes and the C2R
// reads (mode index fastest), so no transposed copy of the modes or of the
// weights is ever made.
//
//   C[m][n][k] = sum_r  opA(A)[m][r][k] * opB(B)[r][n][k],   k < K,
//
// complex64 interleaved, every operand with unit stride along the mode k and
// arbitrary strides along its two channel indices; opA / opB optionally
// conjugate.  The three products of the layer (forward, input gradient,
// weight gradient) are this one kernel with different strides.
//
// Tiling: a 256-thread workgroup owns 16 modes x 16 rows (m) x 16 columns
// (n) of C and walks the reduction index r in chunks of 8.  A chunk of A
// (8 r x 16 m x 16 k) and of B (8 r x 16 n x 16 k) is read with the mode
// index across 16 consecutive lanes (128-byte segments), parked in LDS in
// the same order, and reread by every thread: thread t owns mode k = t % 16
// and the 4 x 4 outputs m = gm + 4i, n = gn + 4j (g = t / 16, gm = g % 4,
// gn = g / 4).  The next chunk is loaded into registers while the current
// one is consumed.  The products are fp32 multiply-adds on the vector ALUs
// (four per complex product), accumulated in fp32 over r.
#include <hip/hip_runtime.h>

#include "bfly_kernels.hpp"

namespace {

constexpr int kKT = 16;  // modes per workgroup (lanes along k)
constexpr int kMT = 16;  // rows of C per workgroup
constexpr int kNT = 16;  // columns of C per workgroup
constexpr int kRT = 8;   // reduction chunk
constexpr int kThreads = 256;
static_assert(kKT * kMT == kThreads && kKT * kNT == kThreads, "one A / B element per thread per r");

__device__ __forceinline__ float2 cmac(float2 acc, float2 a, float2 b) {
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(-a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(a.y, b.x, acc.y);
    return acc;
}

__global__ __launch_bounds__(kThreads) void k_mix(const float2* __restrict__ A, const float2* __restrict__ B,
                                                  float2* __restrict__ C, uint32_t M, uint32_t N, uint32_t R,
                                                  uint32_t K, uint64_t sam, uint64_t sar, uint64_t sbr,
                                                  uint64_t sbn, uint64_t scm, uint64_t scn, uint32_t ktiles,
                                                  uint32_t mtiles, float conj_a, float conj_b) {
    __shared__ float2 As[kRT * kMT * kKT];  // [r][m][k]
    __shared__ float2 Bs[kRT * kNT * kKT];  // [r][n][k]
    const int t = threadIdx.x;
    const int kk = t % kKT, g = t / kKT, gm = g % 4, gn = g / 4;
    // block -> (k tile, m tile, n tile), k tile fastest
    uint32_t b = blockIdx.x;
    const uint32_t kt = b % ktiles;
    b /= ktiles;
    const uint32_t mt = b % mtiles, nt = b / mtiles;
    const uint32_t k = kt * kKT + kk;
    const bool kin = k < K;
    // this thread's loads: A[m0 + g][r0 + j][k], B[r0 + j][n0 + g][k], j < kRT
    const uint32_t la = mt * kMT + g, lb = nt * kNT + g;
    const bool ina = kin && la < M, inb = kin && lb < N;
    const float2* pa = A + (uint64_t)la * sam + k;
    const float2* pb = B + (uint64_t)lb * sbn + k;

    float2 ra[kRT], rb[kRT];
    auto load = [&](uint32_t r0) {
#pragma unroll
        for (int j = 0; j < kRT; j++) {
            const uint32_t r = r0 + j;
            float2 a = make_float2(0.f, 0.f), v = make_float2(0.f, 0.f);
            if (ina && r < R) a = pa[(uint64_t)r * sar];
            if (inb && r < R) v = pb[(uint64_t)r * sbr];
            ra[j] = make_float2(a.x, a.y * conj_a);
            rb[j] = make_float2(v.x, v.y * conj_b);
        }
    };

    float2 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = make_float2(0.f, 0.f);

    load(0);
    for (uint32_t r0 = 0; r0 < R; r0 += kRT) {
#pragma unroll
        for (int j = 0; j < kRT; j++) {
            As[j * kThreads + t] = ra[j];
            Bs[j * kThreads + t] = rb[j];
        }
        __syncthreads();
        if (r0 + kRT < R) load(r0 + kRT);
#pragma unroll
        for (int r = 0; r < kRT; r++) {
            float2 a[4], v[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = As[r * kThreads + (gm + 4 * i) * kKT + kk];
#pragma unroll
            for (int j = 0; j < 4; j++) v[j] = Bs[r * kThreads + (gn + 4 * j) * kKT + kk];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = cmac(acc[i][j], a[i], v[j]);
        }
        __syncthreads();
    }
    if (!kin) return;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t m = mt * kMT + gm + 4 * i;
        if (m >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t n = nt * kNT + gn + 4 * j;
            if (n < N) C[(uint64_t)m * scm + (uint64_t)n * scn + k] = acc[i][j];
        }
    }
}

}  // namespace

namespace bfly_kern {

hipError_t mode_mix(int op, const float* a, const float* b, float* c, uint32_t batch, uint32_t c_in,
                    uint32_t c_out, uint32_t modes, hipStream_t s) {
    const uint64_t K = modes;
    uint32_t M, N, R;
    uint64_t sam, sar, sbr, sbn, scm, scn;
    float ca = 1.f, cb = 1.f;
    if (op == 0) {  // y[b][o] = sum_i x[b][i] w[i][o]
        M = batch, N = c_out, R = c_in;
        sam = c_in * K, sar = K, sbr = c_out * K, sbn = K, scm = c_out * K, scn = K;
    } else if (op == 1) {  // gx[b][i] = sum_o gy[b][o] conj(w[i][o])
        M = batch, N = c_in, R = c_out;
        sam = c_out * K, sar = K, sbr = K, sbn = c_out * K, scm = c_in * K, scn = K;
        cb = -1.f;
    } else if (op == 2) {  // gw[i][o] = sum_b conj(x[b][i]) gy[b][o]
        M = c_in, N = c_out, R = batch;
        sam = K, sar = c_in * K, sbr = c_out * K, sbn = K, scm = c_out * K, scn = K;
        ca = -1.f;
    } else {
        return hipErrorInvalidValue;
    }
    const uint32_t kt = (uint32_t)((K + kKT - 1) / kKT), mt = (M + kMT - 1) / kMT, nt = (N + kNT - 1) / kNT;
    const uint64_t blocks = (uint64_t)kt * mt * nt;
    if (blocks == 0 || blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mix, dim3((unsigned)blocks), dim3(kThreads), 0, s, reinterpret_cast<const float2*>(a),
                       reinterpret_cast<const float2*>(b), reinterpret_cast<float2*>(c), M, N, R, (uint32_t)K,
                       sam, sar, sbr, sbn, scm, scn, kt, mt, ca, cb);
    return hipGetLastError();
}

}  // namespace bfly_kern
