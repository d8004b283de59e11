// bfly_mfma.hip -- C2C with the radix-16 passes on the matrix cores in
// split FP16 (BFLY_VARIANT_SPLIT_MFMA, n = 1024 / 2048 / 4096).
//
// A radix-16 butterfly is the 16x16 complex DFT matrix F times the column of
// its 16 inputs; written over the reals it is a 32x32 matrix times a 32-vector
// (row 2q / 2q+1 = Re / Im of output q, column 2r / 2r+1 = Re / Im of input
// r).  v_mfma_f32_16x16x32_f16 multiplies a 16x32 block of it (outputs q of
// one half) by 16 such columns at once, so two instructions give all 32
// output reals of 16 butterflies.
//
// Split FP16: every operand x is carried as hi = fp16(x) and
// lo = fp16((x - hi) * 2^11), and the product is
//     F x ~ Fhi xhi + 2^-11 (Fhi xlo + Flo xhi)
// with FP32 accumulation (the lo*lo term, ~2^-22 relative, is dropped).
// Each butterfly's inputs are first scaled by a power of two 2^-e that puts
// their largest magnitude in [0.5, 1), so neither half over- nor underflows
// whatever the data's range, and the outputs are scaled back by 2^e (both
// exact).  Inter-pass twiddles stay FP32 on the vector ALUs, as do the
// radix-4 / radix-8 remainder passes.
//
// Lane layout of one MFMA pass: a wave owns 64 butterflies j = 64 w + 16 i
// + c (i < 4, c = lane mod 16); lane group h = lane / 16 holds inputs
// r = 4h .. 4h+3 of each of its four butterflies and, after the product,
// outputs q = 2h, 2h+1, 8+2h, 9+2h.  The LDS exchanges between passes are in
// natural Stockham order, so MFMA passes and vector passes mix freely.
#include <hip/hip_runtime.h>

#include <utility>

#include "bfly_device.hpp"
#include "bfly_hip.h"
#include "bfly_kernels.hpp"
#include "bfly_plans.hpp"
#include "bfly_units.hpp"

using namespace bfly_dev;

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));

constexpr float kLoScale = 2048.0f;        // 2^11
constexpr float kLoUnscale = 1.0f / 2048.0f;

// hi / lo halves of x (|x| < 1 after the butterfly scale).
__device__ __forceinline__ void split16(float x, _Float16& hi, _Float16& lo) {
    hi = (_Float16)x;
    lo = (_Float16)((x - (float)hi) * kLoScale);
}

// The A operands of one lane: rows 4-aligned block `rb` (outputs q = 8 rb ..
// 8 rb + 7), row o = lane mod 16 (q = 8 rb + o / 2, Re for even o), columns
// k = 8 h + kk (input r = 4 h + kk / 2, Re for even kk).
template <int DIR>
__device__ __forceinline__ void dft16_operands(int lane, half8 (&ahi)[2], half8 (&alo)[2]) {
    const int o = lane & 15, h = lane >> 4;
#pragma unroll
    for (int rb = 0; rb < 2; rb++) {
        const int q = 8 * rb + (o >> 1);
        const bool im_row = o & 1;
#pragma unroll
        for (int kk = 0; kk < 8; kk++) {
            const int r = 4 * h + (kk >> 1);
            const bool im_col = kk & 1;
            // F[q][r] = e^{DIR 2 pi i q r / 16} = cr + i ci
            const float cr = cos16(q * r);
            const float ci = DIR < 0 ? -sin16(q * r) : sin16(q * r);
            // Re y += cr Re x - ci Im x ; Im y += ci Re x + cr Im x
            const float a = im_row ? (im_col ? cr : ci) : (im_col ? -ci : cr);
            _Float16 hi, lo;
            split16(a, hi, lo);
            ahi[rb][kk] = hi;
            alo[rb][kk] = lo;
        }
    }
}

// Radix-16 DFT of the lane's four butterflies: in v[4i + rr] = input
// r = 4h + rr of butterfly i; out v[4i + s] = output q = 8 (s / 2) + 2h + s % 2.
__device__ __forceinline__ void dft16_mfma(float2 (&v)[16], const half8 (&ahi)[2], const half8 (&alo)[2]) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
        float m = 0.f;
#pragma unroll
        for (int rr = 0; rr < 4; rr++) m = fmaxf(m, fmaxf(fabsf(v[4 * i + rr].x), fabsf(v[4 * i + rr].y)));
        m = fmaxf(m, __shfl_xor(m, 16));
        m = fmaxf(m, __shfl_xor(m, 32));
        int e = 0;
        (void)frexpf(m, &e);  // m = f 2^e, f in [0.5, 1); e = 0 for m = 0
        const float down = ldexpf(1.0f, -e), up = ldexpf(1.0f, e);
        half8 bhi, blo;
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
            _Float16 h0, l0, h1, l1;
            split16(v[4 * i + rr].x * down, h0, l0);
            split16(v[4 * i + rr].y * down, h1, l1);
            bhi[2 * rr] = h0;
            bhi[2 * rr + 1] = h1;
            blo[2 * rr] = l0;
            blo[2 * rr + 1] = l1;
        }
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
            float4v acc = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi[rb], blo, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo[rb], bhi, acc, 0, 0, 0);
            acc *= kLoUnscale;
            acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi[rb], bhi, acc, 0, 0, 0);
            v[4 * i + 2 * rb] = make_float2(acc[0] * up, acc[1] * up);
            v[4 * i + 2 * rb + 1] = make_float2(acc[2] * up, acc[3] * up);
        }
    }
}

// One radix-16 MFMA pass with Ns points already combined.  In: the lane's
// inputs in v (see dft16_mfma); multiplies by the pass twiddles (Ns > 1),
// runs the DFT and, unless LAST, writes the outputs to `lds` in Stockham
// order.  The caller synchronises.
template <int N, int Ns, int DIR, bool LAST>
__device__ __forceinline__ void pass16_mfma(float2 (&v)[16], float2* lds, const float2* __restrict__ tw,
                                            int w, int lane, const half8 (&ahi)[2], const half8 (&alo)[2]) {
    const int h = lane >> 4, c = lane & 15;
    if constexpr (Ns > 1) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int j = 64 * w + 16 * i + c;
            const int k = j & (Ns - 1);
            const float4* w4 = reinterpret_cast<const float4*>(tw + k * 16 + 4 * h);
#pragma unroll
            for (int q2 = 0; q2 < 2; q2++) {
                const float4 t = w4[q2];
                v[4 * i + 2 * q2] = cmul_tw<DIR>(v[4 * i + 2 * q2], make_float2(t.x, t.y));
                v[4 * i + 2 * q2 + 1] = cmul_tw<DIR>(v[4 * i + 2 * q2 + 1], make_float2(t.z, t.w));
            }
        }
    }
    dft16_mfma(v, ahi, alo);
    if constexpr (!LAST) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int j = 64 * w + 16 * i + c;
            const int k = j & (Ns - 1);
            const int base = (j - k) * 16 + k;
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int q = 8 * (s >> 1) + 2 * h + (s & 1);
                lds[lds_idx(base + q * Ns)] = v[4 * i + s];
            }
        }
    }
}

// Reads the lane's MFMA-pass inputs x[j + r N/16] from `src` (LDS, padded).
template <int N>
__device__ __forceinline__ void read16_lds(float2 (&v)[16], const float2* lds, int w, int lane) {
    const int h = lane >> 4, c = lane & 15;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int rr = 0; rr < 4; rr++)
            v[4 * i + rr] = lds[lds_idx(64 * w + 16 * i + c + (4 * h + rr) * (N / 16))];
}

// Plans [16, 16, R2]: two MFMA passes, then radix R2 -- on the matrix cores
// when R2 = 16, on the vector ALUs otherwise.
template <int I, int DIR>
__global__ __launch_bounds__(bfly_plan::kBlock) void k_c2c_mfma(const float2* __restrict__ in,
                                                              float2* __restrict__ out,
                                                              const float2* __restrict__ tw,
                                                              uint32_t batch) {
    using PL = PlanOf<I>;
    constexpr int N = PL::N, T = PL::T, B = bfly_plan::kBlock / T;
    constexpr int R2 = bfly_plan::kPlans[I].rad[2];
    static_assert(bfly_plan::kPlans[I].nrad == 3 && bfly_plan::kPlans[I].rad[0] == 16 &&
                      bfly_plan::kPlans[I].rad[1] == 16 && T % 64 == 0,
                  "plans [16, 16, R2] with whole waves per transform");
    __shared__ float2 lds_all[B * lds_len(N)];
    const int tid = threadIdx.x;
    const int bl = tid / T, t = tid % T;
    const int w = t >> 6, lane = t & 63, h = lane >> 4, c = lane & 15;
    const uint64_t tr = (uint64_t)blockIdx.x * B + bl;
    const bool active = tr < batch;
    float2* lds = lds_all + bl * lds_len(N);
    half8 ahi[2], alo[2];
    dft16_operands<DIR>(lane, ahi, alo);

    float2 v[16];
    if (active) {
        const float2* src = in + tr * N;
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int rr = 0; rr < 4; rr++)
                v[4 * i + rr] = ld_io<true>(&src[64 * w + 16 * i + c + (4 * h + rr) * (N / 16)]);
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = make_float2(0.f, 0.f);
    }
    // pass 0: Ns = 1
    pass16_mfma<N, 1, DIR, false>(v, lds, tw, w, lane, ahi, alo);
    __syncthreads();
    read16_lds<N>(v, lds, w, lane);
    __syncthreads();
    // pass 1: Ns = 16, table of the pass at tw[0 .. 256)
    pass16_mfma<N, 16, DIR, false>(v, lds, tw, w, lane, ahi, alo);
    __syncthreads();
    const float2* tw2 = tw + 16 * 16;
    if constexpr (R2 == 16) {
        read16_lds<N>(v, lds, w, lane);
        __syncthreads();
        pass16_mfma<N, 256, DIR, true>(v, lds, tw2, w, lane, ahi, alo);
        if (!active) return;
        float2* dst = out + tr * N;
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int q = 8 * (s >> 1) + 2 * h + (s & 1);
                st_io<true>(&dst[64 * w + 16 * i + c + q * (N / 16)], v[4 * i + s]);
            }
    } else {
        // vector radix-R2 pass: thread t owns butterflies j = t + m T
        constexpr int M = 16 / R2;
#pragma unroll
        for (int m = 0; m < M; m++)
#pragma unroll
            for (int r = 0; r < R2; r++) v[m * R2 + r] = lds[lds_idx(t + m * T + r * (N / R2))];
        constexpr int Ns = 256;
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int k = (t + m * T) & (Ns - 1);
#pragma unroll
            for (int r = 1; r < R2; r++) v[m * R2 + r] = cmul_tw<DIR>(v[m * R2 + r], tw2[k * R2 + r]);
            dft<R2, DIR>(&v[m * R2]);
        }
        if (!active) return;
        float2* dst = out + tr * N;
#pragma unroll
        for (int m = 0; m < M; m++)
#pragma unroll
            for (int r = 0; r < R2; r++) st_io<true>(&dst[t + m * T + r * (N / R2)], v[m * R2 + r]);
    }
}

template <int I>
hipError_t launch_mfma(const float2* in, float2* out, const float2* tw, uint32_t batch, int sign,
                       hipStream_t s) {
    constexpr int B = bfly_plan::kBlock / PlanOf<I>::T;
    const dim3 grid((unsigned)((batch + B - 1) / B));
    if (sign < 0)
        hipLaunchKernelGGL((k_c2c_mfma<I, -1>), grid, dim3(bfly_plan::kBlock), 0, s, in, out, tw, batch);
    else
        hipLaunchKernelGGL((k_c2c_mfma<I, 1>), grid, dim3(bfly_plan::kBlock), 0, s, in, out, tw, batch);
    return hipGetLastError();
}

constexpr int idx(int n) { return bfly_plan::plan_index((uint32_t)n); }

}  // namespace

namespace bfly_kern {

hipError_t c2c_mfma(const float* in, float* out, uint32_t n, uint32_t batch, int sign,
                    const float* table, hipStream_t s) {
    const float2* i2 = reinterpret_cast<const float2*>(in);
    float2* o2 = reinterpret_cast<float2*>(out);
    const float2* tw = reinterpret_cast<const float2*>(table);
    switch (n) {
        case 1024: return launch_mfma<idx(1024)>(i2, o2, tw, batch, sign, s);
        case 2048: return launch_mfma<idx(2048)>(i2, o2, tw, batch, sign, s);
        case 4096: return launch_mfma<idx(4096)>(i2, o2, tw, batch, sign, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace bfly_kern
