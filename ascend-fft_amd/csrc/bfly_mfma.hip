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
// Split FP16: every operand x is carried as hi + lo, hi = x rounded toward
// zero to FP16 precision and lo = fp16(x - hi), and the product is
//     F x ~ Fhi xhi + Fhi xlo + Flo xhi
// with FP32 accumulation (the lo*lo term, ~2^-22 relative, is dropped).  The
// data enter scaled by one power of two per transform: its inputs are below
// 2^e (e from the largest component: a wave reduction, plus an LDS exchange
// when a transform spans several waves), every radix-16 pass grows a
// component by less than 16 sqrt(2) < 2^5, so with P matrix passes the
// scale 2^(11 - e - 5(P-1)) keeps every pass's inputs below 2^11 -- hi never
// overflows FP16 whatever the data's range -- and the outputs are scaled
// back by the same power of two (both exact).  Inter-pass twiddles stay FP32
// on the vector ALUs, as do the radix-4 / radix-8 remainder passes.
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

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// cos(2 pi m / 16), m = 0..15, indexed at run time by the lane's (q, r).
__device__ const float kCos16[16] = {
    1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f,
    0.0f, -0.38268343236508977173f, -0.70710678118654752440f, -0.92387953251128675613f,
    -1.0f, -0.92387953251128675613f, -0.70710678118654752440f, -0.38268343236508977173f,
    0.0f, 0.38268343236508977173f, 0.70710678118654752440f, 0.92387953251128675613f};

// The A operands of one lane: rows 4-aligned block `rb` (outputs q = 8 rb ..
// 8 rb + 7), row o = lane mod 16 (q = 8 rb + o / 2, Re for even o), columns
// k = 8 h + kk (input r = 4 h + kk / 2, Re for even kk).  hi = fp16(a),
// lo = fp16(a - hi).
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
            const int m = (q * r) & 15;
            const float cr = kCos16[m];
            const float sn = kCos16[(m - 4) & 15];  // sin(2 pi m / 16)
            const float ci = DIR < 0 ? -sn : sn;
            // Re y += cr Re x - ci Im x ; Im y += ci Re x + cr Im x
            const float a = im_row ? (im_col ? cr : ci) : (im_col ? -ci : cr);
            const _Float16 hi = (_Float16)a;
            ahi[rb][kk] = hi;
            alo[rb][kk] = (_Float16)(a - (float)hi);
        }
    }
}

// Largest |component| of the lane's 16 values over the wave, as the exponent
// e with max < 2^e (0 for all-zero data).
__device__ __forceinline__ int wave_max_exponent(const float2 (&v)[16]) {
    float m = 0.f;
#pragma unroll
    for (int i = 0; i < 16; i++) m = fmaxf(m, fmaxf(fabsf(v[i].x), fabsf(v[i].y)));
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) m = fmaxf(m, __shfl_xor(m, d));
    int e = 0;
    (void)frexpf(m, &e);  // m = f 2^e, f in [0.5, 1)
    return e;
}

// Radix-16 DFT of the lane's four butterflies (data already scaled): in
// v[4i + rr] = input r = 4h + rr of butterfly i; out v[4i + s] = output
// q = 8 (s / 2) + 2h + s % 2.
__device__ __forceinline__ void dft16_mfma(float2 (&v)[16], const half8 (&ahi)[2], const half8 (&alo)[2]) {
    // hi = x with its low 13 mantissa bits cleared (x rounded toward zero to
    // FP16 precision, exact in FP16 for normal magnitudes), lo = x - hi exact
    // in FP32, both packed two per dword
    half8 bhi[4], blo[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        u32x4 ph, pl;
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
            const float x = v[4 * i + rr].x, y = v[4 * i + rr].y;
            const float xh = __uint_as_float(__float_as_uint(x) & 0xFFFFE000u);
            const float yh = __uint_as_float(__float_as_uint(y) & 0xFFFFE000u);
            ph[rr] = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(xh, yh));
            pl[rr] = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(x - xh, y - yh));
        }
        bhi[i] = __builtin_bit_cast(half8, ph);
        blo[i] = __builtin_bit_cast(half8, pl);
    }
    // small terms first; eight independent accumulators per product
    float4v acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
            acc[i][rb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi[rb], blo[i], float4v{0.f, 0.f, 0.f, 0.f},
                                                                0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
            acc[i][rb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo[rb], bhi[i], acc[i][rb], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
            acc[i][rb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi[rb], bhi[i], acc[i][rb], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
            v[4 * i + 2 * rb] = make_float2(acc[i][rb][0], acc[i][rb][1]);
            v[4 * i + 2 * rb + 1] = make_float2(acc[i][rb][2], acc[i][rb][3]);
        }
}

// One radix-16 MFMA pass with Ns points already combined.  In: the lane's
// inputs in v (see dft16_mfma); multiplies by the pass twiddles (Ns > 1;
// entry [k][q] of the pass table at tw[k PITCH + q]), runs the DFT and,
// unless LAST, writes the outputs to `lds` in Stockham order.  The caller
// synchronises.
template <int N, int Ns, int DIR, bool LAST, int PITCH = 16, bool TWO_LEVEL = false>
__device__ __forceinline__ void pass16_mfma(float2 (&v)[16], float2* lds, const float2* __restrict__ tw,
                                            int w, int lane, const half8 (&ahi)[2], const half8 (&alo)[2],
                                            const float2* ab = nullptr) {
    const int h = lane >> 4, c = lane & 15;
    if constexpr (TWO_LEVEL) {
        // W_N^{q k} for the lane's q = 4h .. 4h+3 from the two 64-entry LDS
        // tables ab = [W_N^j, j < 64 | W_N^{64 j}]: the first by one product,
        // the next three by successive multiplication with W_N^k
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int j = 64 * w + 16 * i + c;
            const int k = j & (Ns - 1);
            const int m0 = 4 * h * k;  // < 16 Ns = N
            const float2 a0 = ab[m0 & 63], b0 = ab[64 + (m0 >> 6)];
            const float2 a1 = ab[k & 63], b1 = ab[64 + (k >> 6)];
            float2 wq = make_float2(a0.x * b0.x - a0.y * b0.y, a0.x * b0.y + a0.y * b0.x);
            const float2 ws = make_float2(a1.x * b1.x - a1.y * b1.y, a1.x * b1.y + a1.y * b1.x);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                v[4 * i + e] = cmul_tw<DIR>(v[4 * i + e], wq);
                wq = make_float2(wq.x * ws.x - wq.y * ws.y, wq.x * ws.y + wq.y * ws.x);
            }
        }
    } else if constexpr (Ns > 1) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int j = 64 * w + 16 * i + c;
            const int k = j & (Ns - 1);
            const float4* w4 = reinterpret_cast<const float4*>(tw + k * PITCH + 4 * h);
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
            const int b2 = (j - k) * 16 + k + 2 * h * Ns;  // element q = 2h
#pragma unroll
            for (int s = 0; s < 4; s++) {
                // q - 2h = 8 (s / 2) + s % 2 is a compile-time offset
                const int dq = (8 * (s >> 1) + (s & 1)) * Ns;
                if constexpr (Ns >= 16) {
                    // dq is a multiple of 16: pad(b2 + dq) = pad(b2) + dq * 17 / 16
                    lds[lds_idx(b2) + dq + (dq >> 4)] = v[4 * i + s];
                } else {
                    lds[lds_idx(b2 + dq)] = v[4 * i + s];
                }
            }
        }
    }
}

// Reads the lane's MFMA-pass inputs x[j + r N/16] from `src` (LDS, padded).
template <int N>
__device__ __forceinline__ void read16_lds(float2 (&v)[16], const float2* lds, int w, int lane) {
    const int h = lane >> 4, c = lane & 15;
    // N/16 is a multiple of 16, so every offset past the lane's first element
    // pads by a compile-time amount
    const int b = lds_idx(64 * w + c + 4 * h * (N / 16));
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
            const int d = 16 * i + rr * (N / 16);  // multiple of 16
            v[4 * i + rr] = lds[b + d + (d >> 4)];
        }
}

// The lane's pass-0 inputs x[j + r N/16] straight from global memory.
template <int N>
__device__ __forceinline__ void load16_global(float2 (&v)[16], const float2* __restrict__ src, bool active,
                                              int w, int lane) {
    const int h = lane >> 4, c = lane & 15;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int rr = 0; rr < 4; rr++)
            v[4 * i + rr] = active ? ld_io<true>(&src[64 * w + 16 * i + c + (4 * h + rr) * (N / 16)])
                                   : make_float2(0.f, 0.f);
}

// One transform of plan [16, 16, R2] whose pass-0 inputs are in v: two MFMA
// passes, then radix R2 -- on the matrix cores when R2 = 16, on the vector
// ALUs otherwise -- and the natural-order store to dst (when active).
// `lds` is the transform's region, `s_exp` its per-wave exponent slots; the
// second pass reads its table from tw1 at a row pitch of PITCH1 (tw itself,
// or a copy of those 256 entries, same values, in LDS).
template <int I, int DIR, int PITCH1 = 16>
__device__ __forceinline__ void mfma_transform(float2 (&v)[16], float2* lds, int* s_exp,
                                               const float2* __restrict__ tw, const float2* tw1,
                                               const float2* ab2, const half8 (&ahi)[2],
                                               const half8 (&alo)[2], int t, float2* __restrict__ dst,
                                               bool active) {
    using PL = PlanOf<I>;
    constexpr int N = PL::N, T = PL::T;
    constexpr int R2 = bfly_plan::kPlans[I].rad[2];
    static_assert(bfly_plan::kPlans[I].nrad == 3 && bfly_plan::kPlans[I].rad[0] == 16 &&
                      bfly_plan::kPlans[I].rad[1] == 16 && T % 64 == 0,
                  "plans [16, 16, R2] with whole waves per transform");
    constexpr int WAVES = T / 64;  // waves per transform
    const int w = t >> 6, lane = t & 63, h = lane >> 4, c = lane & 15;
    // one scale per transform: the largest input component over its waves
    constexpr int PM = R2 == 16 ? 3 : 2;  // matrix passes
    int et = wave_max_exponent(v);
    if constexpr (WAVES > 1) {
        if (lane == 0) s_exp[w] = et;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < WAVES; k++) et = max(et, s_exp[k]);
    }
    int sh = 11 - et - 5 * (PM - 1);
    sh = sh > 126 ? 126 : (sh < -126 ? -126 : sh);
    const float down = ldexpf(1.0f, sh), up = ldexpf(1.0f, -sh);
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = make_float2(v[i].x * down, v[i].y * down);
    // pass 0: Ns = 1
    pass16_mfma<N, 1, DIR, false>(v, lds, tw, w, lane, ahi, alo);
    __syncthreads();
    read16_lds<N>(v, lds, w, lane);
    __syncthreads();
    // pass 1: Ns = 16, table of the pass at tw[0 .. 256)
    pass16_mfma<N, 16, DIR, false, PITCH1>(v, lds, tw1, w, lane, ahi, alo);
    __syncthreads();
    const float2* tw2 = tw + 16 * 16;
    if constexpr (R2 == 16) {
        read16_lds<N>(v, lds, w, lane);
        __syncthreads();
        pass16_mfma<N, 256, DIR, true, 16, true>(v, lds, tw2, w, lane, ahi, alo, ab2);
        if (!active) return;
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int q = 8 * (s >> 1) + 2 * h + (s & 1);
                st_io<true>(&dst[64 * w + 16 * i + c + q * (N / 16)],
                            make_float2(v[4 * i + s].x * up, v[4 * i + s].y * up));
            }
    } else {
        // vector radix-R2 pass: thread t owns butterflies j = t + m T
        constexpr int M = 16 / R2;
#pragma unroll
        for (int m = 0; m < M; m++)
#pragma unroll
            for (int r = 0; r < R2; r++) v[m * R2 + r] = lds[lds_idx(t + m * T + r * (N / R2))];
        __syncthreads();  // a caller may reuse `lds` right after
        constexpr int Ns = 256;
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int k = (t + m * T) & (Ns - 1);
#pragma unroll
            for (int r = 1; r < R2; r++) {
                // W_N^{r k} = A[r k mod 64] B[r k / 64] (r k < N)
                const int e = r * k;
                const float2 a = ab2[e & 63], b = ab2[64 + (e >> 6)];
                v[m * R2 + r] = cmul_tw<DIR>(v[m * R2 + r], make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x));
            }
            dft<R2, DIR>(&v[m * R2]);
        }
        if (!active) return;
#pragma unroll
        for (int m = 0; m < M; m++)
#pragma unroll
            for (int r = 0; r < R2; r++)
                st_io<true>(&dst[t + m * T + r * (N / R2)], make_float2(v[m * R2 + r].x * up, v[m * R2 + r].y * up));
    }
}

// One-shot grid: B = 256/T transforms per workgroup.
// The last pass's twiddles come from the two tables [W_N^j, j < 64 |
// W_N^{64 j}, j < N/64] at the end of the local table, copied into LDS once
// per workgroup (64 + N/64 <= 128 entries).
template <int I>
__device__ __forceinline__ void fill_ab2(float2* ab2, const float2* __restrict__ tw, int tid) {
    constexpr int LEN = 64 + PlanOf<I>::N / 64;
    static_assert(LEN <= 128, "two tables of at most 64 entries");
    const float2* ab = tw + bfly_plan::stage_twiddle_len(I);
    for (int i = tid; i < LEN; i += bfly_plan::kBlock) ab2[i] = ab[i];
}

template <int I, int DIR>
__global__ __launch_bounds__(bfly_plan::kBlock) void k_c2c_mfma(const float2* __restrict__ in,
                                                              float2* __restrict__ out,
                                                              const float2* __restrict__ tw,
                                                              uint32_t batch) {
    constexpr int N = PlanOf<I>::N, T = PlanOf<I>::T, B = bfly_plan::kBlock / T, WAVES = T / 64;
    __shared__ float2 lds_all[B * lds_len(N)];
    __shared__ int s_exp[B * WAVES];
    __shared__ float2 ab2[128];
    const int tid = threadIdx.x;
    fill_ab2<I>(ab2, tw, tid);
    // (the transform's first LDS barrier orders these writes before the reads)
    const int bl = tid / T, t = tid % T;
    const uint64_t tr = (uint64_t)blockIdx.x * B + bl;
    const bool active = tr < batch;
    half8 ahi[2], alo[2];
    dft16_operands<DIR>(t & 63, ahi, alo);
    float2 v[16];
    load16_global<N>(v, in + tr * N, active, t >> 6, t & 63);
    mfma_transform<I, DIR>(v, lds_all + bl * lds_len(N), s_exp + bl * WAVES, tw, tw, ab2, ahi, alo, t,
                           out + tr * N, active);
}

// Streaming form: a resident grid walks the batch and loads each
// workgroup's next transforms while the current ones compute.
template <int I, int DIR>
__global__ __launch_bounds__(bfly_plan::kBlock) void k_c2c_mfma_stream(const float2* __restrict__ in,
                                                                     float2* __restrict__ out,
                                                                     const float2* __restrict__ tw,
                                                                     uint32_t batch) {
    constexpr int N = PlanOf<I>::N, T = PlanOf<I>::T, B = bfly_plan::kBlock / T, WAVES = T / 64;
    // the second pass's table copied once per workgroup, rows of 16 entries
    // at a pitch of 18: a lane group's 16-byte reads then hit distinct banks
    constexpr int PITCH1 = 18;
    __shared__ float2 lds_all[B * lds_len(N)];
    __shared__ int s_exp[B * WAVES];
    __shared__ float4 tw1s4[16 * PITCH1 / 2];
    __shared__ float2 ab2[128];
    float2* tw1s = reinterpret_cast<float2*>(tw1s4);
    const int tid = threadIdx.x;
    for (int i = tid; i < 256; i += bfly_plan::kBlock) tw1s[(i / 16) * PITCH1 + i % 16] = tw[i];
    fill_ab2<I>(ab2, tw, tid);
    // (the transform's first LDS barrier orders these writes before the reads)
    const int bl = tid / T, t0 = tid % T;
    const uint32_t groups = (uint32_t)((batch + B - 1) / B);
    half8 ahi[2], alo[2];
    dft16_operands<DIR>(t0 & 63, ahi, alo);
    float2 cur[16], nxt[16];
    uint32_t g = blockIdx.x;
    if (g < groups) {
        const uint64_t tr = (uint64_t)g * B + bl;
        load16_global<N>(cur, in + tr * N, tr < batch, t0 >> 6, t0 & 63);
    }
    for (; g < groups; g += gridDim.x) {
        int t = t0;  // opaque per iteration: keeps index-derived addresses in the loop
        asm volatile("" : "+v"(t));
        const uint32_t gn = g + gridDim.x;
        if (gn < groups) {
            const uint64_t trn = (uint64_t)gn * B + bl;
            load16_global<N>(nxt, in + trn * N, trn < batch, t >> 6, t & 63);
        }
        const uint64_t tr = (uint64_t)g * B + bl;
        mfma_transform<I, DIR, PITCH1>(cur, lds_all + bl * lds_len(N), s_exp + bl * WAVES, tw, tw1s, ab2, ahi,
                                       alo, t, out + tr * N, tr < batch);
#pragma unroll
        for (int i = 0; i < 16; i++) cur[i] = nxt[i];
    }
}


template <int I>
hipError_t launch_mfma(const float2* in, float2* out, const float2* tw, uint32_t batch, int sign,
                       bool stream, hipStream_t s) {
    constexpr int B = bfly_plan::kBlock / PlanOf<I>::T;
    const uint64_t groups = (batch + B - 1) / B;
    if (stream) {
        const void* f = sign < 0 ? reinterpret_cast<const void*>(&k_c2c_mfma_stream<I, -1>)
                                 : reinterpret_cast<const void*>(&k_c2c_mfma_stream<I, 1>);
        static bfly_kern::ResidentCache rc[2];
        const int c = bfly_kern::resident(rc[sign < 0 ? 0 : 1], f, bfly_plan::kBlock);
        if (c <= 0) return hipErrorInvalidValue;
        const dim3 grid((unsigned)(groups < (uint64_t)c ? groups : (uint64_t)c));
        if (sign < 0)
            hipLaunchKernelGGL((k_c2c_mfma_stream<I, -1>), grid, dim3(bfly_plan::kBlock), 0, s, in, out, tw,
                               batch);
        else
            hipLaunchKernelGGL((k_c2c_mfma_stream<I, 1>), grid, dim3(bfly_plan::kBlock), 0, s, in, out, tw,
                               batch);
        return hipGetLastError();
    }
    const dim3 grid((unsigned)groups);
    if (sign < 0)
        hipLaunchKernelGGL((k_c2c_mfma<I, -1>), grid, dim3(bfly_plan::kBlock), 0, s, in, out, tw, batch);
    else
        hipLaunchKernelGGL((k_c2c_mfma<I, 1>), grid, dim3(bfly_plan::kBlock), 0, s, in, out, tw, batch);
    return hipGetLastError();
}

constexpr int idx(int n) { return bfly_plan::plan_index((uint32_t)n); }

}  // namespace

namespace bfly_kern {

hipError_t c2c_mfma(const float* in, float* out, uint32_t n, uint32_t batch, int sign, bool stream,
                    const float* table, hipStream_t s) {
    const float2* i2 = reinterpret_cast<const float2*>(in);
    float2* o2 = reinterpret_cast<float2*>(out);
    const float2* tw = reinterpret_cast<const float2*>(table);
    switch (n) {
        case 1024: return launch_mfma<idx(1024)>(i2, o2, tw, batch, sign, stream, s);
        case 2048: return launch_mfma<idx(2048)>(i2, o2, tw, batch, sign, stream, s);
        case 4096: return launch_mfma<idx(4096)>(i2, o2, tw, batch, sign, stream, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace bfly_kern
