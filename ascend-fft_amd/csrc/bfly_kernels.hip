// bfly_kernels.hip -- gfx950 FFT kernels and their launchers.
//
//   k_c2c_local  C2C, n <= 4096: one HBM read + one HBM write per element;
//                256/T transforms per 256-thread workgroup (T = n/16).
//   k_c2c_stream* persistent streaming forms of the same transform.
//   k_c2c_single C2C, n = 8192 / 16384: one transform per 512 / 1024-thread
//                workgroup, still one HBM round trip.
//   k_r2c        R2C as an n/2-point complex FFT with the split post-twiddle
//                fused behind it; writes the n/2+1 bins at row stride n+2.
//   k_c2r        C2R: Hermitian pre-twiddle fused in front of an n/2-point
//                inverse complex FFT; writes reals scaled 1/n.
//   k_long_cols  four-step, kernel 1: N1-point column FFTs x W_N^{n2 k1},
//                written row-major by k1 into the workspace.
//   k_long_rows  four-step, kernel 2: N2-point row FFTs, natural-order
//                stores X[k1 + N1 k2].
//   k_copy       non-temporal four-deep float4 copy (HBM ceiling probe).
#include <hip/hip_runtime.h>

#include <utility>

#include "bfly_device.hpp"
#include "bfly_hip.h"
#include "bfly_kernels.hpp"
#include "bfly_plans.hpp"
#include "bfly_tables.hpp"
#include "bfly_units.hpp"

using namespace bfly_dev;
using bfly_plan::kPlans;

namespace {

// ------------------------------------------------------------------ C2C
template <int I, int DIR, int LX>
__global__ __launch_bounds__(bfly_plan::kBlock) void k_c2c_local(
    const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw,
    uint32_t batch) {
    constexpr int N = PlanOf<I>::N, B = bfly_plan::kBlock / PlanOf<I>::T;
    __shared__ float2 lds[B * lx_len<LX>(N)];
    c2c_unit<I, DIR, LX, bfly_plan::kBlock>(in, out, tw, batch, blockIdx.x, lds, threadIdx.x);
}

// C2C of a length one 256-thread workgroup cannot hold (n = 8192, 16384):
// one transform per workgroup of T = n/16 threads (512 / 1024), the whole
// transform LDS-resident (68 / 136 KiB), still one HBM round trip.
template <int I, int DIR>
__global__ __launch_bounds__(PlanOf<I>::T) void k_c2c_single(
    const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw,
    uint32_t batch) {
    __shared__ float2 lds[lds_len(PlanOf<I>::N)];
    c2c_unit<I, DIR, LX_PAD, PlanOf<I>::T>(in, out, tw, batch, blockIdx.x, lds, threadIdx.x);
}

// C2C, streaming form: a grid sized to the resident-workgroup capacity walks
// the batch; each thread keeps its twiddles in registers and issues the loads
// of its next transform before computing the current one, so HBM reads stay
// in flight through the LDS exchanges and the register butterflies.
template <int I, int DIR, int LX>
__global__ __launch_bounds__(bfly_plan::kBlock) void k_c2c_stream(
    const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw,
    uint32_t batch) {
    using PL = PlanOf<I>;
    constexpr int N = PL::N, T = PL::T, B = bfly_plan::kBlock / T;
    constexpr int R0 = PL::R0, RL = PL::RL, P = PL::P;
    constexpr int TWR = plan_twiddle_regs<I>(std::make_index_sequence<PL::NRAD>{});
    __shared__ float2 lds[B * lx_len<LX>(N)];
    const int tid = threadIdx.x;
    const int bl = tid / T, t0 = tid % T;
    const uint32_t groups = (uint32_t)((batch + B - 1) / B);
    float2 tr[TWR > 0 ? TWR : 1];
    load_plan_twiddles<I, T>(tr, tw, t0, std::make_index_sequence<PL::NRAD>{});
    float2 cur[P], nxt[P];
    uint32_t g = blockIdx.x;
    if (g < groups) {
        const uint64_t row = (uint64_t)g * B + bl;
        const float2* src = in + row * N;
#pragma unroll
        for (int m = 0; m < P / R0; m++)
#pragma unroll
            for (int r = 0; r < R0; r++)
                cur[m * R0 + r] = row < batch ? ld_io<(T >= 16)>(&src[t0 + m * T + r * (N / R0)]) : make_float2(0.f, 0.f);
    }
    for (; g < groups; g += gridDim.x) {
        int t = t0;  // opaque per iteration (see k_r2c_stream)
        asm volatile("" : "+v"(t));
        const uint32_t gn = g + gridDim.x;
        if (gn < groups) {
            const uint64_t row = (uint64_t)gn * B + bl;
            const float2* src = in + row * N;
#pragma unroll
            for (int m = 0; m < P / R0; m++)
#pragma unroll
                for (int r = 0; r < R0; r++)
                    nxt[m * R0 + r] = row < batch ? ld_io<(T >= 16)>(&src[t + m * T + r * (N / R0)]) : make_float2(0.f, 0.f);
        }
        run_plan_rt<I, T, DIR, LX>(cur, LX == LX_SWZ ? lds : lds + bl * lds_len(N),
                                   LX == LX_SWZ ? bl * N : 0, tr, t,
                                   std::make_index_sequence<PL::NRAD>{});
        const uint64_t row = (uint64_t)g * B + bl;
        if (row < batch) {
            float2* dst = out + row * N;
#pragma unroll
            for (int m = 0; m < P / RL; m++)
#pragma unroll
                for (int r = 0; r < RL; r++) st_io<(T >= 16)>(&dst[t + m * T + r * (N / RL)], cur[m * RL + r]);
        }
#pragma unroll
        for (int i = 0; i < P; i++) cur[i] = nxt[i];
    }
}

// Streaming form with two transforms in flight per workgroup: registers
// hold the transform being computed plus the next two being loaded.
template <int I, int DIR>
__global__ __launch_bounds__(bfly_plan::kBlock) void k_c2c_stream2(
    const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw,
    uint32_t batch) {
    using PL = PlanOf<I>;
    constexpr int N = PL::N, T = PL::T, B = bfly_plan::kBlock / T;
    constexpr int R0 = PL::R0, RL = PL::RL, P = PL::P;
    constexpr int TWR = plan_twiddle_regs<I>(std::make_index_sequence<PL::NRAD>{});
    __shared__ float2 lds[B * N];
    const int tid = threadIdx.x;
    const int bl = tid / T, t = tid % T;
    const uint32_t groups = (uint32_t)((batch + B - 1) / B);
    float2 tr[TWR > 0 ? TWR : 1];
    load_plan_twiddles<I, T>(tr, tw, t, std::make_index_sequence<PL::NRAD>{});
    float2 cur[P], n1[P], n2[P];
    auto fetch = [&](float2 (&v)[P], uint32_t grp) {
        const uint64_t row = (uint64_t)grp * B + bl;
        const float2* src = in + row * N;
#pragma unroll
        for (int m = 0; m < P / R0; m++)
#pragma unroll
            for (int r = 0; r < R0; r++)
                v[m * R0 + r] = row < batch ? ld_io<(T >= 16)>(&src[t + m * T + r * (N / R0)]) : make_float2(0.f, 0.f);
    };
    uint32_t g = blockIdx.x;
    if (g < groups) fetch(cur, g);
    if (g + gridDim.x < groups) fetch(n1, g + gridDim.x);
    for (; g < groups; g += gridDim.x) {
        const uint32_t g2 = g + 2 * gridDim.x;
        if (g2 < groups) fetch(n2, g2);
        run_plan_rt<I, T, DIR, LX_SWZ>(cur, lds, bl * N, tr, t, std::make_index_sequence<PL::NRAD>{});
        const uint64_t row = (uint64_t)g * B + bl;
        if (row < batch) {
            float2* dst = out + row * N;
#pragma unroll
            for (int m = 0; m < P / RL; m++)
#pragma unroll
                for (int r = 0; r < RL; r++) st_io<(T >= 16)>(&dst[t + m * T + r * (N / RL)], cur[m * RL + r]);
        }
#pragma unroll
        for (int i = 0; i < P; i++) {
            cur[i] = n1[i];
            n1[i] = n2[i];
        }
    }
}

// Streaming form with the two-level twiddle tables in LDS (no twiddle
// registers): more workgroups stay resident, so more transforms are in
// flight per CU.  `tw2g` points at [A (64) | B (N/64)] in the table.
// Threads per workgroup of the LDS-twiddle streaming kernel: 256, or one
// whole transform's T = n/16 above 4096 points.
template <int I>
constexpr int lt_wg() { return PlanOf<I>::T > bfly_plan::kBlock ? PlanOf<I>::T : bfly_plan::kBlock; }

template <int I, int DIR>
__global__ __launch_bounds__(lt_wg<I>()) void k_c2c_stream_lt(
    const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw,
    const float2* __restrict__ tw2g, uint32_t batch, float scale, uint64_t idist, uint64_t odist) {
    using PL = PlanOf<I>;
    constexpr int WG = lt_wg<I>();
    constexpr int N = PL::N, T = PL::T, B = WG / T;
    // tw2g: [A (64) | B (N/64)] contiguous up to 4096 points; above, the full
    // W_N^m table of the long layout, from which A = W_N^j and B = W_N^{64 j}
    constexpr bool kFullTable = N > bfly_plan::kMaxLocalN;
    constexpr int R0 = PL::R0, RL = PL::RL, P = PL::P;
    constexpr int NLO = N < TW2_LO ? N : TW2_LO;
    constexpr int NHI = N / NLO;
    __shared__ float2 lds[B * lds_len(N)];
    // the second pass's table (16 R1 entries, the first of the stage table,
    // stored [k R1 + r]), transposed into LDS as [r 16 + k]
    constexpr int R1 = kPlans[I].rad[PL::NRAD > 1 ? 1 : 0];
    constexpr int NT1 = PL::NRAD > 1 ? bfly_plan::kPoints * R1 : 1;
    __shared__ float2 ltw[NLO + NHI + NT1];
    const int tid = threadIdx.x;
    const int bl = tid / T, t0 = tid % T;
    for (int i = tid; i < NLO + NHI; i += WG) ltw[i] = tw2g[(kFullTable && i >= NLO) ? (i - NLO) * NLO : i];
    if constexpr (PL::NRAD > 1)
        for (int i = tid; i < NT1; i += WG)
            ltw[NLO + NHI + (i % R1) * bfly_plan::kPoints + i / R1] = tw[i];
    __syncthreads();
    const uint32_t groups = (uint32_t)((batch + B - 1) / B);
    float2 cur[P], nxt[P];
    uint32_t g = blockIdx.x;
    if (g < groups) {
        const uint64_t row = (uint64_t)g * B + bl;
        const float2* src = in + row * idist;
#pragma unroll
        for (int m = 0; m < P / R0; m++)
#pragma unroll
            for (int r = 0; r < R0; r++)
                cur[m * R0 + r] = row < batch ? ld_io<(T >= 16)>(&src[t0 + m * T + r * (N / R0)]) : make_float2(0.f, 0.f);
    }
    for (; g < groups; g += gridDim.x) {
        // opaque per-iteration thread index (see k_r2c_stream)
        int t = t0;
        asm volatile("" : "+v"(t));
        const uint32_t gn = g + gridDim.x;
        if (gn < groups) {
            const uint64_t row = (uint64_t)gn * B + bl;
            const float2* src = in + row * idist;
#pragma unroll
            for (int m = 0; m < P / R0; m++)
#pragma unroll
                for (int r = 0; r < R0; r++)
                    nxt[m * R0 + r] = row < batch ? ld_io<(T >= 16)>(&src[t + m * T + r * (N / R0)]) : make_float2(0.f, 0.f);
        }
        run_plan_lt<I, T, DIR, LX_PAD>(cur, lds + bl * lds_len(N), 0, ltw, ltw + NLO, ltw + NLO + NHI, t,
                                       std::make_index_sequence<PL::NRAD>{});
        const uint64_t row = (uint64_t)g * B + bl;
        if (scale != 1.0f) {  // uniform: the normalised entry point only
#pragma unroll
            for (int i = 0; i < P; i++) cur[i] = make_float2(cur[i].x * scale, cur[i].y * scale);
        }
        if (row < batch) {
            float2* dst = out + row * odist;
#pragma unroll
            for (int m = 0; m < P / RL; m++)
#pragma unroll
                for (int r = 0; r < RL; r++) st_io<(T >= 16)>(&dst[t + m * T + r * (N / RL)], cur[m * RL + r]);
        }
#pragma unroll
        for (int i = 0; i < P; i++) cur[i] = nxt[i];
    }
}

// One transform per workgroup (n = 8192 / 16384: 512 / 1024 threads, the
// whole transform in LDS, which leaves room for one workgroup per CU) on a
// resident grid that walks the batch.  With a single workgroup per CU nothing
// else hides the memory time, so the next transform's loads are issued before
// the current one's passes and must stay in flight across their LDS barriers.
// They are raw buffer loads through a per-row descriptor whose record count
// is zero when there is no next row (the loads then return 0 and touch no
// memory), and the stores always have a row: no branch joins the prefetch
// and the passes, which would make the compiler wait for the loads first.
// Same per-transform code as k_c2c_stream_lt, so the same bits.
template <int I, int DIR>
__global__ __launch_bounds__(PlanOf<I>::T) void k_c2c_single_pf(
    const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw,
    const float2* __restrict__ tw2g, uint32_t batch, float scale, uint64_t idist, uint64_t odist) {
    using PL = PlanOf<I>;
    constexpr int N = PL::N, T = PL::T;
    static_assert(N > bfly_plan::kMaxLocalN && PL::NRAD > 1, "one transform per workgroup");
    static_assert(T >= 256 && (T & (T - 1)) == 0, "the row relabelling permutes whole 256-thread blocks");
    constexpr int R0 = PL::R0, RL = PL::RL, P = PL::P;
    constexpr int NLO = TW2_LO, NHI = N / NLO;
    constexpr int R1 = kPlans[I].rad[1];
    constexpr int NT1 = bfly_plan::kPoints * R1;
    constexpr int kRowBytes = N * (int)sizeof(float2);
    __shared__ float2 lds[lds_len(N)];
    __shared__ float2 ltw[NLO + NHI + NT1];
    const int tid = threadIdx.x;
    // [W_N^j, j < 64 | W_N^{64 j}] from the full W_N table, and the second
    // pass's table transposed to [r 16 + k]
    for (int i = tid; i < NLO + NHI; i += T) ltw[i] = tw2g[i >= NLO ? (i - NLO) * NLO : i];
    for (int i = tid; i < NT1; i += T) ltw[NLO + NHI + (i % R1) * bfly_plan::kPoints + i / R1] = tw[i];
    __syncthreads();
    auto row_rsrc = [&](uint32_t row, bool valid) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(in + (uint64_t)row * idist), (short)0,
                                                 valid ? kRowBytes : 0, 0x00020000);
    };
    auto load_row = [&](float2 (&v)[P], __amdgpu_buffer_rsrc_t rs, int t) {
#pragma unroll
        for (int m = 0; m < P / R0; m++)
#pragma unroll
            for (int r = 0; r < R0; r++) {
                const auto w = __builtin_amdgcn_raw_buffer_load_b64(rs, (t + m * T + r * (N / R0)) * 8, 0, 2);
                v[m * R0 + r] = __builtin_bit_cast(float2, w);
            }
    };
    float2 cur[P], nxt[P];
    // Row g's transform runs with its threads relabelled t = (tid + 256 g)
    // mod T: the same work and bits per logical thread, but the workgroups
    // on neighbouring rows (rows n complex values apart, a power of two)
    // issue each load and store 2 KiB (8192) or a multiple of 2 KiB (16384)
    // apart inside the row, instead of all at the same offset.  The same
    // rows read and written with a 2 KiB pad between them ran 6 % faster;
    // relabelling by 256 threads per row measured 8192 x 2048 57.7 -> 54.5
    // us, 8192 x 4096 107.7 -> 103.2, 16384 x 1024 62.0 -> 58.1 (other
    // relabelling strides were slower or mixed).
    auto label = [&](uint32_t g) { return (int)((tid + 256u * g) & (uint32_t)(T - 1)); };
    // One step: prefetch the next row, transform and store the current one.
    // The first step is peeled off the loop, so the loop head is reached only
    // with the same memory operations pending (a step's loads, then its
    // stores) and waits for the loads alone, leaving the stores in flight.
    auto step = [&](uint32_t g) {
        int t = label(g);  // opaque per-step thread index (see k_r2c_stream)
        asm volatile("" : "+v"(t));
        const uint32_t gn = g + gridDim.x;
        int tn = label(gn);
        asm volatile("" : "+v"(tn));
        load_row(nxt, row_rsrc(gn < batch ? gn : g, gn < batch), tn);
        run_plan_lt<I, T, DIR, LX_PAD>(cur, lds, 0, ltw, ltw + NLO, ltw + NLO + NHI, t,
                                       std::make_index_sequence<PL::NRAD>{});
        if (scale != 1.0f) {  // uniform: the normalised entry point only
#pragma unroll
            for (int i = 0; i < P; i++) cur[i] = make_float2(cur[i].x * scale, cur[i].y * scale);
        }
        float2* dst = out + (uint64_t)g * odist;
#pragma unroll
        for (int m = 0; m < P / RL; m++)
#pragma unroll
            for (int r = 0; r < RL; r++) st_stream(&dst[t + m * T + r * (N / RL)], cur[m * RL + r]);
#pragma unroll
        for (int i = 0; i < P; i++) cur[i] = nxt[i];
    };
    uint32_t g = blockIdx.x;  // the grid never exceeds the batch
    load_row(cur, row_rsrc(g, true), label(g));
    step(g);
    for (g += gridDim.x; g < batch; g += gridDim.x) step(g);
}

// ------------------------------------------------------------------ R2C
// z[j] = x[2j] + i x[2j+1], Z = FFT_M(z), M = n/2; for k = 0..M:
//   E = (Z[k] + conj Z[M-k]) / 2,  O = (Z[k] - conj Z[M-k]) / (2i),
//   X[k] = E + W_n^k O            (Z[M] == Z[0]).
template <int I>
__global__ __launch_bounds__(bfly_plan::kBlock) void k_r2c(const float* __restrict__ in,
                                                           float* __restrict__ out,
                                                           const float2* __restrict__ tw,
                                                           const float2* __restrict__ wn,
                                                           uint32_t batch, uint32_t kout,
                                                           uint32_t ostride, int adj) {
    constexpr int M = PlanOf<I>::N, B = bfly_plan::kBlock / PlanOf<I>::T;
    __shared__ float2 lds[B * lds_len(M)];
    float2 v[bfly_plan::kPoints];
    r2c_load<I, bfly_plan::kBlock>(v, in, batch, blockIdx.x, threadIdx.x);
    r2c_finish<I, bfly_plan::kBlock>(v, out, RealTwGlobal<I, 0>{tw, wn}, batch, kout, ostride, blockIdx.x, lds,
                                     threadIdx.x, adj != 0);
}

// R2C, streaming form: a grid sized to the resident capacity walks the
// batch, loading the next group of rows while the current one computes.
template <int I>
__global__ __launch_bounds__(bfly_plan::kBlock) void k_r2c_stream(const float* __restrict__ in,
                                                                  float* __restrict__ out,
                                                                  const float2* __restrict__ tw,
                                                                  const float2* __restrict__ wn,
                                                                  uint32_t batch, uint32_t kout,
                                                                  uint32_t ostride, int adj) {
    constexpr int M = PlanOf<I>::N, B = bfly_plan::kBlock / PlanOf<I>::T, P = bfly_plan::kPoints;
    __shared__ float2 lds[B * lds_len(M)];
    const uint32_t groups = (uint32_t)((batch + B - 1) / B);
    const int tid = threadIdx.x;
    float2 cur[P], nxt[P];
    uint32_t g = blockIdx.x;
    if (g < groups) r2c_load<I, bfly_plan::kBlock>(cur, in, batch, g, tid);
    for (; g < groups; g += gridDim.x) {
        const uint32_t gn = g + gridDim.x;
        // an opaque per-iteration copy of the thread index keeps the compiler
        // from hoisting every index-derived address out of the loop (their
        // live ranges would double the register footprint)
        int t2 = tid;
        asm volatile("" : "+v"(t2));
        if (gn < groups) r2c_load<I, bfly_plan::kBlock>(nxt, in, batch, gn, t2);
        r2c_finish<I, bfly_plan::kBlock>(cur, out, RealTwGlobal<I, 0>{tw, wn}, batch, kout, ostride, g, lds, t2,
                                         adj != 0);
        __syncthreads();  // the post-twiddle reads of `lds` end before the next FFT writes it
#pragma unroll
        for (int i = 0; i < P; i++) cur[i] = nxt[i];
    }
}

// ------------------------------------------------------------------ C2R
// Z'[k] = (X[k] + conj X[M-k]) + i W_n^{-k} (X[k] - conj X[M-k]), k < M,
// with Im X[0] = Im X[M] = 0; z = FFT^+_M(Z') / n; x[2j] = Re z, x[2j+1] = Im z.
// Reads bins k < kin of each row at a row stride of istride floats; bins at
// or above kin are zero (the full transform: kin = M+1, istride = 2M+2).
template <int I>
__global__ __launch_bounds__(bfly_plan::kBlock) void k_c2r(const float* __restrict__ in,
                                                           float* __restrict__ out,
                                                           const float2* __restrict__ tw,
                                                           const float2* __restrict__ wn,
                                                           uint32_t batch, float scale, uint32_t kin,
                                                           uint32_t istride, int adj) {
    constexpr int M = PlanOf<I>::N, B = bfly_plan::kBlock / PlanOf<I>::T;
    __shared__ float2 lds[B * lds_len(M)];
    float2 q[bfly_plan::kPoints + 1];
    c2r_load<I, bfly_plan::kBlock>(q, in, batch, kin, istride, blockIdx.x, threadIdx.x);
    c2r_finish<I, bfly_plan::kBlock>(q, out, RealTwGlobal<I, 0>{tw, wn}, batch, scale, blockIdx.x, lds, threadIdx.x,
                                     adj != 0);
}

// C2R, streaming form (see k_r2c_stream).
template <int I>
__global__ __launch_bounds__(bfly_plan::kBlock) void k_c2r_stream(const float* __restrict__ in,
                                                                  float* __restrict__ out,
                                                                  const float2* __restrict__ tw,
                                                                  const float2* __restrict__ wn,
                                                                  uint32_t batch, float scale, uint32_t kin,
                                                                  uint32_t istride, int adj) {
    constexpr int M = PlanOf<I>::N, B = bfly_plan::kBlock / PlanOf<I>::T, P = bfly_plan::kPoints;
    __shared__ float2 lds[B * lds_len(M)];
    const uint32_t groups = (uint32_t)((batch + B - 1) / B);
    const int tid = threadIdx.x;
    float2 cur[P + 1], nxt[P + 1];
    uint32_t g = blockIdx.x;
    if (g < groups) c2r_load<I, bfly_plan::kBlock>(cur, in, batch, kin, istride, g, tid);
    for (; g < groups; g += gridDim.x) {
        const uint32_t gn = g + gridDim.x;
        int t2 = tid;  // opaque per iteration (see k_r2c_stream)
        asm volatile("" : "+v"(t2));
        if (gn < groups) c2r_load<I, bfly_plan::kBlock>(nxt, in, batch, kin, istride, gn, t2);
        if constexpr (PlanOf<I>::T >= 16) {
            // the same steps as c2r_finish through its two halves, which the
            // compiler schedules differently: 512 x 32768 32.3 -> 30.2 us
            // (and 256 x 65536 30.2 -> 34 us, hence only from 16 threads)
            c2r_stage<I, bfly_plan::kBlock>(cur, lds, t2);
            __syncthreads();
            c2r_compute<I, bfly_plan::kBlock>(out, RealTwGlobal<I, 0>{tw, wn}, batch, scale, g, lds, t2, adj != 0);
        } else {
            c2r_finish<I, bfly_plan::kBlock>(cur, out, RealTwGlobal<I, 0>{tw, wn}, batch, scale, g, lds, t2,
                                             adj != 0);
        }
        __syncthreads();  // the last exchange's reads end before the next staging writes
#pragma unroll
        for (int i = 0; i <= P; i++) cur[i] = nxt[i];
    }
}

// R2C / C2R, streaming form with LDS twiddles (the FFT's two-level tables
// and the split twiddle W_n^k as a product of two small tables), launched
// with the grid of bfly_plan::stream_grid.
template <int I>
__global__ __launch_bounds__(bfly_plan::kBlock) void k_r2c_stream_lt(const float* __restrict__ in,
                                                                     float* __restrict__ out,
                                                                     const float2* __restrict__ tw,
                                                                     const float2* __restrict__ wn,
                                                                     uint32_t batch, uint32_t kout,
                                                                     uint32_t ostride, int adj) {
    constexpr int M = PlanOf<I>::N, B = bfly_plan::kBlock / PlanOf<I>::T, P = bfly_plan::kPoints;
    __shared__ float2 lds[B * lds_len(M)];
    __shared__ float2 ltw[real_lds_tw_len<I>()];
    const uint32_t groups = (uint32_t)((batch + B - 1) / B);
    const int tid = threadIdx.x;
    const RealTwLds<I> tws = real_lds_tw<I, bfly_plan::kBlock>(ltw, tw, wn, tid);
    // the FFT's first exchange barrier orders the table writes before use
    float2 cur[P], nxt[P];
    uint32_t g = blockIdx.x;
    if (g < groups) r2c_load<I, bfly_plan::kBlock>(cur, in, batch, g, tid);
    for (; g < groups; g += gridDim.x) {
        const uint32_t gn = g + gridDim.x;
        int t2 = tid;  // opaque per iteration (see k_r2c_stream)
        asm volatile("" : "+v"(t2));
        if (gn < groups) r2c_load<I, bfly_plan::kBlock>(nxt, in, batch, gn, t2);
        r2c_finish<I, bfly_plan::kBlock, 0>(cur, out, tws, batch, kout, ostride, g, lds, t2, adj != 0);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < P; i++) cur[i] = nxt[i];
    }
}

template <int I>
__global__ __launch_bounds__(bfly_plan::kBlock) __attribute__((amdgpu_waves_per_eu(PlanOf<I>::N >= 4096 ? 4 : 1))) void
k_c2r_stream_lt(const float* __restrict__ in,
                                                                     float* __restrict__ out,
                                                                     const float2* __restrict__ tw,
                                                                     const float2* __restrict__ wn,
                                                                     uint32_t batch, float scale,
                                                                     uint32_t kin, uint32_t istride, int adj) {
    constexpr int M = PlanOf<I>::N, B = bfly_plan::kBlock / PlanOf<I>::T, P = bfly_plan::kPoints;
    __shared__ float2 lds[B * lds_len(M)];
    __shared__ float2 ltw[real_lds_tw_len<I>()];
    const uint32_t groups = (uint32_t)((batch + B - 1) / B);
    const int tid = threadIdx.x;
    const RealTwLds<I> tws = real_lds_tw<I, bfly_plan::kBlock>(ltw, tw, wn, tid);
    __syncthreads();  // the pre-twiddle reads the tables before the FFT's first barrier
    float2 cur[P + 1], nxt[P + 1];
    uint32_t g = blockIdx.x;
    if (g < groups) c2r_load<I, bfly_plan::kBlock>(cur, in, batch, kin, istride, g, tid);
    for (; g < groups; g += gridDim.x) {
        const uint32_t gn = g + gridDim.x;
        int t2 = tid;  // opaque per iteration (see k_r2c_stream)
        asm volatile("" : "+v"(t2));
        if constexpr (M >= 4096) {
            // Stage the current rows before the next rows' loads are issued:
            // the wait that the staging needs (for the current rows' loads,
            // issued a step earlier) then no longer holds back the prefetch,
            // which stays in flight through the pre-twiddle and the FFT; with
            // the kernel held to 128 VGPRs (four workgroups per CU) n = 8192
            // x 2048 measured 30.7 -> 27.5-28.5 us.  At n <= 4096 it measured
            // flat (2048, 4096) or slower (512: 29.3 -> 30.1 us).
            c2r_stage<I, bfly_plan::kBlock, 0>(cur, lds, t2);
            __syncthreads();
            if (gn < groups) c2r_load<I, bfly_plan::kBlock>(nxt, in, batch, kin, istride, gn, t2);
            c2r_compute<I, bfly_plan::kBlock, 0>(out, tws, batch, scale, g, lds, t2, adj != 0);
        } else {
            if (gn < groups) c2r_load<I, bfly_plan::kBlock>(nxt, in, batch, kin, istride, gn, t2);
            c2r_finish<I, bfly_plan::kBlock, 0>(cur, out, tws, batch, scale, g, lds, t2, adj != 0);
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i <= P; i++) cur[i] = nxt[i];
    }
}

// ------------------------------------------------------------- long C2C
// n = N2 n1 + n2, k = k1 + N1 k2:
//   Y[k1][n2] = W_N^{n2 k1} * FFT_N1 over n1 of x[N2 n1 + n2]   (k_long_cols)
//   X[k1 + N1 k2] = FFT_N2 over n2 of Y[k1][n2]                 (k_long_rows)
// Both kernels put the C columns (rows) of one workgroup on consecutive
// lanes, so every global access is a run of C consecutive complex values.
// Per-column LDS regions are spaced lds_len + 1 apart to spread banks.  The
// local plans here have two passes; the second pass's table is copied into
// LDS once per workgroup (stored [r 16 + k], as stockham_lt reads it: the
// same values as the global table, so the same bits).
template <int I1, int N2, int DIR, int WG = bfly_plan::kBlock>
__global__ __launch_bounds__(WG) void k_long_cols(
    const float2* __restrict__ in, float2* __restrict__ ws, const float2* __restrict__ tw1,
    const float2* __restrict__ wN, uint64_t idist, int temporal) {
    using PL = PlanOf<I1>;
    constexpr int N1 = PL::N, T = PL::T, C = WG / T;
    constexpr int R0 = PL::R0, RL = PL::RL, P = PL::P;
    constexpr int STRIDE = lds_len(N1) + 1;
    constexpr int GROUPS = N2 / C;
    static_assert(N2 % C == 0, "column groups must tile N2");
    // W_N^m = A[m mod 256] B[m / 256]: A = W_N^j and B = W_N^{256 j} are
    // entries of the full table, gathered from LDS instead of scattered
    // single-element reads of the 8N-byte table
    constexpr int NA = 256, NB = (int)(((uint64_t)N1 * N2) / NA);
    static_assert(PL::NRAD == 2, "two-pass column plans");
    constexpr int R1 = kPlans[I1].rad[1];
    __shared__ float2 lds[C * STRIDE];
    __shared__ float2 tab[NA + NB];
    __shared__ float2 t1s[bfly_plan::kPoints * R1];
    const int tid = threadIdx.x;
    for (int i = tid; i < NA + NB; i += WG) tab[i] = i < NA ? wN[i] : wN[(i - NA) * NA];
    for (int i = tid; i < bfly_plan::kPoints * R1; i += WG)
        t1s[(i % R1) * bfly_plan::kPoints + i / R1] = tw1[i];
    const int c = tid % C, t = tid / C;
    const uint64_t b = blockIdx.x / GROUPS;
    const int n2 = (blockIdx.x % GROUPS) * C + c;
    constexpr uint64_t N = (uint64_t)N1 * N2;
    const float2* src = in + b * idist + n2;
    float2 v[P];
#pragma unroll
    for (int m = 0; m < P / R0; m++)
#pragma unroll
        for (int r = 0; r < R0; r++)
            v[m * R0 + r] = ld_stream(&src[(uint64_t)N2 * (t + m * T + r * (N1 / R0))]);
    // the first LDS exchange of the FFT orders the table writes before any read
    run_plan_lt<I1, T, DIR, LX_PAD>(v, lds + c * STRIDE, 0, t1s, t1s, t1s, t, std::make_index_sequence<2>{});
    float2* dst = ws + b * N + n2;
    // The intermediate is read back by the row kernel right after: from
    // 2^24 points per launch up it is written (and read) with ordinary,
    // temporal accesses (65536 x 256: 98.5 -> 93.3 us, 32768 x 512: 94.9 ->
    // 87.1), below that with the streaming hint (65536 x 64: 28.2 against
    // 30.6 us temporal).  `temporal` is uniform.
    auto store = [&](auto tmp) {
#pragma unroll
        for (int m = 0; m < P / RL; m++)
#pragma unroll
            for (int r = 0; r < RL; r++) {
                const int k1 = t + m * T + r * (N1 / RL);
                const int e = (n2 * k1) & (int)(N - 1);
                const float2 a = tab[e & (NA - 1)], w = tab[NA + (e >> 8)];
                const float2 wv = make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
                const float2 y = cmul_tw<DIR>(v[m * RL + r], wv);
                if constexpr (decltype(tmp)::value) dst[(uint64_t)N2 * k1] = y;
                else st_stream(&dst[(uint64_t)N2 * k1], y);
            }
    };
    if (temporal) store(std::true_type{});
    else store(std::false_type{});
}

// Streaming form of k_long_cols (512-thread workgroups): a resident grid
// walks the units (C-column groups of every transform, in the same order as
// the one-shot grid's workgroups) with the next unit's loads in flight
// through the current unit's passes.  The loads go through a per-unit raw
// buffer descriptor whose record count is zero past the last unit (no
// branch around the prefetch; see k_c2c_single_pf), and the tables are
// staged in LDS once per workgroup instead of once per unit.  Same
// per-column arithmetic as k_long_cols, so the same bits.
template <int I1, int N2, int DIR, bool TEMPORAL>
__global__ __launch_bounds__(512) void k_long_cols_stream(
    const float2* __restrict__ in, float2* __restrict__ ws, const float2* __restrict__ tw1,
    const float2* __restrict__ wN, uint64_t idist, uint32_t units) {
    using PL = PlanOf<I1>;
    constexpr int WG = 512, N1 = PL::N, T = PL::T, C = WG / T;
    constexpr int R0 = PL::R0, RL = PL::RL, P = PL::P;
    constexpr int STRIDE = lds_len(N1) + 1;
    constexpr int GROUPS = N2 / C;
    static_assert(N2 % C == 0 && PL::NRAD == 2, "two-pass column plans tiling N2");
    constexpr int NA = 256, NB = (int)(((uint64_t)N1 * N2) / NA);
    constexpr int R1 = kPlans[I1].rad[1];
    constexpr uint64_t N = (uint64_t)N1 * N2;
    // bytes from a unit's first element to one past its last
    constexpr uint64_t kSpan = ((uint64_t)(N1 - 1) * N2 + C) * sizeof(float2);
    static_assert(kSpan < (1ull << 31), "a unit's span fits a buffer descriptor");
    __shared__ float2 lds[C * STRIDE];
    __shared__ float2 tab[NA + NB];
    __shared__ float2 t1s[bfly_plan::kPoints * R1];
    const int tid = threadIdx.x;
    for (int i = tid; i < NA + NB; i += WG) tab[i] = i < NA ? wN[i] : wN[(i - NA) * NA];
    for (int i = tid; i < bfly_plan::kPoints * R1; i += WG)
        t1s[(i % R1) * bfly_plan::kPoints + i / R1] = tw1[i];
    // the first LDS exchange of the first unit orders the table writes before any read
    const int c = tid % C, t0 = tid / C;
    auto unit_rsrc = [&](uint32_t u, bool valid) {
        const uint64_t b = u / GROUPS;
        const int n2 = (int)(u % GROUPS) * C;
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(in + b * idist + n2), (short)0,
                                                 valid ? (int)kSpan : 0, 0x00020000);
    };
    auto load_unit = [&](float2 (&v)[P], __amdgpu_buffer_rsrc_t rs, int t) {
#pragma unroll
        for (int m = 0; m < P / R0; m++)
#pragma unroll
            for (int r = 0; r < R0; r++) {
                const int row = t + m * T + r * (N1 / R0);
                v[m * R0 + r] = __builtin_bit_cast(
                    float2, __builtin_amdgcn_raw_buffer_load_b64(rs, (row * N2 + c) * (int)sizeof(float2), 0, 2));
            }
    };
    float2 cur[P], nxt[P];
    auto step = [&](uint32_t u) {
        int t = t0;  // opaque per-step thread index (see k_r2c_stream)
        asm volatile("" : "+v"(t));
        const uint32_t un = u + gridDim.x;
        load_unit(nxt, unit_rsrc(un < units ? un : u, un < units), t);
        // the exchange's trailing barrier ends every LDS read before the
        // next unit's first writes
        run_plan_lt<I1, T, DIR, LX_PAD>(cur, lds + c * STRIDE, 0, t1s, t1s, t1s, t, std::make_index_sequence<2>{});
        const uint64_t b = u / GROUPS;
        const int n2 = (int)(u % GROUPS) * C + c;
        float2* dst = ws + b * N + n2;
#pragma unroll
        for (int m = 0; m < P / RL; m++)
#pragma unroll
            for (int r = 0; r < RL; r++) {
                const int k1 = t + m * T + r * (N1 / RL);
                const int e = (n2 * k1) & (int)(N - 1);
                const float2 a = tab[e & (NA - 1)], w = tab[NA + (e >> 8)];
                const float2 wv = make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
                const float2 y = cmul_tw<DIR>(cur[m * RL + r], wv);
                if constexpr (TEMPORAL) dst[(uint64_t)N2 * k1] = y;
                else st_stream(&dst[(uint64_t)N2 * k1], y);
            }
#pragma unroll
        for (int i = 0; i < P; i++) cur[i] = nxt[i];
    };
    uint32_t u = blockIdx.x;  // the grid never exceeds the unit count
    load_unit(cur, unit_rsrc(u, true), t0);
    step(u);
    for (u += gridDim.x; u < units; u += gridDim.x) step(u);
}

template <int I2, int N1, int DIR, int WG = bfly_plan::kBlock>
__global__ __launch_bounds__(WG) void k_long_rows(
    const float2* __restrict__ ws, float2* __restrict__ out, const float2* __restrict__ tw2, float scale,
    uint64_t odist, int temporal) {
    using PL = PlanOf<I2>;
    constexpr int N2 = PL::N, T = PL::T, C = WG / T;
    constexpr int R0 = PL::R0, RL = PL::RL, P = PL::P;
    constexpr int STRIDE = lds_len(N2) + 1;
    constexpr int GROUPS = N1 / C;
    static_assert(N1 % C == 0, "row groups must tile N1");
    static_assert(PL::NRAD == 2, "two-pass row plans");
    constexpr int R1 = kPlans[I2].rad[1];
    __shared__ float2 lds[C * STRIDE];
    __shared__ float2 t1s[bfly_plan::kPoints * R1];
    const int tid = threadIdx.x;
    for (int i = tid; i < bfly_plan::kPoints * R1; i += WG)
        t1s[(i % R1) * bfly_plan::kPoints + i / R1] = tw2[i];
    const uint64_t b = blockIdx.x / GROUPS;
    const int k1_0 = (blockIdx.x % GROUPS) * C;
    constexpr uint64_t N = (uint64_t)N1 * N2;
    // Stage the C contiguous rows through LDS (row-linear, coalesced); the
    // intermediate's access hint follows k_long_cols (`temporal` is uniform).
    const float2* src = ws + b * N + (uint64_t)k1_0 * N2;
    auto stage = [&](auto tmp) {
#pragma unroll
        for (int i = 0; i < C * N2 / WG; i++) {
            const int idx = tid + i * WG;
            const int row = idx / N2, col = idx % N2;
            if constexpr (decltype(tmp)::value) lds[row * STRIDE + lds_idx(col)] = src[idx];
            else lds[row * STRIDE + lds_idx(col)] = ld_stream(&src[idx]);
        }
    };
    if (temporal) stage(std::true_type{});
    else stage(std::false_type{});
    __syncthreads();
    const int c = tid % C, t = tid / C;
    float2* L = lds + c * STRIDE;
    float2 v[P];
#pragma unroll
    for (int m = 0; m < P / R0; m++)
#pragma unroll
        for (int r = 0; r < R0; r++) v[m * R0 + r] = L[lds_idx(t + m * T + r * (N2 / R0))];
    __syncthreads();
    run_plan_lt<I2, T, DIR, LX_PAD>(v, L, 0, t1s, t1s, t1s, t, std::make_index_sequence<2>{});
    float2* dst = out + b * odist + k1_0 + c;
    if (scale != 1.0f) {  // uniform: the normalised entry point only
#pragma unroll
        for (int i = 0; i < P; i++) v[i] = make_float2(v[i].x * scale, v[i].y * scale);
    }
#pragma unroll
    for (int m = 0; m < P / RL; m++)
#pragma unroll
        for (int r = 0; r < RL; r++) {
            const int k2 = t + m * T + r * (N2 / RL);
            st_stream(&dst[(uint64_t)N1 * k2], v[m * RL + r]);
        }
}

// ----------------------------------------------------------------- copy
// The HBM ceiling probe: a streaming copy in which every thread issues four
// non-temporal 16-byte loads (consecutive lanes, consecutive addresses)
// before their four stores, on a grid of up to 8 workgroups per CU.  This is
// the fastest copy form measured here (~6.2 TB/s, against ~5.2 TB/s for a
// plain one-float4-per-thread grid-stride copy).
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src,
                                              float4* __restrict__ dst, size_t n4) {
    typedef float v4f __attribute__((ext_vector_type(4)));
    constexpr int D = 4;
    const v4f* s = reinterpret_cast<const v4f*>(src);
    v4f* d = reinterpret_cast<v4f*>(dst);
    const size_t step = (size_t)gridDim.x * 256 * D;
    size_t base = (size_t)blockIdx.x * 256 * D + threadIdx.x;
    for (; base + 256 * (D - 1) < n4; base += step) {
        v4f v[D];
#pragma unroll
        for (int k = 0; k < D; k++) v[k] = __builtin_nontemporal_load(s + base + 256 * k);
#pragma unroll
        for (int k = 0; k < D; k++) __builtin_nontemporal_store(v[k], d + base + 256 * k);
    }
    // the last, partial step of this thread (nothing lies beyond it)
#pragma unroll
    for (int k = 0; k < D; k++)
        if (base + 256 * k < n4) __builtin_nontemporal_store(__builtin_nontemporal_load(s + base + 256 * k),
                                                             d + base + 256 * k);
}

// ------------------------------------------------------------ dispatch

template <int I, int DIR, int LX>
hipError_t launch_c2c_dir(const float2* in, float2* out, const float2* tw, uint32_t batch,
                          bool stream, hipStream_t s) {
    constexpr int T = PlanOf<I>::T, B = bfly_plan::kBlock / T;
    const uint64_t groups = (batch + B - 1) / B;
    if (stream) {
        static bfly_kern::ResidentCache rc;
        const int cap = bfly_kern::resident(rc, reinterpret_cast<const void*>(&k_c2c_stream<I, DIR, LX>), bfly_plan::kBlock);
        if (cap <= 0) return hipErrorInvalidValue;
        const dim3 grid((unsigned)(groups < (uint64_t)cap ? groups : (uint64_t)cap));
        hipLaunchKernelGGL((k_c2c_stream<I, DIR, LX>), grid, dim3(bfly_plan::kBlock), 0, s, in, out, tw, batch);
    } else {
        hipLaunchKernelGGL((k_c2c_local<I, DIR, LX>), dim3((unsigned)groups), dim3(bfly_plan::kBlock), 0, s,
                           in, out, tw, batch);
    }
    return hipGetLastError();
}

template <int I, int DIR>
hipError_t launch_c2c_lt(const float2* in, float2* out, const float2* tw, const float2* tw2g,
                         uint32_t batch, hipStream_t s, float scale = 1.0f, uint64_t idist = 0,
                         uint64_t odist = 0) {
    constexpr int WG = lt_wg<I>(), B = WG / PlanOf<I>::T;
    // row distances in complex values; 0 = contiguous rows of n
    if (!idist) idist = PlanOf<I>::N;
    if (!odist) odist = PlanOf<I>::N;
    const uint64_t groups = (batch + B - 1) / B;
    if constexpr (PlanOf<I>::N > bfly_plan::kMaxLocalN) {
        // one transform per workgroup: the prefetching kernel on a resident grid
        static bfly_kern::ResidentCache rp;
        const int cap = bfly_kern::resident(rp, reinterpret_cast<const void*>(&k_c2c_single_pf<I, DIR>), WG);
        if (cap <= 0) return hipErrorInvalidValue;
        const dim3 grid((unsigned)(groups < (uint64_t)cap ? groups : (uint64_t)cap));
        hipLaunchKernelGGL((k_c2c_single_pf<I, DIR>), grid, dim3(WG), 0, s, in, out, tw, tw2g, batch, scale, idist,
                           odist);
        return hipGetLastError();
    }
    static bfly_kern::ResidentCache rc;
    const int cap = bfly_kern::resident(rc, reinterpret_cast<const void*>(&k_c2c_stream_lt<I, DIR>), WG);
    if (cap <= 0) return hipErrorInvalidValue;
    const dim3 grid((unsigned)bfly_plan::stream_grid(groups, (uint64_t)cap));
    hipLaunchKernelGGL((k_c2c_stream_lt<I, DIR>), grid, dim3(WG), 0, s, in, out, tw, tw2g, batch, scale, idist,
                       odist);
    return hipGetLastError();
}

template <int I>
hipError_t launch_c2c_local(const float2* in, float2* out, const float2* tw, uint32_t batch,
                            int sign, int variant, hipStream_t s, float scale) {
    if (variant == BFLY_VARIANT_STREAM2) {
        constexpr int T = PlanOf<I>::T, B = bfly_plan::kBlock / T;
        const uint64_t groups = (batch + B - 1) / B;
        const void* f = sign < 0 ? reinterpret_cast<const void*>(&k_c2c_stream2<I, -1>)
                                 : reinterpret_cast<const void*>(&k_c2c_stream2<I, 1>);
        static bfly_kern::ResidentCache rc[2];
        const int c = bfly_kern::resident(rc[sign < 0 ? 0 : 1], f, bfly_plan::kBlock);
        if (c <= 0) return hipErrorInvalidValue;
        const dim3 grid((unsigned)(groups < (uint64_t)c ? groups : (uint64_t)c));
        if (sign < 0)
            hipLaunchKernelGGL((k_c2c_stream2<I, -1>), grid, dim3(bfly_plan::kBlock), 0, s, in, out, tw, batch);
        else
            hipLaunchKernelGGL((k_c2c_stream2<I, 1>), grid, dim3(bfly_plan::kBlock), 0, s, in, out, tw, batch);
        return hipGetLastError();
    }
    if (variant == BFLY_VARIANT_STREAM_LT) {
        const float2* tw2g = tw + bfly_host::table_layout(0, PlanOf<I>::N).extra;
        return sign < 0 ? launch_c2c_lt<I, -1>(in, out, tw, tw2g, batch, s, scale)
                        : launch_c2c_lt<I, 1>(in, out, tw, tw2g, batch, s, scale);
    }
    const bool stream = variant == BFLY_VARIANT_STREAM || variant == BFLY_VARIANT_STREAM_SWZ;
    const bool swz = variant == BFLY_VARIANT_PLAIN_SWZ || variant == BFLY_VARIANT_STREAM_SWZ;
    if (sign < 0)
        return swz ? launch_c2c_dir<I, -1, LX_SWZ>(in, out, tw, batch, stream, s)
                   : launch_c2c_dir<I, -1, LX_PAD>(in, out, tw, batch, stream, s);
    return swz ? launch_c2c_dir<I, 1, LX_SWZ>(in, out, tw, batch, stream, s)
               : launch_c2c_dir<I, 1, LX_PAD>(in, out, tw, batch, stream, s);
}

template <int I>
hipError_t launch_real(int kind, const float* in, float* out, const float2* tw, const float2* wn,
                       uint32_t batch, uint32_t bins, uint32_t stride, int variant, int adj, hipStream_t s) {
    constexpr int T = PlanOf<I>::T, B = bfly_plan::kBlock / T;
    const uint64_t groups = (batch + B - 1) / B;
    // C2R scale: 1/n, or 1/2 for the adjoint (its k = 0 term is doubled in the kernel)
    const float scale = adj ? 0.5f : 1.0f / (float)(2 * PlanOf<I>::N);
    if constexpr (PlanOf<I>::N >= 128) {
        if (variant == BFLY_VARIANT_STREAM_LT) {
            const void* f = kind == 1 ? reinterpret_cast<const void*>(&k_r2c_stream_lt<I>)
                                      : reinterpret_cast<const void*>(&k_c2r_stream_lt<I>);
            static bfly_kern::ResidentCache rc[2];
            const int c = bfly_kern::resident(rc[kind == 1 ? 0 : 1], f, bfly_plan::kBlock);
            if (c <= 0) return hipErrorInvalidValue;
            const dim3 grid((unsigned)bfly_plan::stream_grid(groups, (uint64_t)c));
            if (kind == 1)
                hipLaunchKernelGGL((k_r2c_stream_lt<I>), grid, dim3(bfly_plan::kBlock), 0, s, in, out, tw, wn,
                                   batch, bins, stride, adj);
            else
                hipLaunchKernelGGL((k_c2r_stream_lt<I>), grid, dim3(bfly_plan::kBlock), 0, s, in, out, tw, wn,
                                   batch, scale, bins, stride, adj);
            return hipGetLastError();
        }
    }
    if (variant == BFLY_VARIANT_STREAM_LT) variant = BFLY_VARIANT_STREAM;
    if (variant == BFLY_VARIANT_STREAM) {
        const void* f = kind == 1 ? reinterpret_cast<const void*>(&k_r2c_stream<I>)
                                  : reinterpret_cast<const void*>(&k_c2r_stream<I>);
        static bfly_kern::ResidentCache rc[2];
        const int c = bfly_kern::resident(rc[kind == 1 ? 0 : 1], f, bfly_plan::kBlock);
        if (c <= 0) return hipErrorInvalidValue;
        const dim3 grid((unsigned)(groups < (uint64_t)c ? groups : (uint64_t)c));
        if (kind == 1)
            hipLaunchKernelGGL((k_r2c_stream<I>), grid, dim3(bfly_plan::kBlock), 0, s, in, out, tw, wn,
                               batch, bins, stride, adj);
        else
            hipLaunchKernelGGL((k_c2r_stream<I>), grid, dim3(bfly_plan::kBlock), 0, s, in, out, tw, wn,
                               batch, scale, bins, stride, adj);
        return hipGetLastError();
    }
    const dim3 grid((unsigned)groups);
    if (kind == 1)
        hipLaunchKernelGGL((k_r2c<I>), grid, dim3(bfly_plan::kBlock), 0, s, in, out, tw, wn, batch,
                           bins, stride, adj);
    else
        hipLaunchKernelGGL((k_c2r<I>), grid, dim3(bfly_plan::kBlock), 0, s, in, out, tw, wn, batch,
                           scale, bins, stride, adj);
    return hipGetLastError();
}

template <int I>
hipError_t launch_c2c_single(const float2* in, float2* out, const float2* tw, uint32_t batch,
                             int sign, hipStream_t s) {
    constexpr int T = PlanOf<I>::T;
    if (sign < 0)
        hipLaunchKernelGGL((k_c2c_single<I, -1>), dim3(batch), dim3(T), 0, s, in, out, tw, batch);
    else
        hipLaunchKernelGGL((k_c2c_single<I, 1>), dim3(batch), dim3(T), 0, s, in, out, tw, batch);
    return hipGetLastError();
}

template <int I1, int I2, int DIR>
hipError_t launch_long_dir(const float2* in, float2* out, float2* ws, const float2* tw1,
                           const float2* tw2, const float2* wN, uint32_t batch, hipStream_t s,
                           hipEvent_t mid, float scale, uint64_t idist, uint64_t odist) {
    constexpr int N1 = kPlans[I1].n, N2 = kPlans[I2].n;
    // row distances in complex values; 0 = contiguous rows of n
    if (!idist) idist = (uint64_t)N1 * N2;
    if (!odist) odist = (uint64_t)N1 * N2;
    // Column kernel in 512-thread workgroups when that still fills every
    // resident slot: twice the columns per workgroup, 256-byte runs (65536 x
    // 256: 98.8-100.9 vs 103.1-104.5 us; 65536 x 64: 25.7 vs 31.7; 32768 x
    // 512: 95.1-96.3 vs 98.3-98.8).  With fewer column groups the 256-thread
    // form's wider grid wins (16384 x 64: 11.5 vs 12.6 us).  The row kernel
    // measured slower in 512-thread workgroups.
    constexpr int C1w = 512 / PlanOf<I1>::T, C1 = bfly_plan::kBlock / PlanOf<I1>::T;
    constexpr int C2 = bfly_plan::kBlock / PlanOf<I2>::T;
    const int temporal = (uint64_t)batch * N1 * N2 >= (1ull << 24) ? 1 : 0;
    static bfly_kern::ResidentCache rc;
    const int fit = bfly_kern::resident(rc, reinterpret_cast<const void*>(&k_long_cols<I1, N2, DIR, 512>), 512);
    const uint64_t units = (uint64_t)batch * (N2 / C1w);
    // 256-point columns (n = 65536) with a column group for every resident
    // slot: the streaming form on a resident grid (65536 x 256: 94.5 -> 88.3
    // us, 65536 x 64: 30.4 -> 27.8, same bits); 128-point columns measured
    // slower streamed (32768 x 128: 25.8 -> 27.8 us) and keep the one-shot grid.
    if constexpr (N1 == 256) {
        static bfly_kern::ResidentCache rcs[2];
        const void* fs = temporal ? reinterpret_cast<const void*>(&k_long_cols_stream<I1, N2, DIR, true>)
                                  : reinterpret_cast<const void*>(&k_long_cols_stream<I1, N2, DIR, false>);
        const int fit_s = bfly_kern::resident(rcs[temporal ? 1 : 0], fs, 512);
        if (fit_s > 0 && units >= (uint64_t)fit_s && units <= 0xffffffffull) {
            const dim3 grid((unsigned)fit_s);
            if (temporal)
                hipLaunchKernelGGL((k_long_cols_stream<I1, N2, DIR, true>), grid, dim3(512), 0, s, in, ws, tw1, wN,
                                   idist, (uint32_t)units);
            else
                hipLaunchKernelGGL((k_long_cols_stream<I1, N2, DIR, false>), grid, dim3(512), 0, s, in, ws, tw1, wN,
                                   idist, (uint32_t)units);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            if (mid && (e = hipEventRecord(mid, s)) != hipSuccess) return e;
            hipLaunchKernelGGL((k_long_rows<I2, N1, DIR>), dim3(batch * (N1 / C2)), dim3(bfly_plan::kBlock),
                               0, s, ws, out, tw2, scale, odist, temporal);
            return hipGetLastError();
        }
    }
    if (fit > 0 && units >= (uint64_t)fit)
        hipLaunchKernelGGL((k_long_cols<I1, N2, DIR, 512>), dim3(batch * (N2 / C1w)), dim3(512), 0, s, in, ws,
                           tw1, wN, idist, temporal);
    else
        hipLaunchKernelGGL((k_long_cols<I1, N2, DIR>), dim3(batch * (N2 / C1)), dim3(bfly_plan::kBlock), 0, s,
                           in, ws, tw1, wN, idist, temporal);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (mid && (e = hipEventRecord(mid, s)) != hipSuccess) return e;
    hipLaunchKernelGGL((k_long_rows<I2, N1, DIR>), dim3(batch * (N1 / C2)), dim3(bfly_plan::kBlock),
                       0, s, ws, out, tw2, scale, odist, temporal);
    return hipGetLastError();
}

template <int I1, int I2>
hipError_t launch_long(const float2* in, float2* out, float2* ws, const float2* tw1,
                       const float2* tw2, const float2* wN, uint32_t batch, int sign,
                       hipStream_t s, hipEvent_t mid, float scale, uint64_t idist = 0, uint64_t odist = 0) {
    return sign < 0 ? launch_long_dir<I1, I2, -1>(in, out, ws, tw1, tw2, wN, batch, s, mid, scale, idist, odist)
                    : launch_long_dir<I1, I2, 1>(in, out, ws, tw1, tw2, wN, batch, s, mid, scale, idist, odist);
}

constexpr int idx(int n) { return bfly_plan::plan_index((uint32_t)n); }

}  // namespace

namespace bfly_kern {

int resident(ResidentCache& cache, const void* kernel, int block) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::atomic<int>* slot = dev >= 0 && dev < ResidentCache::kDevices ? &cache.per_device[dev] : nullptr;
    if (slot) {
        const int c = slot->load(std::memory_order_relaxed);
        if (c > 0) return c;
    }
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, 0) != hipSuccess)
        return 0;
    const int c = cus * (per > 0 ? per : 1);
    if (slot) slot->store(c, std::memory_order_relaxed);
    return c;
}

hipError_t c2c(const float* in, float* out, uint32_t n, uint32_t batch, int sign, int variant,
               const float* table, void* work, hipStream_t s, hipEvent_t mid, float scale) {
    if (variant == BFLY_VARIANT_SPLIT_MFMA || variant == BFLY_VARIANT_SPLIT_MFMA_STREAM)
        return c2c_mfma(in, out, n, batch, sign, variant == BFLY_VARIANT_SPLIT_MFMA_STREAM, table, s);
    const float2* i2 = reinterpret_cast<const float2*>(in);
    float2* o2 = reinterpret_cast<float2*>(out);
    const float2* tw = reinterpret_cast<const float2*>(table);
    switch (n) {
        case 16: return launch_c2c_local<idx(16)>(i2, o2, tw, batch, sign, variant, s, scale);
        case 32: return launch_c2c_local<idx(32)>(i2, o2, tw, batch, sign, variant, s, scale);
        case 64: return launch_c2c_local<idx(64)>(i2, o2, tw, batch, sign, variant, s, scale);
        case 128: return launch_c2c_local<idx(128)>(i2, o2, tw, batch, sign, variant, s, scale);
        case 256: return launch_c2c_local<idx(256)>(i2, o2, tw, batch, sign, variant, s, scale);
        case 512: return launch_c2c_local<idx(512)>(i2, o2, tw, batch, sign, variant, s, scale);
        case 1024: return launch_c2c_local<idx(1024)>(i2, o2, tw, batch, sign, variant, s, scale);
        case 2048: return launch_c2c_local<idx(2048)>(i2, o2, tw, batch, sign, variant, s, scale);
        case 4096: return launch_c2c_local<idx(4096)>(i2, o2, tw, batch, sign, variant, s, scale);
        default: break;
    }
    const bfly_host::TableLayout L = bfly_host::table_layout(0, n);
    if (variant == BFLY_VARIANT_STREAM_LT && n <= (uint32_t)bfly_plan::kMaxSingleN) {
        const float2 *ts = tw + L.single, *wN = tw + L.extra;
        switch (n) {
            case 8192:
                return sign < 0 ? launch_c2c_lt<idx(8192), -1>(i2, o2, ts, wN, batch, s, scale)
                                : launch_c2c_lt<idx(8192), 1>(i2, o2, ts, wN, batch, s, scale);
            case 16384:
                return sign < 0 ? launch_c2c_lt<idx(16384), -1>(i2, o2, ts, wN, batch, s, scale)
                                : launch_c2c_lt<idx(16384), 1>(i2, o2, ts, wN, batch, s, scale);
            default: break;
        }
    }
    if (variant != BFLY_VARIANT_FOUR_STEP) {
        switch (n) {
            case 8192: return launch_c2c_single<idx(8192)>(i2, o2, tw + L.single, batch, sign, s);
            case 16384: return launch_c2c_single<idx(16384)>(i2, o2, tw + L.single, batch, sign, s);
            default: break;
        }
    }
    const float2 *t1 = tw + L.stage0, *t2 = tw + L.stage1, *wN = tw + L.extra;
    float2* ws = reinterpret_cast<float2*>(work);
    switch (n) {
        case 8192: return launch_long<idx(64), idx(128)>(i2, o2, ws, t1, t2, wN, batch, sign, s, mid, scale);
        case 16384: return launch_long<idx(128), idx(128)>(i2, o2, ws, t1, t2, wN, batch, sign, s, mid, scale);
        case 32768: return launch_long<idx(128), idx(256)>(i2, o2, ws, t1, t2, wN, batch, sign, s, mid, scale);
        case 65536: return launch_long<idx(256), idx(256)>(i2, o2, ws, t1, t2, wN, batch, sign, s, mid, scale);
        default: return hipErrorInvalidValue;
    }
}

template <int I>
hipError_t launch_rows(const float2* in, uint64_t idist, float2* out, uint64_t odist, const float2* tw,
                       uint32_t batch, int sign, hipStream_t s) {
    constexpr int N = PlanOf<I>::N;
    const bfly_host::TableLayout L = bfly_host::table_layout(0, N);
    // up to 4096 the local table with [A | B] at `extra`; 8192 / 16384 the
    // one-workgroup layout (`single`) with the full W_N table at `extra`
    const float2* ts = N > bfly_plan::kMaxLocalN ? tw + L.single : tw;
    const float2* wN = tw + L.extra;
    return sign < 0 ? launch_c2c_lt<I, -1>(in, out, ts, wN, batch, s, 1.0f, idist, odist)
                    : launch_c2c_lt<I, 1>(in, out, ts, wN, batch, s, 1.0f, idist, odist);
}

hipError_t c2c_rows(const float* in, uint64_t idist, float* out, uint64_t odist, uint32_t n, uint32_t batch,
                    int sign, const float* table, void* work, hipStream_t s) {
    const float2* i2 = reinterpret_cast<const float2*>(in);
    float2* o2 = reinterpret_cast<float2*>(out);
    const float2* tw = reinterpret_cast<const float2*>(table);
    switch (n) {
        case 16: return launch_rows<idx(16)>(i2, idist, o2, odist, tw, batch, sign, s);
        case 32: return launch_rows<idx(32)>(i2, idist, o2, odist, tw, batch, sign, s);
        case 64: return launch_rows<idx(64)>(i2, idist, o2, odist, tw, batch, sign, s);
        case 128: return launch_rows<idx(128)>(i2, idist, o2, odist, tw, batch, sign, s);
        case 256: return launch_rows<idx(256)>(i2, idist, o2, odist, tw, batch, sign, s);
        case 512: return launch_rows<idx(512)>(i2, idist, o2, odist, tw, batch, sign, s);
        case 1024: return launch_rows<idx(1024)>(i2, idist, o2, odist, tw, batch, sign, s);
        case 2048: return launch_rows<idx(2048)>(i2, idist, o2, odist, tw, batch, sign, s);
        case 4096: return launch_rows<idx(4096)>(i2, idist, o2, odist, tw, batch, sign, s);
        case 8192: return launch_rows<idx(8192)>(i2, idist, o2, odist, tw, batch, sign, s);
        case 16384: return launch_rows<idx(16384)>(i2, idist, o2, odist, tw, batch, sign, s);
        default: break;
    }
    // four-step: the column pass reads the strided input rows, the row pass
    // writes the strided output rows; the workspace stays contiguous
    const bfly_host::TableLayout L = bfly_host::table_layout(0, n);
    const float2 *t1 = tw + L.stage0, *t2 = tw + L.stage1, *wN = tw + L.extra;
    float2* ws = reinterpret_cast<float2*>(work);
    switch (n) {
        case 32768:
            return launch_long<idx(128), idx(256)>(i2, o2, ws, t1, t2, wN, batch, sign, s, nullptr, 1.0f, idist, odist);
        case 65536:
            return launch_long<idx(256), idx(256)>(i2, o2, ws, t1, t2, wN, batch, sign, s, nullptr, 1.0f, idist, odist);
        default: return hipErrorInvalidValue;
    }
}

hipError_t real(int kind, const float* in, float* out, uint32_t n, uint32_t batch, uint32_t bins,
                uint32_t stride, int variant, bool adjoint, const float* table, hipStream_t s) {
    const int adj = adjoint ? 1 : 0;
    const float2* tw = reinterpret_cast<const float2*>(table);
    const bfly_host::TableLayout L = bfly_host::table_layout(kind, n);
    const float2* wn = tw + L.extra;
    switch (n / 2) {
        case 16: return launch_real<idx(16)>(kind, in, out, tw, wn, batch, bins, stride, variant, adj, s);
        case 32: return launch_real<idx(32)>(kind, in, out, tw, wn, batch, bins, stride, variant, adj, s);
        case 64: return launch_real<idx(64)>(kind, in, out, tw, wn, batch, bins, stride, variant, adj, s);
        case 128: return launch_real<idx(128)>(kind, in, out, tw, wn, batch, bins, stride, variant, adj, s);
        case 256: return launch_real<idx(256)>(kind, in, out, tw, wn, batch, bins, stride, variant, adj, s);
        case 512: return launch_real<idx(512)>(kind, in, out, tw, wn, batch, bins, stride, variant, adj, s);
        case 1024: return launch_real<idx(1024)>(kind, in, out, tw, wn, batch, bins, stride, variant, adj, s);
        case 2048: return launch_real<idx(2048)>(kind, in, out, tw, wn, batch, bins, stride, variant, adj, s);
        case 4096: return launch_real<idx(4096)>(kind, in, out, tw, wn, batch, bins, stride, variant, adj, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
    const size_t n4 = bytes / 16;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    size_t blocks = (n4 + 1023) / 1024;  // 256 threads x 4 float4 per step
    const size_t cap = (size_t)cus * 8;
    if (blocks > cap) blocks = cap;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s,
                       reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), n4);
    return hipGetLastError();
}

}  // namespace bfly_kern
