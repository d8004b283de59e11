// bfly_fft2d.hip -- 2-D C2C transforms of [batch][n1][n2] interleaved
// complex arrays (row-major: n1 rows of n2 points).
//
// X[k1][k2] = sum_{j1, j2} x[j1][j2] W_n1^{j1 k1} W_n2^{j2 k2} is separable:
// the row pass is the C2C kernels over batch * n1 rows of n2 points; the
// column pass runs k_cols, which takes C = WG / T adjacent columns per
// workgroup of WG threads (T = threads per column transform), so every load
// and store is a run of C consecutive complex values (C * 8 bytes: 128 B at
// 256-point columns, 512 B at 64).  Column lengths up to 2048 are one
// k_cols pass (512 / 1024 / 2048 in 512 / 1024 / 1024-thread workgroups);
// 4096 is split n1 = P * Q = 64 * 64 (four-step along the columns):
//
//   j1 = Q p + q,  k1 = kp + P kq
//   Y[Q kp + q][c] = W_n1^{q kp} * FFT_P over p of x[Q p + q][c]   (k_cols, twiddle)
//   X[kp + P kq][c] = FFT_Q over q of Y[Q kp + q][c]              (k_cols, row map)
//
// so the whole transform is two HBM round trips up to n1 = 2048 and three
// at 4096.  The column transforms use the LDS twiddle tables of the streaming
// C2C kernels (stockham_lt); the inter-pass twiddle is W_n1^e =
// A[e mod 64] * B[e / 64] from two LDS tables.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bfly_hip.h"
#include "bfly_kernels.hpp"
#include "bfly_tables.hpp"
#include "bfly_units.hpp"

namespace {

using namespace bfly_dev;

// Addressing of one column pass.  A batch item of `item` elements holds
// `mats` matrices; matrix m of item b reads rows at in + b*item +
// m*in_mstride (row pitch in_rstride) and writes at out + b*item +
// m*out_mstride (row pitch out_rstride); columns are `ncols` wide, walked
// by `groups` workgroups of C columns.
struct ColView {
    uint64_t item;
    uint64_t in_mstride, out_mstride;
    uint32_t in_rstride, out_rstride;
    uint32_t ncols, groups, mats;
    uint32_t col_div;  // twiddle pass: q = column / col_div
};

template <int I, int DIR, bool TW, int WG = bfly_plan::kBlock>
__global__ __launch_bounds__(WG) void k_cols(const float2* __restrict__ in, float2* __restrict__ out,
                                              const float2* __restrict__ tw, const float2* __restrict__ wab,
                                              ColView cv) {
    using PL = PlanOf<I>;
    using LT = C2cLdsTw<I>;
    constexpr int L = PL::N, T = PL::T, C = WG / T;
    constexpr int R0 = PL::R0, RL = PL::RL, P = PL::P;
    constexpr int STRIDE = lds_len(L) + 1;
    // the twiddle is read after the transform, whose LDS exchange (two or
    // more passes) orders the table writes before it
    static_assert(!TW || PL::NRAD > 1, "twiddle pass needs a multi-pass column plan");
    __shared__ float2 lds[C * STRIDE];
    __shared__ float2 ltw[LT::LEN];
    __shared__ float2 wt[TW ? 128 : 1];
    const int tid = threadIdx.x;
    // the column plan's tables: [its pass tables | A | B] as in the local C2C table
    c2c_lds_tw<I, WG>(ltw, tw, tw + bfly_plan::stage_twiddle_len(I), tid);
    if constexpr (TW)
        for (int i = tid; i < 128; i += WG) wt[i] = wab[i];
    const int c = tid % C, t = tid / C;
    // groups narrower than a 128-byte line (C = 8) share each line with the
    // next group: workgroups are dealt to the 8 XCDs in turn, so within each
    // window of 16 the pair of groups 2x, 2x+1 goes to XCD x (dealt 8 apart)
    // and meets in its L2, while the 8 XCDs still work on adjacent columns
    const uint32_t nb = gridDim.x, b0 = blockIdx.x;
    const uint32_t blk = (C < 16 && (b0 | 15) < nb) ? (b0 & ~15u) | ((b0 & 7u) << 1) | ((b0 >> 3) & 1u) : b0;
    const uint32_t g = blk % cv.groups;
    const uint64_t mi = blk / cv.groups;
    const uint64_t b = mi / cv.mats, m = mi % cv.mats;
    const uint32_t col = g * C + c;
    const bool active = col < cv.ncols;
    const float2* src = in + b * cv.item + m * cv.in_mstride + col;
    float2 v[P];
#pragma unroll
    for (int mm = 0; mm < P / R0; mm++)
#pragma unroll
        for (int r = 0; r < R0; r++)
            v[mm * R0 + r] = active ? ld_stream(&src[(uint64_t)cv.in_rstride * (t + mm * T + r * (L / R0))])
                                    : make_float2(0.f, 0.f);
    // the transform's first LDS exchange orders the table writes before any read
    run_plan_lt<I, T, DIR, LX_PAD>(v, lds + c * STRIDE, 0, ltw, ltw + LT::NLO, ltw + LT::NLO + LT::NHI, t,
                                   std::make_index_sequence<PL::NRAD>{});
    if (!active) return;
    float2* dst = out + b * cv.item + m * cv.out_mstride + col;
    const uint32_t q = TW ? col / cv.col_div : 0u;
#pragma unroll
    for (int mm = 0; mm < P / RL; mm++)
#pragma unroll
        for (int r = 0; r < RL; r++) {
            const int k = t + mm * T + r * (L / RL);
            float2 x = v[mm * RL + r];
            if constexpr (TW) {
                const uint32_t e = q * (uint32_t)k;  // < n1 <= 4096
                const float2 a = wt[e & 63], w = wt[64 + (e >> 6)];
                x = cmul_tw<DIR>(x, make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x));
            }
            st_stream(&dst[(uint64_t)cv.out_rstride * k], x);
        }
}

// Streaming form of k_cols (no twiddle, every column group full): a
// resident grid walks the units (C-column groups of every matrix, in the
// one-shot grid's order) with the next unit's loads in flight through the
// current unit's passes.  Loads and stores go through per-unit raw buffer
// descriptors (zero records past the last unit, so no branch guards the
// prefetch) with the per-row offset in the scalar operand, and the tables
// are staged once per workgroup.  Same per-column arithmetic as k_cols, so
// the same bits.  Used where one workgroup fills a CU (1024-point columns,
// 139 KiB of LDS), which otherwise loads, computes and stores in turn.
template <int I, int DIR, int WG>
__global__ __launch_bounds__(WG) void k_cols_stream(const float2* __restrict__ in, float2* __restrict__ out,
                                                    const float2* __restrict__ tw, ColView cv, uint32_t units,
                                                    uint32_t in_span, uint32_t out_span) {
    using PL = PlanOf<I>;
    using LT = C2cLdsTw<I>;
    constexpr int L = PL::N, T = PL::T, C = WG / T;
    constexpr int R0 = PL::R0, RL = PL::RL, P = PL::P;
    constexpr int STRIDE = lds_len(L) + 1;
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    __shared__ float2 lds[C * STRIDE];
    __shared__ float2 ltw[LT::LEN];
    const int tid = threadIdx.x;
    c2c_lds_tw<I, WG>(ltw, tw, tw + bfly_plan::stage_twiddle_len(I), tid);
    // (the first unit's first LDS exchange orders these writes before the reads)
    const int c = tid % C, t0 = tid / C;
    // element offsets of unit u's first column, input and output side
    auto base = [&](uint32_t u, uint64_t& oin, uint64_t& oout) {
        const uint32_t g = u % cv.groups, mi = u / cv.groups;
        const uint32_t b = mi / cv.mats, m = mi % cv.mats;
        oin = (uint64_t)b * cv.item + (uint64_t)m * cv.in_mstride + (uint64_t)g * C;
        oout = (uint64_t)b * cv.item + (uint64_t)m * cv.out_mstride + (uint64_t)g * C;
    };
    auto load_unit = [&](float2 (&v)[P], uint32_t u, bool valid, int t) {
        uint64_t oin, oout;
        base(u, oin, oout);
        const auto rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(in + oin), (short)0,
                                                          valid ? (int)in_span : 0, 0x00020000);
        const uint32_t voff = ((uint32_t)t * cv.in_rstride + (uint32_t)c) * 8u;
#pragma unroll
        for (int mm = 0; mm < P / R0; mm++)
#pragma unroll
            for (int r = 0; r < R0; r++) {
                const uint32_t soff = (uint32_t)(mm * T + r * (L / R0)) * cv.in_rstride * 8u;
                v[mm * R0 + r] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 2));
            }
    };
    float2 cur[P], nxt[P];
    auto step = [&](uint32_t u) {
        int t = t0;  // opaque per-step thread index (see k_r2c_stream)
        asm volatile("" : "+v"(t));
        const uint32_t un = u + gridDim.x;
        load_unit(nxt, un < units ? un : u, un < units, t);
        run_plan_lt<I, T, DIR, LX_PAD>(cur, lds + c * STRIDE, 0, ltw, ltw + LT::NLO, ltw + LT::NLO + LT::NHI, t,
                                       std::make_index_sequence<PL::NRAD>{});
        uint64_t oin, oout;
        base(u, oin, oout);
        const auto os = __builtin_amdgcn_make_buffer_rsrc(out + oout, (short)0, (int)out_span, 0x00020000);
        const uint32_t voff = ((uint32_t)t * cv.out_rstride + (uint32_t)c) * 8u;
#pragma unroll
        for (int mm = 0; mm < P / RL; mm++)
#pragma unroll
            for (int r = 0; r < RL; r++) {
                const uint32_t soff = (uint32_t)(mm * T + r * (L / RL)) * cv.out_rstride * 8u;
                __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, cur[mm * RL + r]), os, voff, soff, 2);
            }
        // the stores issue before the register copy waits for the prefetch
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < P; i++) cur[i] = nxt[i];
    };
    uint32_t u = blockIdx.x;  // the grid never exceeds the unit count
    load_unit(cur, u, true, t0);
    step(u);
    for (u += gridDim.x; u < units; u += gridDim.x) step(u);
}

template <int I, bool TW, int WG = bfly_plan::kBlock>
hipError_t launch_cols(const float2* in, float2* out, const float2* tw, const float2* wab, ColView cv,
                       uint64_t batch, int sign, hipStream_t s) {
    constexpr int C = WG / PlanOf<I>::T;
    cv.groups = (cv.ncols + C - 1) / C;
    const uint64_t blocks = batch * cv.mats * cv.groups;
    if (blocks == 0 || blocks > 0x7fffffffull) return hipErrorInvalidValue;
    // 1024-point columns: the streaming form when every group is full, a
    // unit's span fits a buffer descriptor and there is a unit for every
    // resident slot (2-D 1024x1024 x16: 110.4 -> 106.4 us, 1024x128 x128:
    // 95.6 -> 91.7; at 512 and 256 points, where two or more workgroups
    // share a CU, it measured 0.5-2 % slower)
    if constexpr (!TW && PlanOf<I>::N == 1024 && C >= 16) {
        const uint64_t in_span = ((uint64_t)(PlanOf<I>::N - 1) * cv.in_rstride + C) * sizeof(float2);
        const uint64_t out_span = ((uint64_t)(PlanOf<I>::N - 1) * cv.out_rstride + C) * sizeof(float2);
        if (cv.ncols % C == 0 && in_span < (1ull << 31) && out_span < (1ull << 31)) {
            static bfly_kern::ResidentCache rc[2];
            const void* f = sign < 0 ? reinterpret_cast<const void*>(&k_cols_stream<I, -1, WG>)
                                     : reinterpret_cast<const void*>(&k_cols_stream<I, 1, WG>);
            const int fit = bfly_kern::resident(rc[sign < 0 ? 0 : 1], f, WG);
            if (fit > 0 && blocks >= (uint64_t)fit) {
                if (sign < 0)
                    hipLaunchKernelGGL((k_cols_stream<I, -1, WG>), dim3((unsigned)fit), dim3(WG), 0, s, in, out, tw,
                                       cv, (uint32_t)blocks, (uint32_t)in_span, (uint32_t)out_span);
                else
                    hipLaunchKernelGGL((k_cols_stream<I, 1, WG>), dim3((unsigned)fit), dim3(WG), 0, s, in, out, tw,
                                       cv, (uint32_t)blocks, (uint32_t)in_span, (uint32_t)out_span);
                return hipGetLastError();
            }
        }
    }
    if (sign < 0)
        hipLaunchKernelGGL((k_cols<I, -1, TW, WG>), dim3((unsigned)blocks), dim3(WG), 0, s, in, out, tw, wab, cv);
    else
        hipLaunchKernelGGL((k_cols<I, 1, TW, WG>), dim3((unsigned)blocks), dim3(WG), 0, s, in, out, tw, wab, cv);
    return hipGetLastError();
}

template <bool TW>
hipError_t cols(uint32_t len, const float2* in, float2* out, const float2* tw, const float2* wab, ColView cv,
                uint64_t batch, int sign, hipStream_t s) {
    constexpr auto idx = bfly_plan::plan_index;
    // the twiddle pass runs the first factor of a split length (32 or 64)
    if constexpr (!TW)
        if (len == 16) return launch_cols<idx(16), TW>(in, out, tw, wab, cv, batch, sign, s);
    switch (len) {
        case 32: return launch_cols<idx(32), TW>(in, out, tw, wab, cv, batch, sign, s);
        case 64: return launch_cols<idx(64), TW>(in, out, tw, wab, cv, batch, sign, s);
        case 128: return launch_cols<idx(128), TW>(in, out, tw, wab, cv, batch, sign, s);
        case 256: return launch_cols<idx(256), TW>(in, out, tw, wab, cv, batch, sign, s);
        default: break;
    }
    // 512 / 1024 / 2048-point columns in one pass: 16 / 16 / 8 columns per
    // workgroup in 512 / 1024 / 1024-thread workgroups (70 / 139 / 139 KiB
    // of LDS)
    if constexpr (!TW) {
        if (len == 512) return launch_cols<idx(512), TW, 512>(in, out, tw, wab, cv, batch, sign, s);
        if (len == 1024) return launch_cols<idx(1024), TW, 1024>(in, out, tw, wab, cv, batch, sign, s);
        if (len == 2048) return launch_cols<idx(2048), TW, 1024>(in, out, tw, wab, cv, batch, sign, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace

namespace bfly_kern {

hipError_t columns(const float* in, float* out, uint32_t n1, uint32_t ncols, uint32_t batch, int sign,
                   const float* table_cols, void* work, hipStream_t s) {
    const float2* i2 = reinterpret_cast<const float2*>(in);
    float2* o2 = reinterpret_cast<float2*>(out);
    float2* ws = reinterpret_cast<float2*>(work);
    const float2* tc = reinterpret_cast<const float2*>(table_cols);
    const bfly_host::Fft2dLayout L = bfly_host::fft2d_layout(n1);
    const uint64_t item = (uint64_t)n1 * ncols;
    if (L.q == 1) {
        ColView cv{item, 0, 0, ncols, ncols, ncols, 0, 1, 1};
        return cols<false>(n1, i2, o2, tc + L.p_table, nullptr, cv, batch, sign, s);
    }
    // P-point pass with the twiddle: in -> ws; Q-point pass with the row map: ws -> out
    const uint32_t qn = L.q * ncols;
    ColView a{item, 0, 0, qn, qn, qn, 0, 1, ncols};
    hipError_t e = cols<true>(L.p, i2, ws, tc + L.p_table, tc + L.wab, a, batch, sign, s);
    if (e != hipSuccess) return e;
    ColView b{item, qn, ncols, ncols, L.p * ncols, ncols, 0, L.p, 1};
    return cols<false>(L.q, ws, o2, tc + L.q_table, nullptr, b, batch, sign, s);
}

hipError_t fft2d(const float* in, float* out, uint32_t n1, uint32_t n2, uint32_t batch, int sign,
                 const float* table_rows, const float* table_cols, void* work, hipStream_t s) {
    // the row pass takes the C2C kernel the library picks for (n2, batch * n1)
    const int rows = bfly_host::auto_variant(n2, batch * n1);
    if (bfly_host::fft2d_layout(n1).q == 1) {
        // rows in -> work, columns work -> out
        float* w = static_cast<float*>(work);
        hipError_t e = c2c(in, w, n2, batch * n1, sign, rows, table_rows, nullptr, s);
        return e != hipSuccess ? e : columns(w, out, n1, n2, batch, sign, table_cols, nullptr, s);
    }
    // rows in -> out, columns out -> work -> out
    hipError_t e = c2c(in, out, n2, batch * n1, sign, rows, table_rows, nullptr, s);
    return e != hipSuccess ? e : columns(out, out, n1, n2, batch, sign, table_cols, work, s);
}

}  // namespace bfly_kern
