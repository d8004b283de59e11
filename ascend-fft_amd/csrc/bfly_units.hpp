// bfly_units.hpp -- the per-workgroup work units of the gfx950 kernels.
//
// A work unit is B = WG/T consecutive transforms of one batch (T = n/16
// threads each) taken through one HBM read and one HBM write.  The one-shot
// kernels run one unit per workgroup; the grouped kernel runs units of many
// requests from one persistent launch.  Both call these functions, so a
// transform computes the same bits whichever kernel runs it.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>

#include "bfly_device.hpp"
#include "bfly_plans.hpp"

namespace bfly_dev {

using bfly_plan::kPlans;

// Plan I under radix schedule SC (bfly_plan::radix_at).
template <int I, int SC = 0>
struct PlanOf {
    static constexpr int N = kPlans[I].n;
    static constexpr int NRAD = kPlans[I].nrad;
    static constexpr int R0 = bfly_plan::radix_at(I, SC, 0);
    static constexpr int RL = bfly_plan::radix_at(I, SC, NRAD - 1);
    static constexpr int P = bfly_plan::kPoints;
    static constexpr int T = N / P;  // threads per transform
};

template <int I, int T, int DIR, int LX, int SC = 0, size_t... S>
__device__ __forceinline__ void run_plan(float2 (&v)[bfly_plan::kPoints], float2* lds, int rb,
                                         const float2* __restrict__ tw, int t,
                                         std::index_sequence<S...>) {
    stockham<kPlans[I].n, bfly_plan::kPoints, T, DIR, LX, 1, bfly_plan::radix_at(I, SC, S)...>(
        v, lds, rb, tw, t);
}

template <int I, int T, int DIR, int LX, size_t... S>
__device__ __forceinline__ void run_plan_rt(float2 (&v)[bfly_plan::kPoints], float2* lds, int rb,
                                            const float2* tr, int t, std::index_sequence<S...>) {
    stockham_rt<kPlans[I].n, bfly_plan::kPoints, T, DIR, LX, 1, kPlans[I].rad[S]...>(v, lds, rb, tr, t);
}

template <int I, int T, int DIR, int LX, size_t... S>
__device__ __forceinline__ void run_plan_lt(float2 (&v)[bfly_plan::kPoints], float2* lds, int rb,
                                            const float2* tlo, const float2* thi, const float2* t1, int t,
                                            std::index_sequence<S...>) {
    stockham_lt<kPlans[I].n, bfly_plan::kPoints, T, DIR, LX, 1, kPlans[I].rad[S]...>(v, lds, rb, tlo, thi,
                                                                                     t1, t);
}

template <int I, int T, size_t... S>
__device__ __forceinline__ void load_plan_twiddles(float2* tr, const float2* __restrict__ tw, int t,
                                                   std::index_sequence<S...>) {
    load_twiddles<kPlans[I].n, bfly_plan::kPoints, T, 1, kPlans[I].rad[S]...>(tr, tw, t);
}

template <int I, size_t... S>
constexpr int plan_twiddle_regs(std::index_sequence<S...>) {
    return TwRegCount<bfly_plan::kPoints, 1, kPlans[I].rad[S]...>::value;
}

// Padded-region form used by the real and long kernels: `lds` is the
// transform's own region.
template <int I, int T, int DIR, int SC = 0>
__device__ __forceinline__ void fft_local(float2 (&v)[bfly_plan::kPoints], float2* lds,
                                          const float2* __restrict__ tw, int t) {
    run_plan<I, T, DIR, LX_PAD, SC>(v, lds, 0, tw, t, std::make_index_sequence<kPlans[I].nrad>{});
}

// ------------------------------------------------------------------ C2C
// One work unit: the WG/T transforms g*B .. g*B+B-1 of a batch (B = WG/T),
// one HBM read and one HBM write per element.  `lds` holds B regions; `tid`
// is the thread's index in the workgroup.
template <int I, int DIR, int LX, int WG, int SC = 0>
__device__ __forceinline__ void c2c_unit(const float2* __restrict__ in, float2* __restrict__ out,
                                         const float2* __restrict__ tw, uint32_t batch, uint32_t g,
                                         float2* lds, int tid) {
    using PL = PlanOf<I, SC>;
    constexpr int N = PL::N, T = PL::T, B = WG / T;
    constexpr int R0 = PL::R0, RL = PL::RL, P = PL::P;
    static_assert(WG % T == 0, "a work unit holds whole transforms");
    const int bl = tid / T, t = tid % T;
    const uint64_t tr = (uint64_t)g * B + bl;
    const bool active = tr < batch;
    float2 v[P];
    if (active) {
        const float2* src = in + tr * N;
#pragma unroll
        for (int m = 0; m < P / R0; m++)
#pragma unroll
            for (int r = 0; r < R0; r++) v[m * R0 + r] = ld_io<(T >= 16)>(&src[t + m * T + r * (N / R0)]);
    } else {
#pragma unroll
        for (int i = 0; i < P; i++) v[i] = make_float2(0.f, 0.f);
    }
    run_plan<I, T, DIR, LX, SC>(v, LX == LX_SWZ ? lds : lds + bl * lds_len(N),
                                LX == LX_SWZ ? bl * N : 0, tw, t, std::make_index_sequence<PL::NRAD>{});
    if (active) {
        float2* dst = out + tr * N;
#pragma unroll
        for (int m = 0; m < P / RL; m++)
#pragma unroll
            for (int r = 0; r < RL; r++) st_io<(T >= 16)>(&dst[t + m * T + r * (N / RL)], v[m * RL + r]);
    }
}

// The same unit with the inter-pass twiddles read from LDS tables (the
// per-transform code of k_c2c_stream_lt): `ltw` = [A (NLO) | B (NHI) | the
// second pass's table stored [r 16 + k]], filled by c2c_lds_tw.
template <int I>
struct C2cLdsTw {
    static constexpr int N = kPlans[I].n;
    static constexpr int NLO = N < TW2_LO ? N : TW2_LO;
    static constexpr int NHI = N / NLO;
    static constexpr int R1 = kPlans[I].rad[kPlans[I].nrad > 1 ? 1 : 0];
    static constexpr int NT1 = kPlans[I].nrad > 1 ? bfly_plan::kPoints * R1 : 1;
    static constexpr int LEN = NLO + NHI + NT1;
};

// Fills `buf` (C2cLdsTw<I>::LEN entries) from the plan's pass tables `tw` and
// `ab` = [W_N^j, j < NLO | W_N^{NLO j}, j < NHI] in device memory.  The
// caller synchronises before the tables are read.
template <int I, int WG>
__device__ __forceinline__ void c2c_lds_tw(float2* buf, const float2* __restrict__ tw,
                                           const float2* __restrict__ ab, int tid) {
    using L = C2cLdsTw<I>;
    for (int i = tid; i < L::NLO + L::NHI; i += WG) buf[i] = ab[i];
    if constexpr (kPlans[I].nrad > 1)
        for (int i = tid; i < L::NT1; i += WG) buf[L::NLO + L::NHI + (i % L::R1) * bfly_plan::kPoints + i / L::R1] = tw[i];
}

template <int I, int DIR, int WG>
__device__ __forceinline__ void c2c_unit_lt(const float2* __restrict__ in, float2* __restrict__ out,
                                            const float2* ltw, uint32_t batch, uint32_t g, float2* lds, int tid) {
    using PL = PlanOf<I>;
    using L = C2cLdsTw<I>;
    constexpr int N = PL::N, T = PL::T, B = WG / T;
    constexpr int R0 = PL::R0, RL = PL::RL, P = PL::P;
    static_assert(WG % T == 0, "a work unit holds whole transforms");
    const int bl = tid / T, t = tid % T;
    const uint64_t tr = (uint64_t)g * B + bl;
    const bool active = tr < batch;
    float2 v[P];
    if (active) {
        const float2* src = in + tr * N;
#pragma unroll
        for (int m = 0; m < P / R0; m++)
#pragma unroll
            for (int r = 0; r < R0; r++) v[m * R0 + r] = ld_io<(T >= 16)>(&src[t + m * T + r * (N / R0)]);
    } else {
#pragma unroll
        for (int i = 0; i < P; i++) v[i] = make_float2(0.f, 0.f);
    }
    run_plan_lt<I, T, DIR, LX_PAD>(v, lds + bl * lds_len(N), 0, ltw, ltw + L::NLO, ltw + L::NLO + L::NHI, t,
                                   std::make_index_sequence<PL::NRAD>{});
    if (active) {
        float2* dst = out + tr * N;
#pragma unroll
        for (int m = 0; m < P / RL; m++)
#pragma unroll
            for (int r = 0; r < RL; r++) st_io<(T >= 16)>(&dst[t + m * T + r * (N / RL)], v[m * RL + r]);
    }
}

// --------------------------------------------------- real twiddle sources
// Where a real unit takes its twiddles from: the plan's pass tables (for
// the n/2-point FFT) and W_n^k, k <= n/2 (for the split twiddle).
// RealTwGlobal reads both tables from global memory; RealTwLds reads LDS
// copies made once per workgroup: the two-level tables of the FFT
// (stockham_lt) and W_n^k = wlo[k mod 64] * whi[k / 64].
template <int I, int SC>
struct RealTwGlobal {
    const float2* __restrict__ tw;
    const float2* __restrict__ wn;
    template <int T, int DIR>
    __device__ __forceinline__ void fft(float2 (&v)[bfly_plan::kPoints], float2* L, int t) const {
        fft_local<I, T, DIR, SC>(v, L, tw, t);
    }
    __device__ __forceinline__ float2 w(int k) const { return wn[k]; }
};

template <int I>
struct RealTwLds {
    const float2* tlo;  // W_M^j, j < 64
    const float2* thi;  // W_M^{64 j}, j < M / 64
    const float2* t1;   // second pass's table, [r 16 + k]
    const float2* wlo;  // W_n^j, j < 64
    const float2* whi;  // W_n^{64 j}, 64 j <= n / 2
    template <int T, int DIR>
    __device__ __forceinline__ void fft(float2 (&v)[bfly_plan::kPoints], float2* L, int t) const {
        run_plan_lt<I, T, DIR, LX_PAD>(v, L, 0, tlo, thi, t1, t, std::make_index_sequence<kPlans[I].nrad>{});
    }
    __device__ __forceinline__ float2 w(int k) const {
        const float2 a = wlo[k & 63], b = whi[k >> 6];
        return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
    }
};

// LDS entries RealTwLds needs for plan I (n = 2 N_I), and the copy that fills
// them from the global tables of the real kind: tw (pass tables) and wn
// (W_n^k, k <= n/2; W_n^e for e > n/2 is conj W_n^{n-e}).
template <int I>
constexpr int real_lds_tw_len() {
    constexpr int M = kPlans[I].n;
    return 64 + M / 64 + bfly_plan::kPoints * kPlans[I].rad[1] + 64 + M / 64 + 1;
}
template <int I, int WG>
__device__ __forceinline__ RealTwLds<I> real_lds_tw(float2* buf, const float2* __restrict__ tw,
                                                    const float2* __restrict__ wn, int tid) {
    constexpr int M = kPlans[I].n, n = 2 * M, R1 = kPlans[I].rad[1];
    static_assert(M >= 128 && kPlans[I].rad[0] == 16, "two-level tables need M >= 128");
    float2* tlo = buf;
    float2* thi = tlo + 64;
    float2* t1 = thi + M / 64;
    float2* wlo = t1 + 16 * R1;
    float2* whi = wlo + 64;
    auto wfull = [&](int e) {  // W_n^e, 0 <= e < n
        if (e <= M) return wn[e];
        const float2 c = wn[n - e];
        return make_float2(c.x, -c.y);
    };
    for (int i = tid; i < real_lds_tw_len<I>(); i += WG) {
        if (i < 64) tlo[i] = wfull(2 * i);
        else if (i < 64 + M / 64) thi[i - 64] = wfull(128 * (i - 64));
        else if (i < 64 + M / 64 + 16 * R1) {
            const int j = i - 64 - M / 64;  // stage-table entry [k R1 + r]
            t1[(j % R1) * 16 + j / R1] = tw[j];
        } else if (i < 64 + M / 64 + 16 * R1 + 64) {
            wlo[i - 64 - M / 64 - 16 * R1] = wn[i - 64 - M / 64 - 16 * R1];
        } else {
            whi[i - 128 - M / 64 - 16 * R1] = wn[64 * (i - 128 - M / 64 - 16 * R1)];
        }
    }
    return RealTwLds<I>{tlo, thi, t1, wlo, whi};
}

// ------------------------------------------------------------------ R2C
// z[j] = x[2j] + i x[2j+1], Z = FFT_M(z), M = n/2; for k = 0..M:
//   E = (Z[k] + conj Z[M-k]) / 2,  O = (Z[k] - conj Z[M-k]) / (2i),
//   X[k] = E + W_n^k O            (Z[M] == Z[0]).
// Writes bins k < kout of each row at a row stride of ostride floats (the
// full transform: kout = M+1, ostride = 2M+2; a mode-truncated one writes
// only its first kout bins).
// Load phase: the real row of transform g*B + bl read as M complex values
// (zeros past the batch).
template <int I, int WG, int SC = 0>
__device__ __forceinline__ void r2c_load(float2 (&v)[bfly_plan::kPoints], const float* __restrict__ in,
                                         uint32_t batch, uint32_t g, int tid) {
    using PL = PlanOf<I, SC>;
    constexpr int M = PL::N, T = PL::T, B = WG / T, R0 = PL::R0, P = PL::P;
    const int bl = tid / T, t = tid % T;
    const uint64_t tr = (uint64_t)g * B + bl;
    if (tr < batch) {
        const float2* src = reinterpret_cast<const float2*>(in + tr * (uint64_t)(2 * M));
#pragma unroll
        for (int m = 0; m < P / R0; m++)
#pragma unroll
            for (int r = 0; r < R0; r++) v[m * R0 + r] = ld_io<(T >= 16)>(&src[t + m * T + r * (M / R0)]);
    } else {
#pragma unroll
        for (int i = 0; i < P; i++) v[i] = make_float2(0.f, 0.f);
    }
}

// Compute phase: FFT_M, spectrum parked in LDS, split post-twiddle, store.
// `adj` (uniform) stores the adjoint of the C2R instead: X[k] v_k / n with
// v = 1 at bins 0 and M (imaginary part zero; the C2R reads only the real
// parts there) and 2 elsewhere.  Every factor is a power of two, so this is
// bit for bit the forward spectrum rescaled afterwards.
// Ends with LDS reads still in flight in other threads: a caller that reuses
// `lds` synchronises first.
template <int I, int WG, int SC = 0, class TWS = RealTwGlobal<I, SC>>
__device__ __forceinline__ void r2c_finish(float2 (&v)[bfly_plan::kPoints], float* __restrict__ out,
                                           const TWS& tws, uint32_t batch, uint32_t kout, uint32_t ostride,
                                           uint32_t g, float2* lds, int tid, bool adj = false) {
    using PL = PlanOf<I, SC>;
    constexpr int M = PL::N, T = PL::T, B = WG / T, RL = PL::RL, P = PL::P;
    static_assert(WG % T == 0, "a work unit holds whole transforms");
    const int bl = tid / T, t = tid % T;
    const uint64_t tr = (uint64_t)g * B + bl;
    float2* L = lds + bl * lds_len(M);
    tws.template fft<T, -1>(v, L, t);
    // Every LDS read of the last exchange completed before the barrier that
    // ends it, so the spectrum can be parked immediately.
#pragma unroll
    for (int m = 0; m < P / RL; m++)
#pragma unroll
        for (int r = 0; r < RL; r++) L[lds_idx(t + m * T + r * (M / RL))] = v[m * RL + r];
    __syncthreads();
    if constexpr (T >= 64) {
        // Bins j and M - j from one pair of reads: with E, O the even / odd
        // spectra at j, X[j] = E + W_n^j O and X[M-j] = conj(E - W_n^j O)
        // (E and O at M - j are the conjugates of those at j, W_n^{M-j} =
        // -conj W_n^j), so each thread reads half the bins and one twiddle
        // per pair.  j = M/2 is its own mirror and is stored once.  Measured
        // (MI355X, LDS twiddles): 4096 x 4096 27.9 -> 26.9 us, 8192 x 2048
        // 28.9 -> 27.0, 2048 x 8192 28.2 -> 27.7; at T = 32 the mirrored
        // stores, which cannot be line-aligned together with the ascending
        // ones, made 1024 x 16384 slower (31.9 against 27.6 us below).
        // The one-shot and streaming forms share this code, so they stay
        // bit-equal to each other.
        if (tr < batch) {
            float2* dst = reinterpret_cast<float2*>(out + tr * (uint64_t)ostride);
            constexpr int H = M / 2;
            // the ascending stores start on a 128-byte line (j shifted by the
            // row's offset in its line, as below; the mirrored ones cannot
            // be aligned at the same time): 2048 x 8192 27.5 -> 26.7 us,
            // 4096 x 4096 27.1 -> 26.3; at T = 256 (8192 x 2048) it measured
            // 27.0 -> 27.3 and is left out
            constexpr bool kShift = T <= 128;
            const int s = kShift ? (int)((reinterpret_cast<uintptr_t>(dst) >> 3) & 15) : 0;
#pragma unroll
            for (int i = 0; i <= H / T + (kShift ? 1 : 0); i++) {
                const int j = t + i * T - s;
                // truncated spectra (kout <= M) skip the pairs they do not store
                if (j >= 0 && j <= H && (j < (int)kout || M - j < (int)kout)) {
                    const float2 a = L[lds_idx(j & (M - 1))];
                    const float2 b = L[lds_idx((M - j) & (M - 1))];
                    const float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
                    const float2 o = make_float2(0.5f * (a.y + b.y), -0.5f * (a.x - b.x));
                    const float2 wo = cmul_tw<-1>(o, tws.w(j));
                    float2 x = cadd(e, wo);
                    float2 y = make_float2(e.x - wo.x, wo.y - e.y);
                    if (adj) {
                        const float f = 1.0f / (float)M;
                        if (j == 0) {
                            x = make_float2(x.x * (0.5f * f), 0.f);
                            y = make_float2(y.x * (0.5f * f), 0.f);
                        } else {
                            x = make_float2(x.x * f, x.y * f);
                            y = make_float2(y.x * f, y.y * f);
                        }
                    }
                    if (j < (int)kout) st_io<true>(&dst[j], x);
                    if (j != H && M - j < (int)kout) st_io<true>(&dst[M - j], y);
                }
            }
        }
        return;
    }
    if (tr < batch) {
        float2* dst = reinterpret_cast<float2*>(out + tr * (uint64_t)ostride);
        // Rows of n+2 floats start 8 bytes further into a 128-byte line each
        // time.  Where a wave's store covers only a few lines (T <= 32 here),
        // the bins are shifted by the row's offset in its line, so that every
        // store starts on a line boundary (same bins, same arithmetic, other
        // lanes): R2C 1024 x 16384 31.5 -> 27.7 us.  A transform of T < 16
        // threads stores T 8-byte bins at a time, aligned to 8 T bytes the
        // same way: 64 x 262144 52.9 -> 48.8 us, 128 x 131072 36.7 -> 35.9
        // (at T = 8, 256 x 65536, it measured neutral and is left out).
        const uintptr_t off = reinterpret_cast<uintptr_t>(dst) >> 3;
        const int s = T >= 16 ? (int)(off & 15) : T <= 4 ? (int)(off & (T - 1)) : 0;
#pragma unroll
        for (int i = 0; i <= P; i++) {
            const int k = t + i * T - s;
            if (k >= 0 && k <= M && k < (int)kout) {
                const float2 a = L[lds_idx(k & (M - 1))];
                const float2 b = L[lds_idx((M - k) & (M - 1))];
                const float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
                // O = (a - conj b) / (2i) = (0.5 (a.y + b.y), -0.5 (a.x - b.x))
                const float2 o = make_float2(0.5f * (a.y + b.y), -0.5f * (a.x - b.x));
                const float2 w = tws.w(k);
                const float2 wo = cmul_tw<-1>(o, w);
                float2 x = cadd(e, wo);
                if (adj) {
                    const bool edge = k == 0 || k == M;
                    const float f = edge ? 0.5f / (float)M : 1.0f / (float)M;
                    x = make_float2(x.x * f, edge ? 0.f : x.y * f);
                }
                st_io<(T >= 16)>(&dst[k], x);
            }
        }
    }
}

template <int I, int WG, int SC = 0>
__device__ __forceinline__ void r2c_unit(const float* __restrict__ in, float* __restrict__ out,
                                         const float2* __restrict__ tw, const float2* __restrict__ wn,
                                         uint32_t batch, uint32_t kout, uint32_t ostride, uint32_t g,
                                         float2* lds, int tid) {
    float2 v[bfly_plan::kPoints];
    r2c_load<I, WG, SC>(v, in, batch, g, tid);
    r2c_finish<I, WG, SC>(v, out, RealTwGlobal<I, SC>{tw, wn}, batch, kout, ostride, g, lds, tid);
}

// ------------------------------------------------------------------ C2R
// Z'[k] = (X[k] + conj X[M-k]) + i W_n^{-k} (X[k] - conj X[M-k]), k < M,
// with Im X[0] = Im X[M] = 0; z = FFT^+_M(Z') / n; x[2j] = Re z, x[2j+1] = Im z.
// Reads bins k < kin of each row at a row stride of istride floats; bins at
// or above kin are zero (the full transform: kin = M+1, istride = 2M+2).
// Load phase: bins 0..M of transform g*B + bl (k < kin, zero above), bin
// k = t + i T in q[i] (i <= P; only t = 0 holds bin M).
template <int I, int WG, int SC = 0>
__device__ __forceinline__ void c2r_load(float2 (&q)[bfly_plan::kPoints + 1], const float* __restrict__ in,
                                         uint32_t batch, uint32_t kin, uint32_t istride, uint32_t g,
                                         int tid) {
    using PL = PlanOf<I, SC>;
    constexpr int M = PL::N, T = PL::T, B = WG / T, P = PL::P;
    const int bl = tid / T, t = tid % T;
    const uint64_t tr = (uint64_t)g * B + bl;
    const bool active = tr < batch;
    static_assert(P * T == M, "bins t + i T, i < P, cover 0 .. M-1");
    const float2* src = reinterpret_cast<const float2*>(in + tr * (uint64_t)istride);
    if constexpr (T < 16) {
        // short rows: bin M loaded as the others
#pragma unroll
        for (int i = 0; i <= P; i++) {
            const int k = t + i * T;
            q[i] = (active && k <= M && k < (int)kin) ? ld_io<false>(&src[k]) : make_float2(0.f, 0.f);
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < P; i++) {
        const int k = t + i * T;
        q[i] = (active && k < (int)kin) ? ld_io<true>(&src[k]) : make_float2(0.f, 0.f);
    }
    // bin M (thread 0 only): its imaginary part is never read (the
    // pre-twiddle zeroes it), so only the real part is loaded and held.  The
    // other threads read the last bin of their own row instead (an input
    // row past the batch reads the first row) and drop it: no branch around
    // the load, whose join would make the compiler wait for every load in
    // flight, the prefetch included.  Measured with the AUTO kernels: C2R
    // 1024 x 16384 27.6 -> 26.8 us; it also brings the 8192-point
    // LDS-twiddle kernel within 128 VGPRs without spilling (four
    // workgroups per CU).
    // A truncated input (kin <= M, uniform) has no bin M and skips the load.
    q[P] = make_float2(0.f, 0.f);
    if ((int)kin > M) {
        const int kM = t + P * T;
        const bool okM = active && kM == M;
        const float* fM = reinterpret_cast<const float*>(active ? &src[okM ? kM : M] : reinterpret_cast<const float2*>(in));
        const float xm = __builtin_nontemporal_load(fM);
        q[P].x = okM ? xm : 0.f;
    }
}

// The two halves of c2r_finish, for a caller that issues other work (the
// next rows' loads) between them, or whose kernel the compiler schedules
// better through them.  c2r_finish keeps its own single body: built from
// these two it compiled differently and measured slower for the
// register-twiddle streaming kernel at 8 threads per transform (256 x 65536
// C2R 30.2 -> 34 us).  Staging: the thread's bins into its transform's LDS
// region.
template <int I, int WG, int SC = 0>
__device__ __forceinline__ void c2r_stage(const float2 (&q)[bfly_plan::kPoints + 1], float2* lds, int tid) {
    using PL = PlanOf<I, SC>;
    constexpr int M = PL::N, T = PL::T, P = PL::P;
    static_assert(lds_len(M) >= M + 1, "staging needs M+1 bins");
    static_assert(WG % T == 0, "a work unit holds whole transforms");
    const int bl = tid / T, t = tid % T;
    float2* L = lds + bl * lds_len(M);
#pragma unroll
    for (int i = 0; i <= P; i++) {
        const int k = t + i * T;
        if (k <= M) L[k] = q[i];
    }
}

// The rest of the compute phase, after the staging has been synchronised:
// pre-twiddle from the staged bins, inverse FFT_M, reals scaled and stored.
template <int I, int WG, int SC = 0, class TWS = RealTwGlobal<I, SC>>
__device__ __forceinline__ void c2r_compute(float* __restrict__ out, const TWS& tws, uint32_t batch, float scale,
                                            uint32_t g, float2* lds, int tid, bool adj) {
    using PL = PlanOf<I, SC>;
    constexpr int M = PL::N, T = PL::T, B = WG / T, R0 = PL::R0, RL = PL::RL, P = PL::P;
    const int bl = tid / T, t = tid % T;
    const uint64_t tr = (uint64_t)g * B + bl;
    float2* L = lds + bl * lds_len(M);
    float2 v[P];
#pragma unroll
    for (int m = 0; m < P / R0; m++)
#pragma unroll
        for (int r = 0; r < R0; r++) {
            const int k = t + m * T + r * (M / R0);
            float2 a = L[k], b = L[M - k];
            if (k == 0) {
                a.y = 0.f;
                b.y = 0.f;
                if (adj) { a.x *= 2.f; b.x *= 2.f; }
            }
            const float2 s = make_float2(a.x + b.x, a.y - b.y);
            const float2 d = make_float2(a.x - b.x, a.y + b.y);
            const float2 wd = cmul_tw<1>(d, tws.w(k));  // W_n^{-k} d
            v[m * R0 + r] = make_float2(s.x - wd.y, s.y + wd.x);
        }
    __syncthreads();
    tws.template fft<T, 1>(v, L, t);
    if (tr >= batch) return;
    float2* dst = reinterpret_cast<float2*>(out + tr * (uint64_t)(2 * M));
#pragma unroll
    for (int m = 0; m < P / RL; m++)
#pragma unroll
        for (int r = 0; r < RL; r++) {
            const float2 z = v[m * RL + r];
            st_io<(T >= 16)>(&dst[t + m * T + r * (M / RL)], make_float2(z.x * scale, z.y * scale));
        }
}

// Compute phase: bins staged in LDS, Hermitian pre-twiddle, inverse FFT_M,
// reals scaled and stored.  `adj` (uniform) doubles the k = 0 term of the
// pre-twiddle (bins 0 and M); with scale = 1/2 the result is the adjoint of
// the R2C, n * C2R(w X) with w = 1 at bins 0 and M and 1/2 elsewhere, bit for
// bit (every factor is a power of two).  Starts by writing `lds`: a caller
// that reuses it synchronises first.
template <int I, int WG, int SC = 0, class TWS = RealTwGlobal<I, SC>>
__device__ __forceinline__ void c2r_finish(float2 (&q)[bfly_plan::kPoints + 1], float* __restrict__ out,
                                           const TWS& tws, uint32_t batch, float scale, uint32_t g,
                                           float2* lds, int tid, bool adj = false) {
    using PL = PlanOf<I, SC>;
    constexpr int M = PL::N, T = PL::T, B = WG / T, R0 = PL::R0, RL = PL::RL, P = PL::P;
    static_assert(lds_len(M) >= M + 1, "staging needs M+1 bins");
    static_assert(WG % T == 0, "a work unit holds whole transforms");
    const int bl = tid / T, t = tid % T;
    const uint64_t tr = (uint64_t)g * B + bl;
    float2* L = lds + bl * lds_len(M);
#pragma unroll
    for (int i = 0; i <= P; i++) {
        const int k = t + i * T;
        if (k <= M) L[k] = q[i];
    }
    __syncthreads();
    float2 v[P];
#pragma unroll
    for (int m = 0; m < P / R0; m++)
#pragma unroll
        for (int r = 0; r < R0; r++) {
            const int k = t + m * T + r * (M / R0);
            float2 a = L[k], b = L[M - k];
            if (k == 0) {
                a.y = 0.f;
                b.y = 0.f;
                if (adj) { a.x *= 2.f; b.x *= 2.f; }
            }
            const float2 s = make_float2(a.x + b.x, a.y - b.y);
            const float2 d = make_float2(a.x - b.x, a.y + b.y);
            const float2 wd = cmul_tw<1>(d, tws.w(k));  // W_n^{-k} d
            v[m * R0 + r] = make_float2(s.x - wd.y, s.y + wd.x);
        }
    __syncthreads();
    tws.template fft<T, 1>(v, L, t);
    if (tr >= batch) return;
    float2* dst = reinterpret_cast<float2*>(out + tr * (uint64_t)(2 * M));
#pragma unroll
    for (int m = 0; m < P / RL; m++)
#pragma unroll
        for (int r = 0; r < RL; r++) {
            const float2 z = v[m * RL + r];
            st_io<(T >= 16)>(&dst[t + m * T + r * (M / RL)], make_float2(z.x * scale, z.y * scale));
        }
}

template <int I, int WG, int SC = 0>
__device__ __forceinline__ void c2r_unit(const float* __restrict__ in, float* __restrict__ out,
                                         const float2* __restrict__ tw, const float2* __restrict__ wn,
                                         uint32_t batch, float scale, uint32_t kin, uint32_t istride,
                                         uint32_t g, float2* lds, int tid) {
    float2 q[bfly_plan::kPoints + 1];
    c2r_load<I, WG, SC>(q, in, batch, kin, istride, g, tid);
    c2r_finish<I, WG, SC>(q, out, RealTwGlobal<I, SC>{tw, wn}, batch, scale, g, lds, tid);
}

}  // namespace bfly_dev
