// bfly_device.hpp -- device building blocks for the gfx950 FFT kernels.
//
// Complex values are float2 (x = re, y = im).  DIR is the exponent sign of the
// transform: -1 forward (X[k] = sum x[j] e^{-2 pi i jk/n}), +1 unnormalised
// inverse.  Every twiddle table holds forward values W = e^{-2 pi i m/L}; the
// inverse conjugates them on use, so one table serves both directions.
//
// The Stockham core keeps one transform's working set in LDS between radix
// passes and in registers inside a pass: each thread owns P points, a pass of
// radix R runs P/R butterflies per thread, and only the exchanges between
// passes touch LDS.  The first pass reads its operands straight from the
// caller (global memory or a staging buffer) and the last pass leaves its
// results in registers in natural order, so a plain C2C is one HBM read and
// one HBM write per element.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bfly_plans.hpp"

namespace bfly_dev {

// ------------------------------------------------------- streaming access
// Transform data is touched once per launch: non-temporal loads and stores
// keep it from displacing the twiddle tables in L2 and let the memory
// controller stream it (measured on MI355X: a 4-deep float4 copy goes from
// 5.6 to 6.2 TB/s with non-temporal access).
typedef float bfly_v2f __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ld_stream(const float2* p) {
    const bfly_v2f v = __builtin_nontemporal_load(reinterpret_cast<const bfly_v2f*>(p));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ void st_stream(float2* p, float2 x) {
    bfly_v2f v;
    v.x = x.x;
    v.y = x.y;
    __builtin_nontemporal_store(v, reinterpret_cast<bfly_v2f*>(p));
}

// Streaming access only pays when a row segment covers whole 128-byte lines
// (16 consecutive complex values or more); shorter segments rely on L2
// write-combining, which non-temporal stores bypass.
template <bool NT>
__device__ __forceinline__ float2 ld_io(const float2* p) {
    if constexpr (NT) return ld_stream(p);
    else return *p;
}
template <bool NT>
__device__ __forceinline__ void st_io(float2* p, float2 x) {
    if constexpr (NT) st_stream(p, x);
    else *p = x;
}

// ------------------------------------------------------------ complex ops
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// a * w for DIR = -1, a * conj(w) for DIR = +1 (w is a forward twiddle).
template <int DIR>
__device__ __forceinline__ float2 cmul_tw(float2 a, float2 w) {
    if constexpr (DIR < 0)
        return make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
    else
        return make_float2(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y);
}

// Multiply by e^{DIR * 2 pi i m / 16}.  m is a compile-time constant after
// unrolling, so every branch folds away; the quarter turns cost no flops and
// the eighth turns two multiplies.
__device__ __forceinline__ float cos16(int m) {
    switch (m & 15) {
        case 0: return 1.0f;
        case 1: case 15: return 0.92387953251128675613f;
        case 2: case 14: return 0.70710678118654752440f;
        case 3: case 13: return 0.38268343236508977173f;
        case 4: case 12: return 0.0f;
        case 5: case 11: return -0.38268343236508977173f;
        case 6: case 10: return -0.70710678118654752440f;
        case 7: case 9: return -0.92387953251128675613f;
        default: return -1.0f;
    }
}
__device__ __forceinline__ float sin16(int m) { return cos16(m - 4); }

template <int DIR>
__device__ __forceinline__ float2 rot16(float2 a, int m) {
    m &= 15;
    if (m == 0) return a;
    if (m == 8) return make_float2(-a.x, -a.y);
    // e^{DIR i pi/2} = DIR*i: (x + iy) * (DIR i) = -DIR y + i DIR x
    if (m == 4) return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
    if (m == 12) return DIR < 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
    const float c = cos16(m), s = DIR < 0 ? -sin16(m) : sin16(m);
    if ((m & 3) == 2) {
        // |c| == |s| == sqrt(1/2): two adds then two multiplies
        const float h = c;  // +-sqrt(1/2)
        const float g = s / c;  // +-1 exactly
        return make_float2(h * (a.x - g * a.y), h * (g * a.x + a.y));
    }
    return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
}

// ---------------------------------------------------------- small DFTs
// In-place DFT of R points held in registers, natural order in and out.
template <int R, int DIR>
__device__ __forceinline__ void dft(float2* v);

template <>
__device__ __forceinline__ void dft<1, -1>(float2*) {}
template <>
__device__ __forceinline__ void dft<1, 1>(float2*) {}

template <int DIR>
__device__ __forceinline__ void dft2(float2* v) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}
template <>
__device__ __forceinline__ void dft<2, -1>(float2* v) { dft2<-1>(v); }
template <>
__device__ __forceinline__ void dft<2, 1>(float2* v) { dft2<1>(v); }

template <int DIR>
__device__ __forceinline__ void dft4(float2* v) {
    const float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const float2 t2 = cadd(v[1], v[3]);
    const float2 d = csub(v[1], v[3]);
    const float2 t3 = DIR < 0 ? make_float2(d.y, -d.x) : make_float2(-d.y, d.x);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
}
template <>
__device__ __forceinline__ void dft<4, -1>(float2* v) { dft4<-1>(v); }
template <>
__device__ __forceinline__ void dft<4, 1>(float2* v) { dft4<1>(v); }

// R = A * B with A = 4: n = B n1 + n2, k = k1 + A k2.
//   Y[k1][n2] = W_R^{n2 k1} * DFT_A over n1 of x[B n1 + n2]
//   X[k1 + A k2] = DFT_B over n2 of Y[k1][n2]
template <int R, int DIR>
__device__ __forceinline__ void dft_split(float2* v) {
    constexpr int A = 4, B = R / 4;
    float2 t[R];
#pragma unroll
    for (int n2 = 0; n2 < B; n2++) {
        float2 c[A];
#pragma unroll
        for (int n1 = 0; n1 < A; n1++) c[n1] = v[B * n1 + n2];
        dft4<DIR>(c);
#pragma unroll
        for (int k1 = 0; k1 < A; k1++) t[B * k1 + n2] = rot16<DIR>(c[k1], (n2 * k1) * (16 / R));
    }
#pragma unroll
    for (int k1 = 0; k1 < A; k1++) {
        float2 c[B];
#pragma unroll
        for (int n2 = 0; n2 < B; n2++) c[n2] = t[B * k1 + n2];
        dft<B, DIR>(c);
#pragma unroll
        for (int k2 = 0; k2 < B; k2++) v[k1 + A * k2] = c[k2];
    }
}
template <>
__device__ __forceinline__ void dft<8, -1>(float2* v) { dft_split<8, -1>(v); }
template <>
__device__ __forceinline__ void dft<8, 1>(float2* v) { dft_split<8, 1>(v); }
template <>
__device__ __forceinline__ void dft<16, -1>(float2* v) { dft_split<16, -1>(v); }
template <>
__device__ __forceinline__ void dft<16, 1>(float2* v) { dft_split<16, 1>(v); }

// ------------------------------------------------------------ LDS layout
// One pad slot every 2^PAD_SHIFT complex values breaks the power-of-two
// strides of the Stockham writes across LDS banks.
constexpr int PAD_SHIFT = 4;
__device__ __forceinline__ int lds_idx(int i) { return i + (i >> PAD_SHIFT); }
__host__ __device__ constexpr int lds_len(int n) { return n + (n >> PAD_SHIFT); }

// Two exchange layouts.  LX_PAD: region pointer + padded local index (one
// slot per 16 complex).  LX_SWZ: no padding; complex index i of the whole
// workgroup buffer lives at i ^ ((i >> 5) & 31), a bijection that spreads
// the stride-16 and stride-256 Stockham writes over all 64 banks (each
// 8-byte slot owns two of them) while consecutive reads stay conflict-free.
// Without padding a 256-thread workgroup needs exactly 32 KiB, so five fit
// the 160 KiB of LDS of a CU instead of four.
constexpr int LX_PAD = 0;
constexpr int LX_SWZ = 1;
template <int LX>
__device__ __forceinline__ int lx(int rb, int i) {
    if constexpr (LX == LX_SWZ) {
        const int g = rb + i;
        return g ^ ((g >> 5) & 31);
    } else {
        return lds_idx(i);
    }
}
template <int LX>
__host__ __device__ constexpr int lx_len(int n) { return LX == LX_SWZ ? n : lds_len(n); }

// Address of element r of a Stockham write group (base + r*Ns) and of a read
// group (u + r*S).  In the padded layout, when every offset r*Ns (Ns >= 16)
// or r*S (S >= 16) is a multiple of 16, pad(x + c) = pad(x) + c*17/16 and the
// whole group is one per-thread base plus compile-time offsets (LDS
// immediates); for Ns == 1 the group never crosses a 16-slot boundary
// (R | 16), so pad(base + r) = pad(base) + r.  Other strides (a small radix
// ahead of radix 16) pad every address on its own.
template <int Q>
constexpr bool pad_aligned() { return Q % (1 << PAD_SHIFT) == 0; }
template <int LX, int Ns>
__device__ __forceinline__ int lx_wbase(int rb, int base) {
    if constexpr (LX == LX_PAD) return (Ns == 1 || pad_aligned<Ns>()) ? lds_idx(base) : base;
    else return rb + base;
}
template <int LX, int Ns>
__device__ __forceinline__ int lx_wat(int wb, int r) {
    if constexpr (LX == LX_PAD) {
        if constexpr (Ns == 1) return wb + r;
        else if constexpr (pad_aligned<Ns>()) return wb + r * Ns + (r * Ns >> PAD_SHIFT);
        else return lds_idx(wb + r * Ns);
    } else {
        const int g = wb + r * Ns;
        return g ^ ((g >> 5) & 31);
    }
}
template <int LX, int S>
__device__ __forceinline__ int lx_rat(int rbase, int r) {
    if constexpr (LX == LX_PAD) {
        if constexpr (pad_aligned<S>()) return rbase + r * S + (r * S >> PAD_SHIFT);
        else return lds_idx(rbase + r * S);
    } else {
        const int g = rbase + r * S;
        return g ^ ((g >> 5) & 31);
    }
}
template <int LX, int S>
__device__ __forceinline__ int lx_rbase(int rb, int u) {
    if constexpr (LX == LX_PAD) return pad_aligned<S>() ? lds_idx(u) : u;
    else return rb + u;
}

// ------------------------------------------------------- Stockham passes
// Radix plans: the list of radices of one length, first pass first.
template <int... R>
struct Radices {};

template <int... R>
struct RadixProduct;
template <>
struct RadixProduct<> { static constexpr int value = 1; };
template <int R0, int... R>
struct RadixProduct<R0, R...> { static constexpr int value = R0 * RadixProduct<R...>::value; };

// Twiddle-table length (complex) of the passes after the first: pass s reads
// W_{Ns R_s}^{r k} at [k * R_s + r] for k < Ns, r < R_s.
template <int Ns, int... R>
struct TwiddleLen;
template <int Ns>
struct TwiddleLen<Ns> { static constexpr int value = 0; };
template <int Ns, int R0, int... R>
struct TwiddleLen<Ns, R0, R...> {
    static constexpr int value = (Ns > 1 ? Ns * R0 : 0) + TwiddleLen<Ns * R0, R...>::value;
};

template <int R0, int... R>
struct First { static constexpr int value = R0; };

// Runs the passes of radix list <R, Rest...> with Ns points already combined.
// Thread t of a transform owns butterflies j = t + m*T (m < P/R) of each
// pass.  On entry v[m*R + r] = x[t + m*T + r*N/R] (the first pass's operands);
// on exit v[m*RL + r] = X[t + m*T + r*N/RL] for the last radix RL.
// `lds` is this transform's region (lds_len(N) complex); `tw` the table of
// the current pass (unused for the first).
template <int N, int P, int T, int DIR, int LX, int Ns, int R, int... Rest>
__device__ __forceinline__ void stockham(float2 (&v)[P], float2* lds, int rb,
                                         const float2* __restrict__ tw, int t) {
    static_assert(P % R == 0, "radix must divide the points per thread");
    constexpr int M = P / R;
    if constexpr (Ns > 1) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int k = (t + m * T) & (Ns - 1);
            const float4* w4 = reinterpret_cast<const float4*>(tw + k * R);
#pragma unroll
            for (int q = 0; q < R / 2; q++) {
                const float4 w = w4[q];
                if (q > 0) v[m * R + 2 * q] = cmul_tw<DIR>(v[m * R + 2 * q], make_float2(w.x, w.y));
                v[m * R + 2 * q + 1] = cmul_tw<DIR>(v[m * R + 2 * q + 1], make_float2(w.z, w.w));
            }
        }
    }
#pragma unroll
    for (int m = 0; m < M; m++) dft<R, DIR>(&v[m * R]);
    if constexpr (sizeof...(Rest) > 0) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int j = t + m * T;
            const int k = j & (Ns - 1);
            const int wb = lx_wbase<LX, Ns>(rb, (j - k) * R + k);
#pragma unroll
            for (int r = 0; r < R; r++) lds[lx_wat<LX, Ns>(wb, r)] = v[m * R + r];
        }
        __syncthreads();
        constexpr int R2 = First<Rest...>::value;
        constexpr int M2 = P / R2;
#pragma unroll
        for (int m = 0; m < M2; m++) {
            const int rbase = lx_rbase<LX, N / R2>(rb, t + m * T);
#pragma unroll
            for (int r = 0; r < R2; r++) v[m * R2 + r] = lds[lx_rat<LX, N / R2>(rbase, r)];
        }
        __syncthreads();
        stockham<N, P, T, DIR, LX, Ns * R, Rest...>(v, lds, rb, tw + (Ns > 1 ? Ns * R : 0), t);
    }
}

// ------------------------------------------- register-resident twiddles
// The twiddles a thread multiplies by depend only on its index t inside the
// transform (k = (t + m T) mod Ns), never on which transform it works on.
// A thread that processes a stream of transforms loads them once into
// registers, so the passes issue no global loads at all -- which lets the
// next transform's input stay in flight across the whole compute phase.
template <int P, int Ns, int... R>
struct TwRegCount;
template <int P, int Ns>
struct TwRegCount<P, Ns> { static constexpr int value = 0; };
template <int P, int Ns, int R0, int... R>
struct TwRegCount<P, Ns, R0, R...> {
    static constexpr int value = (Ns > 1 ? (P / R0) * (R0 - 1) : 0) + TwRegCount<P, Ns * R0, R...>::value;
};

// Fill tr[] with every twiddle of passes s >= 1 in pass order, [m][r-1].
template <int N, int P, int T, int Ns, int R, int... Rest>
__device__ __forceinline__ void load_twiddles(float2* tr, const float2* __restrict__ tw, int t) {
    constexpr int M = P / R;
    if constexpr (Ns > 1) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int k = (t + m * T) & (Ns - 1);
#pragma unroll
            for (int r = 1; r < R; r++) tr[m * (R - 1) + r - 1] = tw[k * R + r];
        }
    }
    if constexpr (sizeof...(Rest) > 0)
        load_twiddles<N, P, T, Ns * R, Rest...>(tr + (Ns > 1 ? M * (R - 1) : 0),
                                                tw + (Ns > 1 ? Ns * R : 0), t);
}

// Same passes as stockham(), twiddles taken from registers.
template <int N, int P, int T, int DIR, int LX, int Ns, int R, int... Rest>
__device__ __forceinline__ void stockham_rt(float2 (&v)[P], float2* lds, int rb, const float2* tr,
                                            int t) {
    constexpr int M = P / R;
    if constexpr (Ns > 1) {
#pragma unroll
        for (int m = 0; m < M; m++)
#pragma unroll
            for (int r = 1; r < R; r++)
                v[m * R + r] = cmul_tw<DIR>(v[m * R + r], tr[m * (R - 1) + r - 1]);
    }
#pragma unroll
    for (int m = 0; m < M; m++) dft<R, DIR>(&v[m * R]);
    if constexpr (sizeof...(Rest) > 0) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int j = t + m * T;
            const int k = j & (Ns - 1);
            const int wb = lx_wbase<LX, Ns>(rb, (j - k) * R + k);
#pragma unroll
            for (int r = 0; r < R; r++) lds[lx_wat<LX, Ns>(wb, r)] = v[m * R + r];
        }
        __syncthreads();
        constexpr int R2 = First<Rest...>::value;
        constexpr int M2 = P / R2;
#pragma unroll
        for (int m = 0; m < M2; m++) {
            const int rbase = lx_rbase<LX, N / R2>(rb, t + m * T);
#pragma unroll
            for (int r = 0; r < R2; r++) v[m * R2 + r] = lds[lx_rat<LX, N / R2>(rbase, r)];
        }
        __syncthreads();
        stockham_rt<N, P, T, DIR, LX, Ns * R, Rest...>(v, lds, rb, tr + (Ns > 1 ? M * (R - 1) : 0), t);
    }
}

// ------------------------------------------- two-level LDS twiddles
// W_N^m = A[m mod 64] * B[m / 64] with A[j] = W_N^j (j < 64) and
// B[j] = W_N^{64 j} (j < N/64), both FP64-generated and held in LDS (at most
// 128 complex).  One extra complex product per twiddle replaces the
// register-resident twiddle file of stockham_rt, which frees the registers
// for more resident workgroups; the product adds about one ulp.
constexpr int TW2_LO = 64;
template <int N>
__device__ __forceinline__ float2 tw2(const float2* tlo, const float2* thi, int m) {
    if constexpr (N <= TW2_LO) {
        return tlo[m];
    } else {
        const float2 a = tlo[m & (TW2_LO - 1)], b = thi[m / TW2_LO];
        return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
    }
}

// The R-1 twiddles W^{r e} of one butterfly: W^e, W^{2e}, W^{4e}, W^{8e}
// from the two-level tables, every other power as the product of its
// highest power-of-two factor and the rest (w3 = w2 w1, ..., w15 = w8 w7:
// at most three products deep).  A radix-16 butterfly then reads 8 LDS
// values instead of 30 for the same complex products; the added rounding
// is a few ulp per twiddle (transform error 1.27e-7 -> 1.40e-7 at 4096).
template <int N, int R>
__device__ __forceinline__ void tw_powers(float2 (&w)[R], const float2* tlo, const float2* thi, int e) {
#pragma unroll
    for (int r = 1; r < R; r *= 2) w[r] = tw2<N>(tlo, thi, r * e);
#pragma unroll
    for (int r = 3; r < R; r++) {
        if ((r & (r - 1)) == 0) continue;
        int hb = 1;
        while (hb * 2 <= r) hb *= 2;
        const float2 a = w[hb], b = w[r - hb];
        w[r] = make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
    }
}

// Twiddles of the passes after the second built by tw_powers (fewer LDS reads
// in the LDS-latency-bound one-workgroup 16384 kernel: 65.8 -> 61.9 us at
// batch 1024; 47.0 -> 46.6 us at 4096 x 4096), except at 8192, where taking
// each twiddle from the tables measured faster (57.5 vs 59.2 us at batch 2048).
template <int N>
constexpr bool kTwPowers = N != 8192;

// `t1` holds the second pass's own table in LDS, transposed to [r Ns + k]
// (Ns = the first radix) so that the lanes of a wave, whose k are
// consecutive, read consecutive banks: one read per twiddle there; later
// passes combine the two-level tables.
template <int N, int P, int T, int DIR, int LX, int Ns, int R, int... Rest>
__device__ __forceinline__ void stockham_lt(float2 (&v)[P], float2* lds, int rb, const float2* tlo,
                                            const float2* thi, const float2* t1, int t) {
    constexpr int M = P / R;
    if constexpr (Ns > 1) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int k = (t + m * T) & (Ns - 1);
            if constexpr (Ns == bfly_plan::kPoints) {
#pragma unroll
                for (int r = 1; r < R; r++) v[m * R + r] = cmul_tw<DIR>(v[m * R + r], t1[r * Ns + k]);
            } else if constexpr (kTwPowers<N>) {
                float2 w[R];
                tw_powers<N, R>(w, tlo, thi, k * (N / (Ns * R)));
#pragma unroll
                for (int r = 1; r < R; r++) v[m * R + r] = cmul_tw<DIR>(v[m * R + r], w[r]);
            } else {
#pragma unroll
                for (int r = 1; r < R; r++)
                    v[m * R + r] = cmul_tw<DIR>(v[m * R + r], tw2<N>(tlo, thi, r * k * (N / (Ns * R))));
            }
        }
    }
#pragma unroll
    for (int m = 0; m < M; m++) dft<R, DIR>(&v[m * R]);
    if constexpr (sizeof...(Rest) > 0) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int j = t + m * T;
            const int k = j & (Ns - 1);
            const int wb = lx_wbase<LX, Ns>(rb, (j - k) * R + k);
#pragma unroll
            for (int r = 0; r < R; r++) lds[lx_wat<LX, Ns>(wb, r)] = v[m * R + r];
        }
        __syncthreads();
        constexpr int R2 = First<Rest...>::value;
        constexpr int M2 = P / R2;
#pragma unroll
        for (int m = 0; m < M2; m++) {
            const int rbase = lx_rbase<LX, N / R2>(rb, t + m * T);
#pragma unroll
            for (int r = 0; r < R2; r++) v[m * R2 + r] = lds[lx_rat<LX, N / R2>(rbase, r)];
        }
        __syncthreads();
        stockham_lt<N, P, T, DIR, LX, Ns * R, Rest...>(v, lds, rb, tlo, thi, t1, t);
    }
}

}  // namespace bfly_dev
