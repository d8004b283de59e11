// bfly_mix.hip -- per-mode complex channel mixing of the 1-D spectral
// convolution on the matrix cores, in the layouts the mode-truncated R2C
// writes and the C2R reads (mode index fastest), so no transposed copy of the
// modes or of the weights is ever made.
//
//   C[m][n][k] = sum_r  opA(A)[m][r][k] * opB(B)[r][n][k],   k < K,
//
// complex64 interleaved, every operand with unit stride along the mode k and
// arbitrary strides along its two channel indices; opA / opB optionally
// conjugate.  The three products of the layer (forward, input gradient,
// weight gradient) are this one kernel with different strides.
//
// Per mode the product is a small complex GEMM, done as four real fp32
// v_mfma_f32_16x16x4_f32 products per 16x16 tile and 4-deep slice of r:
//   Cre += Are Bre,  Cre += (-Aim) Bim,  Cim += Are Bim,  Cim += Aim Bre
// (fp32 operands, fp32 accumulation).
//
// Workgroup tile: 8 modes x 16 MH rows (m) x 16 NH columns (n) of C, 256
// threads; wave w owns modes 2w, 2w+1 and all MH x NH 16x16 tiles of each.
// The reduction runs in chunks of 16 r: the chunk of A (16 MH m x 16 r x
// 8 k) and of B (16 r x 16 NH n x 8 k) is read with two modes per lane
// (16 bytes, four lanes per 64-byte run along k), prefetched into registers
// during the previous chunk's products and parked in LDS as one plane per
// (mode, 16-row half): plane[r][16] -- so the operand of one matrix
// instruction (lane l: row l % 16, r = 4s + l / 16) is 64 consecutive
// complex values.  The plane pad puts the two modes a half-wave parks in
// opposite LDS bank halves.  Results are staged through the same LDS and
// stored along the mode index (64-byte runs per (m, n)).
//
// Measured at batch 64, 128 x 128 channels, 256 modes: ~38 us per product,
// against ~29 us for the matrix instructions alone and ~25 us for the
// operand loads alone (the fp32 16x16x4 instruction retires at about half
// the fp32 matrix peak here, so the products cost about what fp32 vector
// multiply-adds would).
#include <hip/hip_runtime.h>

#include "bfly_kernels.hpp"

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kKT = 8;           // modes per workgroup
constexpr int kRT = 16;          // reduction chunk: four 4-deep matrix steps
constexpr int kThreads = 256;
constexpr int kStage = kKT + 1;  // staging pitch (complex) of one (m, n) row of modes

// LDS plane pitch (complex) of an operand with H 16-row halves: 16 r x 16
// rows plus a pad chosen so that the two modes a half-wave parks land in
// opposite LDS bank halves (2 H pitch complex = 32 banks mod 64).
template <int H>
constexpr int plane_pitch() { return 16 * kRT + 8 / H; }

// One operand chunk of H 16-row halves: lane (mm = t % 16, kp = (t / 16) %
// 4, q = t / 64) loads modes 2kp, 2kp+1 of row 16 h + mm (h = p / 16) at
// reduction index p % 16, for p = q + 4 j, j < 4H.  `lead` = the row
// stride, `red` = the r stride.  Every load is issued unconditionally (an
// element outside the operand reads the operand's first element instead)
// and `ok` records which halves are real: a load under a branch would make
// the compiler wait for it right there, serialising the chunk's loads and
// the prefetch with them.
template <int H, bool VEC>
__device__ __forceinline__ void load_chunk(float4 (&reg)[4 * H], uint32_t& ok, const float2* __restrict__ base,
                                           uint32_t rows, uint32_t R, uint32_t K, uint64_t lead, uint64_t red,
                                           uint32_t row0, uint32_t r0, uint32_t k, int t) {
    const int mm = t % 16, q = t / 64;
    ok = 0;
#pragma unroll
    for (int j = 0; j < 4 * H; j++) {
        const int p = q + 4 * j;
        const uint32_t row = row0 + 16 * (p / 16) + mm, r = r0 + p % 16;
        const bool in0 = row < rows && r < R && k < K, in1 = in0 && k + 1 < K;
        const float2* src = base + (in0 ? (uint64_t)row * lead + (uint64_t)r * red + k : 0);
        if constexpr (VEC) {
            reg[j] = *reinterpret_cast<const float4*>(src);
        } else {
            const float2 a = src[0], b = src[in1 ? 1 : 0];
            reg[j] = make_float4(a.x, a.y, b.x, b.y);
        }
        ok |= (in0 ? 1u : 0u) << (2 * j) | (in1 ? 1u : 0u) << (2 * j + 1);
    }
}

// Plane (mode, h) holds [r][16 rows].  The masking and the conjugation
// (cj = -1) are applied here, not at the load, so nothing waits on the
// prefetch before the next chunk is parked.
template <int H>
__device__ __forceinline__ void park_chunk(float2* __restrict__ lds, const float4 (&reg)[4 * H], uint32_t ok,
                                           int t, float cj) {
    constexpr int P = plane_pitch<H>();
    const int mm = t % 16, kp = (t / 16) % 4, q = t / 64;
#pragma unroll
    for (int j = 0; j < 4 * H; j++) {
        const int p = q + 4 * j, h = p / 16, r = p % 16;
        const float2 z = make_float2(0.f, 0.f);
        const float2 a = make_float2(reg[j].x, reg[j].y * cj), b = make_float2(reg[j].z, reg[j].w * cj);
        lds[(H * (2 * kp) + h) * P + 16 * r + mm] = (ok >> (2 * j)) & 1 ? a : z;          // mode 2kp
        lds[(H * (2 * kp + 1) + h) * P + 16 * r + mm] = (ok >> (2 * j + 1)) & 1 ? b : z;  // mode 2kp+1
    }
}

// MH / NH = 16-row halves of the tile along m / n (tile 16 MH x 16 NH).
template <int MH, int NH, bool VEC>
__global__ __launch_bounds__(kThreads) void k_mix(const float2* __restrict__ A, const float2* __restrict__ B,
                                                  float2* __restrict__ C, uint32_t M, uint32_t N, uint32_t R,
                                                  uint32_t K, uint64_t sam, uint64_t sar, uint64_t sbr,
                                                  uint64_t sbn, uint64_t scm, uint64_t scn, uint32_t ktiles,
                                                  uint32_t mtiles, float conj_a, float conj_b) {
    constexpr int PA = plane_pitch<MH>(), PB = plane_pitch<NH>(), TM = 16 * MH, TN = 16 * NH;
    static_assert(16 * TN * kStage <= kKT * (MH * PA + NH * PB), "a 16-row slice of the tile fits the staging area");
    __shared__ float2 lds[kKT * (MH * PA + NH * PB)];
    float2* As = lds;                   // plane (mode, m half): [r][16 m]
    float2* Bs = lds + kKT * MH * PA;   // plane (mode, n half): [r][16 n]
    const int t = threadIdx.x, lane = t % 64, w = t / 64;
    // block -> (k tile, m tile, n tile), k tile fastest
    uint32_t b = blockIdx.x;
    const uint32_t kt = b % ktiles;
    b /= ktiles;
    const uint32_t m0 = (b % mtiles) * TM, n0 = (b / mtiles) * TN, k0 = kt * kKT;
    const uint32_t kl = k0 + 2 * ((t / 16) % 4);  // first of this lane's two load modes

    float4 ra[4 * MH], rb[4 * NH];
    uint32_t oka, okb;
    f32x4 acc[2][MH][NH][2];  // [mode][m half][n half][re, im]
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int i = 0; i < MH; i++)
#pragma unroll
            for (int j = 0; j < NH; j++)
#pragma unroll
                for (int c = 0; c < 2; c++) acc[a][i][j][c] = f32x4{0.f, 0.f, 0.f, 0.f};

    load_chunk<MH, VEC>(ra, oka, A, M, R, K, sam, sar, m0, 0, kl, t);
    load_chunk<NH, VEC>(rb, okb, B, N, R, K, sbn, sbr, n0, 0, kl, t);
    for (uint32_t r0 = 0; r0 < R; r0 += kRT) {
        park_chunk<MH>(As, ra, oka, t, conj_a);
        park_chunk<NH>(Bs, rb, okb, t, conj_b);
        __syncthreads();
        if (r0 + kRT < R) {
            load_chunk<MH, VEC>(ra, oka, A, M, R, K, sam, sar, m0, r0 + kRT, kl, t);
            load_chunk<NH, VEC>(rb, okb, B, N, R, K, sbn, sbr, n0, r0 + kRT, kl, t);
        }
#pragma unroll
        for (int a = 0; a < 2; a++) {
            const int mode = 2 * w + a;
#pragma unroll
            for (int s = 0; s < kRT / 4; s++) {
                float2 x[MH], y[NH];
#pragma unroll
                for (int h = 0; h < MH; h++) x[h] = As[(MH * mode + h) * PA + 64 * s + lane];
#pragma unroll
                for (int h = 0; h < NH; h++) y[h] = Bs[(NH * mode + h) * PB + 64 * s + lane];
                // independent accumulators between the two halves of the
                // complex product, so consecutive matrix instructions never
                // wait on each other
#pragma unroll
                for (int i = 0; i < MH; i++)
#pragma unroll
                    for (int j = 0; j < NH; j++) {
                        acc[a][i][j][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[i].x, y[j].x, acc[a][i][j][0], 0, 0, 0);
                        acc[a][i][j][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[i].x, y[j].y, acc[a][i][j][1], 0, 0, 0);
                    }
#pragma unroll
                for (int i = 0; i < MH; i++)
#pragma unroll
                    for (int j = 0; j < NH; j++) {
                        acc[a][i][j][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(-x[i].y, y[j].y, acc[a][i][j][0], 0, 0, 0);
                        acc[a][i][j][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[i].y, y[j].x, acc[a][i][j][1], 0, 0, 0);
                    }
            }
        }
        __syncthreads();
    }
    // Results leave through LDS, one 16-row slice of the tile at a time, so
    // that the stores run along the mode index (64-byte runs) instead of one
    // complex value per lane.  Staging layout [m][n][mode], n pitch 9.
    // Lane l holds rows 4 (l / 16) + v (v < 4) and column l % 16 of each
    // 16x16 tile.
    constexpr int RPP = kThreads / (4 * TN);  // rows stored per pass
    const int kp = t % 4, sn = (t / 4) % TN, sm = t / (4 * TN);
    const uint32_t ks = k0 + 2 * kp, n = n0 + sn;
#pragma unroll
    for (int i = 0; i < MH; i++) {
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int j = 0; j < NH; j++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    lds[((4 * (lane / 16) + v) * TN + 16 * j + lane % 16) * kStage + 2 * w + a] =
                        make_float2(acc[a][i][j][0][v], acc[a][i][j][1][v]);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 16 / RPP; e++) {
            const int ml = sm + RPP * e;
            const uint32_t m = m0 + 16 * i + ml;
            const float2 c0 = lds[(ml * TN + sn) * kStage + 2 * kp];
            const float2 c1 = lds[(ml * TN + sn) * kStage + 2 * kp + 1];
            if (m < M && n < N && ks < K) {
                float2* dst = C + (uint64_t)m * scm + (uint64_t)n * scn + ks;
                if constexpr (VEC) {
                    *reinterpret_cast<float4*>(dst) = make_float4(c0.x, c0.y, c1.x, c1.y);
                } else {
                    dst[0] = c0;
                    if (ks + 1 < K) dst[1] = c1;
                }
            }
        }
        __syncthreads();
    }
}

template <int MH, int NH>
hipError_t launch_mix(const float2* A, const float2* B, float2* C, uint32_t M, uint32_t N, uint32_t R, uint32_t K,
                      uint64_t sam, uint64_t sar, uint64_t sbr, uint64_t sbn, uint64_t scm, uint64_t scn,
                      float ca, float cb, bool vec, hipStream_t s) {
    const uint32_t kt = (K + kKT - 1) / kKT, mt = (M + 16 * MH - 1) / (16 * MH), nt = (N + 16 * NH - 1) / (16 * NH);
    const uint64_t blocks = (uint64_t)kt * mt * nt;
    if (blocks == 0 || blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    if (vec)
        hipLaunchKernelGGL((k_mix<MH, NH, true>), dim3((unsigned)blocks), dim3(kThreads), 0, s, A, B, C, M, N, R, K,
                           sam, sar, sbr, sbn, scm, scn, kt, mt, ca, cb);
    else
        hipLaunchKernelGGL((k_mix<MH, NH, false>), dim3((unsigned)blocks), dim3(kThreads), 0, s, A, B, C, M, N, R,
                           K, sam, sar, sbr, sbn, scm, scn, kt, mt, ca, cb);
    return hipGetLastError();
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
    // two modes per 16-byte access: needs an even mode count and 16-byte aligned operands
    const bool vec = K % 2 == 0 && (reinterpret_cast<uintptr_t>(a) % 16) == 0 &&
                     (reinterpret_cast<uintptr_t>(b) % 16) == 0 && (reinterpret_cast<uintptr_t>(c) % 16) == 0;
    const auto* A = reinterpret_cast<const float2*>(a);
    const auto* B = reinterpret_cast<const float2*>(b);
    auto* C = reinterpret_cast<float2*>(c);
    // tile shapes measured at batch 64, 128 x 128 channels, 256 modes: 16 x 32
    // for the forward and input-gradient products (38-39 us vs 40-43 us with
    // 32 x 32), 32 x 32 for the weight gradient (36-37 us vs 42-43 us)
    if (op == 2)
        return launch_mix<2, 2>(A, B, C, M, N, R, (uint32_t)K, sam, sar, sbr, sbn, scm, scn, ca, cb, vec, s);
    return launch_mix<1, 2>(A, B, C, M, N, R, (uint32_t)K, sam, sar, sbr, sbn, scm, scn, ca, cb, vec, s);
}

}  // namespace bfly_kern
