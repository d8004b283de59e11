// bfly_tables.cpp -- host-side shape rules, table layout and table generation
// of the C ABI (include/bfly_hip.h).  Pure host C++: linked into
// libbfly_hip.so and into the CPU stub library, so both agree byte for byte.
//
// Twiddles are evaluated in FP64 from the exact rational angle 2*pi*m/L
// (integer m reduced modulo L first) and rounded once to fp32.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>

#include "bfly_hip.h"
#include "bfly_plans.hpp"
#include "bfly_tables.hpp"

namespace bfly_host {

static inline bool pow2(uint32_t n) { return n && !(n & (n - 1)); }

bool shape_ok(int kind, uint32_t n) {
    if (!pow2(n)) return false;
    switch (kind) {
        case BFLY_KIND_C2C:
            if (n <= (uint32_t)bfly_plan::kMaxLocalN) return bfly_plan::plan_index(n) >= 0;
            return bfly_plan::long_split(n, nullptr, nullptr) && n <= 65536u;
        case BFLY_KIND_R2C:
        case BFLY_KIND_C2R:
            return n >= 32u && n <= 2u * (uint32_t)bfly_plan::kMaxLocalN;
        default:
            return false;
    }
}

// FP64 e^{-2 pi i m / L} rounded to fp32.
void twiddle(uint64_t m, uint64_t L, float* out) {
    m %= L;
    const double ang = -2.0 * M_PI * (double)m / (double)L;
    out[0] = (float)std::cos(ang);
    out[1] = (float)std::sin(ang);
}

// Inter-pass table of local plan i under radix schedule `sched`, complex
// entries at [k * R_s + r].
size_t fill_stage(int i, float* dst, int sched) {
    const auto& p = bfly_plan::kPlans[i];
    size_t off = 0;
    uint64_t ns = 1;
    for (int s = 0; s < p.nrad; s++) {
        const uint64_t R = (uint64_t)bfly_plan::radix_at(i, sched, s);
        if (s > 0) {
            for (uint64_t k = 0; k < ns; k++)
                for (uint64_t r = 0; r < R; r++) {
                    if (dst) twiddle(r * k, ns * R, dst + 2 * (off + k * R + r));
                }
            off += ns * R;
        }
        ns *= R;
    }
    return off;
}

TableLayout table_layout(int kind, uint32_t n) {
    TableLayout t{};
    if (!shape_ok(kind, n)) return t;
    if (kind == BFLY_KIND_C2C && n <= (uint32_t)bfly_plan::kMaxLocalN) {
        t.stage0 = 0;
        // [stage table][two-level table: W_n^j, j < 64 | W_n^{64 j}, j < n/64]
        t.extra = bfly_plan::stage_twiddle_len(bfly_plan::plan_index(n));
        const size_t lo = n < 64 ? n : 64;
        t.complex_len = t.extra + lo + n / lo;
    } else if (kind == BFLY_KIND_C2C) {
        uint32_t n1 = 0, n2 = 0;
        bfly_plan::long_split(n, &n1, &n2);
        t.stage0 = 0;
        t.stage1 = bfly_plan::stage_twiddle_len(bfly_plan::plan_index(n1));
        t.extra = t.stage1 + bfly_plan::stage_twiddle_len(bfly_plan::plan_index(n2));
        t.complex_len = t.extra + n;  // W_N^m, m < N
        if (bfly_plan::plan_index(n) >= 0) {
            t.single = t.complex_len;
            t.complex_len += bfly_plan::stage_twiddle_len(bfly_plan::plan_index(n));
        }
    } else {
        const uint32_t m = n / 2;
        t.stage0 = 0;
        t.extra = bfly_plan::stage_twiddle_len(bfly_plan::plan_index(m));
        t.complex_len = t.extra + m + 2;  // W_n^k, k = 0..m, padded to even
    }
    return t;
}

bool fft2d_ok(uint32_t n1, uint32_t n2) {
    return pow2(n1) && pow2(n2) && n1 >= 16u && n2 >= 16u && n1 <= (uint32_t)bfly_plan::kMaxLocalN &&
           n2 <= (uint32_t)bfly_plan::kMaxLocalN;
}

Fft2dLayout fft2d_layout(uint32_t n1) {
    Fft2dLayout t{};
    if (!pow2(n1) || n1 < 16u || n1 > (uint32_t)bfly_plan::kMaxLocalN) return t;
    if (n1 <= 2048u) {
        t.p = n1;
        t.q = 1;
    } else {
        t.p = 64;
        t.q = n1 / t.p;
    }
    t.p_table = 0;
    t.q_table = table_layout(BFLY_KIND_C2C, t.p).complex_len;
    t.wab = t.q_table + (t.q > 1 ? table_layout(BFLY_KIND_C2C, t.q).complex_len : 0);
    t.complex_len = t.wab + (t.q > 1 ? 128 : 0);
    return t;
}

int fft2d_fill(uint32_t n1, float* dst) {
    const Fft2dLayout t = fft2d_layout(n1);
    if (!dst || !t.complex_len) return BFLY_ERR_SHAPE;
    for (size_t i = 0; i < 2 * t.complex_len; i++) dst[i] = 0.f;
    int rc = fill_table(BFLY_KIND_C2C, t.p, dst + 2 * t.p_table);
    if (rc == BFLY_OK && t.q > 1) rc = fill_table(BFLY_KIND_C2C, t.q, dst + 2 * t.q_table);
    if (rc != BFLY_OK || t.q == 1) return rc;
    for (uint64_t j = 0; j < 64; j++) twiddle(j, n1, dst + 2 * (t.wab + j));
    for (uint64_t j = 0; j < n1 / 64; j++) twiddle(64 * j, n1, dst + 2 * (t.wab + 64 + j));
    return BFLY_OK;
}

int check_columns(int sign, const void* in, const void* out, uint32_t n1, uint32_t ncols, uint32_t batch,
                  const void* table_cols, const void* work) {
    const Fft2dLayout L = fft2d_layout(n1);
    if (!L.complex_len || ncols == 0 || batch == 0 || !in || !out) return BFLY_ERR_SHAPE;
    if (sign != -1 && sign != 1) return BFLY_ERR_SHAPE;
    if (!table_cols) return BFLY_ERR_TABLE;
    if (L.q > 1 && !work) return BFLY_ERR_SHAPE;
    const uint64_t item = (uint64_t)n1 * ncols;
    // widest column grid: batch * q * ceil(p * ncols / 16) workgroups at most
    if ((uint64_t)batch * item / 16 + (uint64_t)batch * L.p > 0x7fffffffull) return BFLY_ERR_SHAPE;
    const uint64_t bytes = 8 * item * batch;
    const auto lo = [](const void* p) { return reinterpret_cast<uintptr_t>(p); };
    const auto apart = [&](const void* a, const void* b) { return lo(a) + bytes <= lo(b) || lo(b) + bytes <= lo(a); };
    if (!apart(in, out)) return BFLY_ERR_SHAPE;
    if (L.q > 1 && (!apart(in, work) || !apart(out, work))) return BFLY_ERR_SHAPE;
    return BFLY_OK;
}

static bool grid_ok(int kind, uint32_t n, uint32_t batch);

int check_rows(int sign, const void* in, uint64_t idist, const void* out, uint64_t odist, uint32_t n,
               uint32_t batch, const void* table, const void* work) {
    if (!in || !out || batch == 0 || !shape_ok(BFLY_KIND_C2C, n) || !grid_ok(BFLY_KIND_C2C, n, batch))
        return BFLY_ERR_SHAPE;
    const bool four_step = n > (uint32_t)bfly_plan::kMaxSingleN;
    if (four_step && !work) return BFLY_ERR_SHAPE;
    if (sign != -1 && sign != 1) return BFLY_ERR_SHAPE;
    if (idist < n || odist < n) return BFLY_ERR_SHAPE;
    const auto lo = [](const void* p) { return reinterpret_cast<uintptr_t>(p); };
    if ((lo(in) | lo(out)) % 8) return BFLY_ERR_SHAPE;  // float2 rows
    // extents in complex values, kept below 2^58 so byte offsets fit
    constexpr uint64_t kMaxExtent = 1ull << 58;
    const uint64_t b1 = batch - 1ull;
    if (b1 && (idist > (kMaxExtent - n) / b1 || odist > (kMaxExtent - n) / b1)) return BFLY_ERR_SHAPE;
    if (!table) return BFLY_ERR_TABLE;
    const uint64_t in_bytes = 8 * (b1 * idist + n), out_bytes = 8 * (b1 * odist + n);
    const auto apart = [&](const void* p, uint64_t pb, const void* q, uint64_t qb) {
        return lo(p) + pb <= lo(q) || lo(q) + qb <= lo(p);
    };
    if (!apart(in, in_bytes, out, out_bytes)) return BFLY_ERR_SHAPE;
    const uint64_t work_bytes = 8ull * n * batch;  // contiguous [batch][n]
    if (four_step && (!apart(work, work_bytes, in, in_bytes) || !apart(work, work_bytes, out, out_bytes)))
        return BFLY_ERR_SHAPE;
    return BFLY_OK;
}

int check_fft2d(int sign, const void* in, const void* out, uint32_t n1, uint32_t n2, uint32_t batch,
                const void* table_rows, const void* table_cols, const void* work) {
    if (!fft2d_ok(n1, n2) || batch == 0 || !in || !out || !work) return BFLY_ERR_SHAPE;
    if (sign != -1 && sign != 1) return BFLY_ERR_SHAPE;
    if (!table_rows || !table_cols) return BFLY_ERR_TABLE;
    const uint64_t item = (uint64_t)n1 * n2;
    // row pass: batch * n1 transforms in one launch; column pass grids
    if ((uint64_t)batch * n1 > 0xffffffffull || (uint64_t)batch * item / 16 > 0x7fffffffull)
        return BFLY_ERR_SHAPE;
    const uint64_t bytes = 8 * item * batch;
    const auto lo = [](const void* p) { return reinterpret_cast<uintptr_t>(p); };
    const auto apart = [&](const void* a, const void* b) { return lo(a) + bytes <= lo(b) || lo(b) + bytes <= lo(a); };
    if (!apart(in, out) || !apart(in, work) || !apart(out, work)) return BFLY_ERR_SHAPE;
    return BFLY_OK;
}

int fill_table(int kind, uint32_t n, float* dst) {
    if (!dst || !shape_ok(kind, n)) return BFLY_ERR_SHAPE;
    const TableLayout t = table_layout(kind, n);
    for (size_t i = 0; i < 2 * t.complex_len; i++) dst[i] = 0.f;
    if (kind == BFLY_KIND_C2C && n <= (uint32_t)bfly_plan::kMaxLocalN) {
        fill_stage(bfly_plan::plan_index(n), dst);
        const uint64_t lo = n < 64 ? n : 64;
        for (uint64_t j = 0; j < lo; j++) twiddle(j, n, dst + 2 * (t.extra + j));
        for (uint64_t j = 0; j < n / lo; j++) twiddle(lo * j, n, dst + 2 * (t.extra + lo + j));
    } else if (kind == BFLY_KIND_C2C) {
        uint32_t n1 = 0, n2 = 0;
        bfly_plan::long_split(n, &n1, &n2);
        fill_stage(bfly_plan::plan_index(n1), dst + 2 * t.stage0);
        fill_stage(bfly_plan::plan_index(n2), dst + 2 * t.stage1);
        for (uint64_t m = 0; m < n; m++) twiddle(m, n, dst + 2 * (t.extra + m));
        if (t.single) fill_stage(bfly_plan::plan_index(n), dst + 2 * t.single);
    } else {
        const uint32_t m = n / 2;
        fill_stage(bfly_plan::plan_index(m), dst);
        for (uint64_t k = 0; k <= m; k++) twiddle(k, n, dst + 2 * (t.extra + k));
    }
    return BFLY_OK;
}

// Largest grid any launcher issues for (kind, n, batch) must fit a 32-bit x
// dimension.
bool split_mfma_ok(uint32_t n) { return n == 1024 || n == 2048 || n == 4096; }

bool needs_work(int kind, uint32_t n, uint32_t batch, int variant) {
    if (kind != BFLY_KIND_C2C || n <= (uint32_t)bfly_plan::kMaxLocalN) return false;
    if (variant == BFLY_VARIANT_AUTO) variant = auto_variant(n, batch);
    return n > (uint32_t)bfly_plan::kMaxSingleN || variant == BFLY_VARIANT_FOUR_STEP;
}

static bool grid_ok(int kind, uint32_t n, uint32_t batch) {
    uint64_t blocks = 0;
    const uint64_t b = batch;
    if (kind == BFLY_KIND_C2C && n > (uint32_t)bfly_plan::kMaxLocalN) {
        uint32_t n1 = 0, n2 = 0;
        bfly_plan::long_split(n, &n1, &n2);
        const uint64_t c1 = bfly_plan::kBlock / bfly_plan::threads_per_transform(n1);
        const uint64_t c2 = bfly_plan::kBlock / bfly_plan::threads_per_transform(n2);
        const uint64_t g1 = b * (n2 / c1), g2 = b * (n1 / c2);
        blocks = g1 > g2 ? g1 : g2;
        // single-workgroup path: one workgroup of n/16 threads per transform
        if (n <= (uint32_t)bfly_plan::kMaxSingleN && b * (n / 16) > 0xffffffffull) return false;
    } else {
        const uint32_t local = kind == BFLY_KIND_C2C ? n : n / 2;
        const uint64_t per = bfly_plan::kBlock / bfly_plan::threads_per_transform(local);
        blocks = (b + per - 1) / per;
    }
    return blocks >= 1 && blocks <= 0x7fffffffull;
}

bool disjoint(int kind, const void* in, const void* out, const void* work, uint32_t n, uint32_t batch,
              uint32_t bins, uint32_t stride) {
    const uint64_t b = batch, spec = (b - 1) * stride + 2ull * bins;  // half-spectrum floats
    const uint64_t in_f = kind == BFLY_KIND_C2C ? 2ull * n * b : kind == BFLY_KIND_R2C ? (uint64_t)n * b : spec;
    const uint64_t out_f = kind == BFLY_KIND_C2C ? 2ull * n * b : kind == BFLY_KIND_R2C ? spec : (uint64_t)n * b;
    auto apart = [](const void* p, uint64_t pf, const void* q, uint64_t qf) {
        const uintptr_t a = (uintptr_t)p, c = (uintptr_t)q;
        return a + 4 * pf <= c || c + 4 * qf <= a;
    };
    if (!apart(in, in_f, out, out_f)) return false;
    if (work && (!apart(work, 2ull * n * b, in, in_f) || !apart(work, 2ull * n * b, out, out_f))) return false;
    return true;
}

int check_execute(int kind, int sign, const void* in, const void* out, uint32_t n,
                  uint32_t batch, const void* table, const void* work, int variant) {
    if (!in || !out || batch == 0 || !shape_ok(kind, n)) return BFLY_ERR_SHAPE;
    if (sign != -1 && sign != 1) return BFLY_ERR_SHAPE;
    if (!grid_ok(kind, n, batch)) return BFLY_ERR_SHAPE;
    if (!table) return BFLY_ERR_TABLE;
    if (needs_work(kind, n, batch, variant) && !work) return BFLY_ERR_SHAPE;
    if (kind == BFLY_KIND_C2C &&
        (variant == BFLY_VARIANT_SPLIT_MFMA || variant == BFLY_VARIANT_SPLIT_MFMA_STREAM) && !split_mfma_ok(n))
        return BFLY_ERR_SHAPE;
    return BFLY_OK;
}

// Measured on MI355X (C2C, 4 rotated buffer pairs, HIP events): the
// streaming form with LDS twiddles wins from n = 256 once the batch fills
// several waves of resident workgroups (n=4096 b=4096: 47.5 us vs 62.2 us
// one-shot; n=256 b=65536: 46.1 vs 48.5 us); below that the one-shot grid
// is as fast or faster (n=128 b=131072: 49.4 vs 50.3 us).
// One workgroup per transform at 8192 / 16384 (its LDS-twiddle streaming
// form) beats the four-step chain once the batch fills the compute units
// (8192 b=2048: 56.5 vs 93.3 us; 8192 b=64: 9.4 vs 11.9 us; 16384 b=1024:
// 74.6 vs 98.1 us; 16384 b=256: 19.4 vs 31.1 us); with few 16384-point
// transforms the four-step chain's wider grid wins (b=64: 11.6 vs 13.8 us).
int auto_variant(uint32_t n, uint32_t batch) {
    const uint64_t points = (uint64_t)n * batch;
    if (n >= 256 && n <= (uint32_t)bfly_plan::kMaxLocalN && points >= (1ull << 22))
        return BFLY_VARIANT_STREAM_LT;
    if (n == 8192) return BFLY_VARIANT_STREAM_LT;
    if (n == 16384) return batch < 256 ? BFLY_VARIANT_FOUR_STEP : BFLY_VARIANT_STREAM_LT;
    return BFLY_VARIANT_PLAIN;
}

int scaled_variant(uint32_t n, uint32_t batch) {
    return n <= (uint32_t)bfly_plan::kMaxLocalN ? BFLY_VARIANT_STREAM_LT : auto_variant(n, batch);
}

// Real transforms, once the batch fills several rounds of resident
// workgroups: streaming with LDS twiddles from n = 512 (measured, MI355X:
// R2C n=8192 b=4096 54.3 us vs 74.1 one-shot / 74.8 streaming; C2R n=4096
// b=4096 31.2 vs 35.1 / 33.9; R2C n=512 b=32768 27.2 vs 32.9 / 30.0, C2R
// 30.0 vs 34.2 / 30.0), the register-table streaming form below (n=256
// b=65536: R2C 33.4 vs 37.5 LDS twiddles / 35.9 one-shot, C2R 30.2 vs
// 34.5 / 34.4).
int auto_real_variant(uint32_t n, uint32_t batch) {
    if ((uint64_t)n * batch < (1ull << 22)) return BFLY_VARIANT_PLAIN;
    return n >= 512 ? BFLY_VARIANT_STREAM_LT : BFLY_VARIANT_STREAM;
}

int check_pipelined(int kind, int sign, int variant, const void* host_in, const void* host_out, uint32_t n,
                    uint32_t batch, const void* table, const void* dev_in, const void* dev_out, const void* work,
                    uint32_t chunk_rows, PipeShape* shape) {
    if (!shape) return BFLY_ERR_SHAPE;
    if (kind != BFLY_KIND_C2C && kind != BFLY_KIND_R2C && kind != BFLY_KIND_C2R) return BFLY_ERR_SHAPE;
    if (kind != BFLY_KIND_C2C) sign = kind == BFLY_KIND_R2C ? -1 : 1;
    if (!host_in || !host_out || !dev_in || !dev_out) return BFLY_ERR_SHAPE;
    if (!shape_ok(kind, n)) return BFLY_ERR_SHAPE;
    if (variant < BFLY_VARIANT_AUTO || variant > BFLY_VARIANT_LAST) return BFLY_ERR_SHAPE;
    if (batch == 0) return BFLY_ERR_SHAPE;
    const bool c2c = kind == BFLY_KIND_C2C;
    shape->row_in = c2c ? 2 * (size_t)n : kind == BFLY_KIND_R2C ? (size_t)n : (size_t)n + 2;
    shape->row_out = c2c ? 2 * (size_t)n : kind == BFLY_KIND_R2C ? (size_t)n + 2 : (size_t)n;
    // about 8 MiB of input per chunk unless the caller says otherwise
    uint64_t cr = chunk_rows ? chunk_rows : (8ull << 20) / (shape->row_in * sizeof(float));
    if (cr < 1) cr = 1;
    if (cr > batch) cr = batch;
    shape->chunk = cr;
    // the kernels run on chunks, so AUTO picks the variant for a chunk's batch
    if (c2c) {
        if (variant == BFLY_VARIANT_AUTO) variant = auto_variant(n, (uint32_t)cr);
    } else {
        if (variant == BFLY_VARIANT_AUTO) variant = auto_real_variant(n, (uint32_t)cr);
        if (variant != BFLY_VARIANT_PLAIN && variant != BFLY_VARIANT_STREAM && variant != BFLY_VARIANT_STREAM_LT)
            return BFLY_ERR_SHAPE;
    }
    if (int g = check_execute(kind, sign, dev_in, dev_out, n, batch, table, c2c ? work : nullptr, variant))
        return g;
    if (!disjoint(kind, dev_in, dev_out, nullptr, n, batch, n / 2 + 1, n + 2)) return BFLY_ERR_SHAPE;
    // the workspace holds one chunk and lies apart from both whole-batch buffers
    if (c2c && needs_work(BFLY_KIND_C2C, n, batch, variant)) {
        const uintptr_t w = (uintptr_t)work, wb = 8ull * n * cr, ab = 8ull * n * batch;
        for (const void* q : {dev_in, dev_out}) {
            const uintptr_t a = (uintptr_t)q;
            if (!(w + wb <= a || a + ab <= w)) return BFLY_ERR_SHAPE;
        }
    }
    shape->sign = sign;
    shape->variant = variant;
    return BFLY_OK;
}

bool bins_ok(uint32_t n, uint32_t bins, uint32_t stride) {
    // 1 <= bins <= n/2+1, the row holds the bins, rows stay 8-byte aligned
    return bins >= 1 && bins <= n / 2 + 1 && stride >= 2 * bins && (stride % 2) == 0;
}

int check_mode_mix(int op, const float* a, const float* b, const float* c, uint32_t batch, uint32_t c_in,
                   uint32_t c_out, uint32_t modes) {
    if (op < BFLY_MIX_FORWARD || op > BFLY_MIX_GRAD_WEIGHT) return BFLY_ERR_SHAPE;
    if (!a || !b || !c || batch == 0 || c_in == 0 || c_out == 0 || modes == 0) return BFLY_ERR_SHAPE;
    for (const float* p : {a, b, c})
        if ((uintptr_t)p % 8) return BFLY_ERR_SHAPE;
    // element counts, refused when a product leaves 64 bits (no device holds them)
    uint64_t x, w, y;
    if (__builtin_mul_overflow((uint64_t)batch * c_in, (uint64_t)modes, &x) ||
        __builtin_mul_overflow((uint64_t)c_in * c_out, (uint64_t)modes, &w) ||
        __builtin_mul_overflow((uint64_t)batch * c_out, (uint64_t)modes, &y) || x >> 60 || w >> 60 || y >> 60)
        return BFLY_ERR_SHAPE;
    // complex elements of a, b, c and the (m, n) extents of c for each op
    const uint64_t na = op == BFLY_MIX_GRAD_INPUT ? y : x;
    const uint64_t nb = op == BFLY_MIX_GRAD_WEIGHT ? y : w;
    const uint64_t nc = op == BFLY_MIX_FORWARD ? y : op == BFLY_MIX_GRAD_INPUT ? x : w;
    const uint64_t m = op == BFLY_MIX_GRAD_WEIGHT ? c_in : batch;
    const uint64_t n = op == BFLY_MIX_GRAD_INPUT ? c_in : c_out;
    // tiles of 16 modes x 16 x 16; each factor < 2^29, so no product overflows
    const uint64_t tiles2 = ((uint64_t)modes + 15) / 16 * ((m + 15) / 16);
    if (tiles2 > 0x7FFFFFFFull || tiles2 * ((n + 15) / 16) > 0x7FFFFFFFull) return BFLY_ERR_SHAPE;
    auto overlap = [](const float* p, uint64_t np, const float* q, uint64_t nq) {
        const uintptr_t p0 = (uintptr_t)p, q0 = (uintptr_t)q;
        return p0 < q0 + 8 * nq && q0 < p0 + 8 * np;
    };
    if (overlap(c, nc, a, na) || overlap(c, nc, b, nb)) return BFLY_ERR_SHAPE;
    return BFLY_OK;
}

bool group_ok(int kind, uint32_t n, int wg) {
    if (wg != 512 && wg != 1024) return false;
    if (!pow2(n)) return false;
    const uint32_t m = group_local_len(kind, n);
    if (kind == BFLY_KIND_C2C) {
        if (m < 16) return false;
    } else if (kind == BFLY_KIND_R2C || kind == BFLY_KIND_C2R) {
        if (n < 32 || m > (uint32_t)bfly_plan::kMaxLocalN) return false;
    } else {
        return false;
    }
    return bfly_plan::plan_index(m) >= 0 && m / bfly_plan::kPoints <= (uint32_t)wg;
}

int group_plan(const bfly_group_request* reqs, int nreq, const bfly_group_config* cfg, GroupPlan* out) {
    if (!reqs || nreq <= 0 || !out) return BFLY_ERR_SHAPE;
    bfly_group_config c{512, 0, 0, 0, 0, 0};
    if (cfg) c = *cfg;
    if (c.workgroup == 0) c.workgroup = 512;
    if ((c.workgroup != 512 && c.workgroup != 1024) || (c.schedule != 0 && c.schedule != 1) ||
        c.order < 0 || c.order > 3 || (c.dynamic != 0 && c.dynamic != 1) || c.grid < 0 ||
        (c.twiddles != 0 && c.twiddles != 1) || (c.twiddles == 1 && c.schedule != 0))
        return BFLY_ERR_SHAPE;
    GroupPlan& P = *out;
    P = GroupPlan{};
    P.cfg = c;
    P.per.resize(nreq);
    P.count.resize(nreq);
    for (int r = 0; r < nreq; r++) {
        const bfly_group_request& q = reqs[r];
        if (!q.in || !q.out || q.batch == 0 || !group_ok(q.kind, q.n, c.workgroup)) return BFLY_ERR_SHAPE;
        if (!disjoint(q.kind, q.in, q.out, nullptr, q.n, q.batch, q.n / 2 + 1, q.n + 2)) return BFLY_ERR_SHAPE;
        if (q.kind == BFLY_KIND_C2C && q.sign != -1 && q.sign != 1) return BFLY_ERR_SHAPE;
        P.per[r] = (uint32_t)c.workgroup / (group_local_len(q.kind, q.n) / bfly_plan::kPoints);
        P.count[r] = (uint32_t)(((uint64_t)q.batch + P.per[r] - 1) / P.per[r]);
        P.total += P.count[r];
    }
    if (P.total == 0 || P.total > 0x7fffffffull) return BFLY_ERR_SHAPE;
    P.unit_req.reserve(P.total);
    P.unit_grp.reserve(P.total);
    auto push = [&](int r, uint32_t g) {
        P.unit_req.push_back((uint32_t)r);
        P.unit_grp.push_back(g);
    };
    if (c.order == 2) {
        // round robin: unit g of every request before unit g+1 of any
        uint32_t most = 0;
        for (int r = 0; r < nreq; r++) most = std::max(most, P.count[r]);
        for (uint32_t g = 0; g < most; g++)
            for (int r = 0; r < nreq; r++)
                if (g < P.count[r]) push(r, g);
        return BFLY_OK;
    }
    std::vector<int> ord(nreq);
    for (int r = 0; r < nreq; r++) ord[r] = r;
    if (c.order == 3) {
        // least total work first: the short requests complete early, which
        // minimises the mean and median completion time of the group
        auto total = [&](int r) {
            const uint64_t m = group_local_len(reqs[r].kind, reqs[r].n);
            uint32_t lg = 0;
            while ((1u << lg) < m) lg++;
            return (uint64_t)reqs[r].batch * m * (uint64_t)lg;
        };
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return total(a) < total(b); });
    }
    if (c.order == 1) {
        // most work per unit first: filled transforms * n' * log2 n'
        auto work = [&](int r) {
            const uint32_t m = group_local_len(reqs[r].kind, reqs[r].n);
            const uint64_t full = std::min<uint64_t>(reqs[r].batch, P.per[r]);
            uint32_t lg = 0;
            while ((1u << lg) < m) lg++;
            return full * m * (uint64_t)lg;
        };
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return work(a) > work(b); });
    }
    for (int r : ord)
        for (uint32_t g = 0; g < P.count[r]; g++) push(r, g);
    return BFLY_OK;
}

}  // namespace bfly_host

extern "C" {

int bfly_supported(int kind, uint32_t n) { return bfly_host::shape_ok(kind, n) ? 1 : 0; }

size_t bfly_table_floats(int kind, uint32_t n) {
    return 2 * bfly_host::table_layout(kind, n).complex_len;
}

int bfly_table_fill(int kind, uint32_t n, float* host_table) {
    return bfly_host::fill_table(kind, n, host_table);
}

size_t bfly_workspace_bytes(int kind, uint32_t n, uint32_t batch) {
    if (!bfly_host::shape_ok(kind, n) || !bfly_host::needs_work(kind, n, batch, BFLY_VARIANT_AUTO))
        return 0;
    return (size_t)n * (size_t)batch * 8u;
}

}  // extern "C"
