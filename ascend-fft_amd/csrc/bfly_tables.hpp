// bfly_tables.hpp -- host-side table layout shared by the C ABI, the kernel
// launchers and the CPU stub.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "bfly_hip.h"

namespace bfly_host {

// Offsets in complex entries inside the table of one (kind, n).
//   local C2C:  [stage table of n]
//   long C2C:   [stage table of N1][stage table of N2][W_N^m, m < N]
//   R2C / C2R:  [stage table of n/2][W_n^k, k = 0..n/2, padded to even]
//   long C2C with a single-workgroup plan (n <= 16384) appends
//               [stage table of the single-workgroup plan of n]
struct TableLayout {
    size_t stage0 = 0;      // first stage table
    size_t stage1 = 0;      // second stage table (long C2C only)
    size_t extra = 0;       // W_N (long) or W_n (real) table
    size_t single = 0;      // single-workgroup pass table (long C2C, n <= 16384)
    size_t complex_len = 0; // total complex entries (0 = unsupported)
};

bool shape_ok(int kind, uint32_t n);
// FP64 e^{-2 pi i m / L} (m reduced modulo L) rounded to fp32 into out[0..1].
void twiddle(uint64_t m, uint64_t L, float* out);
// Writes the inter-pass twiddle table of plan `plan` under radix schedule
// `sched` (interleaved fp32, when dst is not null); returns its complex length.
size_t fill_stage(int plan, float* dst, int sched = 0);
TableLayout table_layout(int kind, uint32_t n);
int fill_table(int kind, uint32_t n, float* dst);
// The buffers of one execute call do not overlap: input, output and (long
// C2C) workspace, at the extents the call reads and writes (real kinds:
// `bins` bins at a row stride of `stride` floats on the half-spectrum side).
bool disjoint(int kind, const void* in, const void* out, const void* work, uint32_t n, uint32_t batch,
              uint32_t bins, uint32_t stride);
// Argument guards of the execute entry points: BFLY_OK or the status code.
// `variant` matters only for long C2C, where it decides whether a workspace
// is needed (BFLY_VARIANT_FOUR_STEP, or n above the single-workgroup limit).
int check_execute(int kind, int sign, const void* in, const void* out, uint32_t n,
                  uint32_t batch, const void* table, const void* work,
                  int variant = BFLY_VARIANT_AUTO);
// True when a C2C of (n, batch) run as `variant` (AUTO resolved by
// auto_variant) goes through the four-step path and its n*batch complex
// workspace.
bool needs_work(int kind, uint32_t n, uint32_t batch, int variant);
// Lengths the split-FP16 matrix-core variant runs (plans [16, 16, R]).
bool split_mfma_ok(uint32_t n);

// 2-D C2C of [batch][n1][n2]: n1, n2 powers of two in [16, 4096].  Column
// lengths up to 2048 run as one pass (p = n1, q = 1); 4096 = p * q = 64 * 64.  The column table
// of n1 is [local C2C table of p][local C2C table of q][W_n1^j, j < 64 |
// W_n1^{64 j}, j < n1 / 64, zero-padded to 128 entries] (complex offsets).
struct Fft2dLayout {
    uint32_t p = 0, q = 0;
    size_t p_table = 0, q_table = 0, wab = 0, complex_len = 0;
};
bool fft2d_ok(uint32_t n1, uint32_t n2);
Fft2dLayout fft2d_layout(uint32_t n1);
int fft2d_fill(uint32_t n1, float* dst);
// Guards of bfly_c2c_columns (n1 a power of two in [16, 4096], ncols and
// batch >= 1, disjoint buffers, a workspace when n1 > 256, grids that fit).
int check_columns(int sign, const void* in, const void* out, uint32_t n1, uint32_t ncols, uint32_t batch,
                  const void* table_cols, const void* work);
// Guards of bfly_c2c_rows: n a power of two in [16, 65536], batch >= 1,
// row distances >= n (complex values), 8-byte aligned buffers whose spans
// [p, p + 8 ((batch-1) dist + n)) do not overlap, extents below 2^58; above
// 16384 a workspace of 8 n batch bytes overlapping neither span.
int check_rows(int sign, const void* in, uint64_t idist, const void* out, uint64_t odist, uint32_t n,
               uint32_t batch, const void* table, const void* work);
// Guards of bfly_fft2d: BFLY_OK or the status code.
int check_fft2d(int sign, const void* in, const void* out, uint32_t n1, uint32_t n2, uint32_t batch,
                const void* table_rows, const void* table_cols, const void* work);

// Mode-truncated real transforms: 1 <= bins <= n/2+1, stride (floats) even
// and at least 2*bins.
bool bins_ok(uint32_t n, uint32_t bins, uint32_t stride);
// The variant BFLY_VARIANT_AUTO resolves to for a C2C of (n, batch).
int auto_variant(uint32_t n, uint32_t batch);
// The variant BFLY_VARIANT_AUTO resolves to for a real transform of (n, batch).
int auto_real_variant(uint32_t n, uint32_t batch);
// Guards and shape of one bfly_host_pipelined call (shared by the HIP library
// and the CPU stub): resolves the direction of the real kinds and the AUTO
// variant for one chunk's batch, checks the whole batch as one execute call
// on the device buffers, and gives the row sizes (floats) and transforms per chunk.
struct PipeShape {
    int sign = -1, variant = 0;
    size_t row_in = 0, row_out = 0;
    uint64_t chunk = 0;
};
int check_pipelined(int kind, int sign, int variant, const void* host_in, const void* host_out, uint32_t n,
                    uint32_t batch, const void* table, const void* dev_in, const void* dev_out, const void* work,
                    uint32_t chunk_rows, PipeShape* shape);
// The variant bfly_c2c_scaled runs: one whose last store applies the scale
// (the LDS-twiddle streaming kernel up to n = 8192, AUTO's choice above).
int scaled_variant(uint32_t n, uint32_t batch);

// Guards of bfly_mode_mix: op 0..2, non-null 8-byte aligned operands, every
// size >= 1, an output overlapping neither input, a grid that fits.
// BFLY_OK or BFLY_ERR_SHAPE.
int check_mode_mix(int op, const float* a, const float* b, const float* c, uint32_t batch, uint32_t c_in,
                   uint32_t c_out, uint32_t modes);

// ------------------------------------------------------- grouped launch
// Whether a grouped launch with `wg` threads per workgroup takes (kind, n).
bool group_ok(int kind, uint32_t n, int wg);
// Points per transform of the complex plan a request runs (n, or n/2 for
// the real kinds).
inline uint32_t group_local_len(int kind, uint32_t n) { return kind == BFLY_KIND_C2C ? n : n / 2; }

// Host side of a grouped launch: the resolved configuration, transforms per
// unit and unit count of every request, and the unit list (request, group
// index) in launch order.
struct GroupPlan {
    bfly_group_config cfg{};
    std::vector<uint32_t> per, count;
    std::vector<uint32_t> unit_req, unit_grp;
    uint64_t total = 0;
};
// Validates requests and configuration and builds the unit list: BFLY_OK or
// BFLY_ERR_SHAPE.
int group_plan(const bfly_group_request* reqs, int nreq, const bfly_group_config* cfg, GroupPlan* out);

}  // namespace bfly_host
