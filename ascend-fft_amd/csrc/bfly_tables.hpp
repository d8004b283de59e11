// bfly_tables.hpp -- host-side table layout shared by the C ABI, the kernel
// launchers and the CPU stub.
#pragma once
#include <cstddef>
#include <cstdint>

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
TableLayout table_layout(int kind, uint32_t n);
int fill_table(int kind, uint32_t n, float* dst);
// Argument guards of the execute entry points: BFLY_OK or the status code.
int check_execute(int kind, int sign, const void* in, const void* out, uint32_t n,
                  uint32_t batch, const void* table, const void* work);

// Mode-truncated real transforms: 1 <= bins <= n/2+1, stride (floats) even
// and at least 2*bins.
bool bins_ok(uint32_t n, uint32_t bins, uint32_t stride);
// The variant BFLY_VARIANT_AUTO resolves to for a C2C of (n, batch).
int auto_variant(uint32_t n, uint32_t batch);

}  // namespace bfly_host
