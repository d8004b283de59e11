// bfly_plans.hpp -- the radix plans the gfx950 kernels are instantiated for,
// shared by the device code (template arguments) and the host (twiddle-table
// layout, launch geometry), so the two can never disagree.
//
// A plan for length N lists its radices first pass first.  Every thread owns
// P = 16 points, so a transform uses T = N/16 threads and a 256-thread
// workgroup holds 256/T transforms (one for N = 4096).  Radix 16 everywhere it
// divides, the remainder as the last pass: 4096 = 16*16*16 is three register
// passes with two LDS exchanges.
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace bfly_plan {

constexpr int kPoints = 16;       // points per thread
constexpr int kBlock = 256;       // threads per workgroup for N <= 4096
constexpr int kMaxRadices = 5;
constexpr int kMinN = 16;
constexpr int kMaxLocalN = 4096;  // largest length held by one 256-thread workgroup
constexpr int kMaxSingleN = 16384;  // largest length held by one workgroup (1024 threads)

struct Plan {
    int n;
    int nrad;
    int rad[kMaxRadices];
};

constexpr Plan kPlans[] = {
    {16, 1, {16, 0, 0, 0, 0}},
    {32, 2, {16, 2, 0, 0, 0}},
    {64, 2, {16, 4, 0, 0, 0}},
    {128, 2, {16, 8, 0, 0, 0}},
    {256, 2, {16, 16, 0, 0, 0}},
    {512, 3, {16, 16, 2, 0, 0}},
    {1024, 3, {16, 16, 4, 0, 0}},
    {2048, 3, {16, 16, 8, 0, 0}},
    {4096, 3, {16, 16, 16, 0, 0}},
    // Single-workgroup plans of the long lengths (T = n/16 = 512 / 1024
    // threads, the whole transform LDS-resident: 68 / 136 KiB).
    {8192, 4, {16, 16, 16, 2, 0}},
    {16384, 4, {16, 16, 16, 4, 0}},
};
constexpr int kNumPlans = (int)(sizeof(kPlans) / sizeof(kPlans[0]));

// Radix schedules of one plan: 0 runs the radices as listed (radix 16 first,
// the remainder radix last), 1 moves the remainder radix to the front.  Both
// use the same radices and differ only in pass order (and so in rounding).
constexpr int kSchedules = 2;
constexpr int radix_at(int i, int sched, int s) {
    return (sched == 0 || kPlans[i].nrad == 1)
               ? kPlans[i].rad[s]
               : (s == 0 ? kPlans[i].rad[kPlans[i].nrad - 1] : kPlans[i].rad[s - 1]);
}

constexpr int plan_index(uint32_t n) {
    for (int i = 0; i < kNumPlans; i++)
        if ((uint32_t)kPlans[i].n == n) return i;
    return -1;
}

// Complex entries of the inter-pass twiddle table of plan i:
// sum over passes s >= 1 of Ns * R_s, Ns = R_0 * ... * R_{s-1}.
constexpr size_t stage_twiddle_len(int i, int sched = 0) {
    size_t len = 0, ns = 1;
    for (int s = 0; s < kPlans[i].nrad; s++) {
        if (s > 0) len += ns * (size_t)radix_at(i, sched, s);
        ns *= (size_t)radix_at(i, sched, s);
    }
    return len;
}

constexpr int threads_per_transform(uint32_t n) { return (int)(n / kPoints); }

// Workgroups the LDS-twiddle streaming kernels (C2C, R2C, C2R) launch for
// `groups` units of work with `resident` workgroups fitting the GPU at once.
// Above two units per resident slot: about 1.5 units per workgroup.  More
// workgroups than fit means the dispatcher starts new ones as others drain,
// so the first loads and the last computes of different workgroups overlap
// instead of every workgroup filling and draining its pipeline at the same
// time (measured on MI355X, C2C 4096 x 4096: 46.6-47.9 us at 1.45-1.5 units
// per workgroup vs 49.5 us with exactly the resident count; 4096 x 16384:
// 173.8 vs 194 us; R2C / C2R 8192 x 4096: 54.7 / 57.0 vs 57.8 / 59.1 us at two
// units per workgroup).  Up to two units per slot: the resident count (C2C
// 4096 x 2048: 26.1 vs 27.4-28.9 us at 1.5; 8192 x 1024: 29.5 vs 32.7-33.9;
// R2C / C2R 4096 x 4096: 28.5 / 28.0 vs 31.0 / 30.2; R2C 1024 x 16384: 32.4
// vs 34.4).  The register-twiddle and matrix-core streaming kernels keep the
// resident count.
constexpr uint64_t stream_grid(uint64_t groups, uint64_t resident) {
    if (groups <= 2 * resident) return groups < resident ? groups : resident;
    return (groups * 2 + 2) / 3;
}

// Four-step split of a long length into two local lengths N1 * N2 (column
// FFTs of N1 first, then row FFTs of N2).  Balanced, N2 >= N1.
constexpr bool long_split(uint32_t n, uint32_t* n1, uint32_t* n2) {
    if (n <= (uint32_t)kMaxLocalN || n > (uint32_t)kMaxLocalN * (uint32_t)kMaxLocalN) return false;
    uint32_t l = 0;
    while ((1u << l) < n) l++;
    if ((1u << l) != n) return false;
    const uint32_t a = 1u << (l / 2), b = n / a;
    if (n1) *n1 = a;
    if (n2) *n2 = b;
    return a >= 16 && b <= (uint32_t)kMaxLocalN;
}

}  // namespace bfly_plan
