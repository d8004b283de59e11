// The reference's plan-descriptor integer rules, kept as pure functions so
// descriptor records computed here match the reference's bit for bit
// (include/butterfly/fft_k.hpp:45-144 of the reference).  They describe the
// Ascend row kernel (K-plane factor, batch fold D); the gfx950 kernels do not
// consume them -- their own plan is bfly_plan::kPlans.
#pragma once
#include <stdint.h>

namespace bfly {

namespace detail {
inline uint32_t log2ceil(uint32_t n) {
    uint32_t l = 0;
    while ((1u << l) < n) l++;
    return l;
}
}  // namespace detail

// K in {8,16,32} (K*8 <= n) minimising the Ascend vector-op count; for
// n <= 1024 only K with n/K <= 64 are considered; ties go to the larger K.
inline uint32_t planeKFor(uint32_t n) {
    const uint32_t need = (n + 63u) / 64u;
    const uint32_t first = (n <= 1024u && need > 8u) ? need : 8u;
    uint32_t best = first;
    uint64_t bestOps = UINT64_MAX;
    const uint32_t logn = detail::log2ceil(n);
    for (uint32_t K = first; K <= 32u && K * 8u <= n; K *= 2u) {
        const uint32_t logK = detail::log2ceil(K);
        uint64_t ops = 7;  // per-row copies and gathers
        uint32_t s = 0;
        for (; s + 1u < logK; s += 2u) {  // fused radix-4 plane stages
            const uint64_t h = 1ull << s;
            ops += (K / (4u * h)) * (16u + (h - 1u) * 28u);
        }
        if (s < logK) {  // trailing radix-2 plane stage
            const uint64_t h = 1ull << s, trivial = K / (2u * h);
            ops += trivial * 6u + (K / 2u - trivial) * 8u;
        }
        for (uint32_t st = logK; st < logn; st++) {  // planar stages
            const uint32_t h = 1u << st, groups = n / (2u * h), slices = (h + 63u) / 64u;
            const bool merged = h >= 8u && (h / 4u) <= 255u && (h / 8u) <= 255u &&
                                groups <= 255u && slices <= groups;
            ops += (merged ? 8ull * slices : 8ull * groups) + 1u;
        }
        if (ops <= bestOps) {
            bestOps = ops;
            best = K;
        }
    }
    return best;
}

inline uint32_t planeLogKFor(uint32_t n) { return detail::log2ceil(planeKFor(n)); }

// Batch fold D in 1..4: n <= 1024, D * n/(2K) <= 255, and (when nblk > 1)
// at least nblk row groups of D rows.
inline uint32_t foldDFor(uint32_t n, uint32_t batch = 0xFFFFFFFFu, uint32_t nblk = 1u) {
    const uint32_t K = planeKFor(n);
    if (K == 0u || n % K != 0u || n / K == 0u || n / 8u > 255u) return 1u;
    const uint32_t groupsMax = n >> (detail::log2ceil(K) + 1u);
    if (groupsMax == 0u) return 1u;
    uint32_t D = 4u;
    while (D > 1u && D * groupsMax > 255u) D--;
    if (nblk > 1u)
        while (D > 1u && (batch - 1u) / D + 1u < nblk) D--;
    return D;
}

}  // namespace bfly
