// Design objects H / G / A / P / L / F / Q of the plan-selection framework,
// re-derived for gfx950 (MI355X).  Type and member names follow the
// reference's include/butterfly/objects.hpp:13-113 so code written against
// that API compiles unchanged; each member documents its gfx950 meaning.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "butterfly/fft_k.hpp"

namespace bfly {

// ----------------------------------------------------------------- H
// Loaded from the gfx950 hardware profile (ascend-fft_amd/config/
// gfx950_profile.json, produced from bfly_probe_device + bfly_probe_copy).
struct Hardware {
    std::string soc;               // "gfx950"
    std::string cannVersion;       // runtime stack version (ROCm); name kept for source compatibility
    int aicoreNum = 0;             // compute units
    int vectorCoreNum = 0;         // compute units (upper bound of Mapping::udCore)
    int ubBytesPerCore = 0;        // LDS bytes per CU
    int vecLaneFp32 = 64;          // wavefront width (fp32 lanes per instruction)
    int l2Bytes = 0;               // L2 bytes (per XCD)
    bool simt = true;              // per-lane programming model with cross-lane DPP/permute
    int subBlockNum = 4;           // SIMDs per CU
    int ubOffsetAlignFloats = 4;   // 16-byte alignment of 128-bit LDS / global accesses
    int minBlockOutputBytes = 0;   // no minimum output granule on gfx950
    bool matrixUnitUsable = true;  // MFMA matrix cores
    // gfx950-only facts
    int ldsBytesPerWorkgroup = 0;  // largest LDS allocation of one workgroup
    int xcds = 8;                  // accelerator complex dies (one L2 each)
    long long mallBytes = 0;       // last-level (Infinity) cache bytes
    long long hbmBytes = 0;        // device memory bytes
    double hbmPeakGBs = 0.0;       // published HBM peak
    double hbmCopyGBs = 0.0;       // measured streaming-copy bandwidth

    // Throws std::exception for an unreadable, malformed or invalid profile.
    static Hardware load(const std::string& profilePath);
    bool supportsShuffleExchange() const { return simt; }
};

// ----------------------------------------------------------------- G
// The reference's input-independent tables (butterfly.cpp:96-166), kept as
// closed-form generators for plan-descriptor parity.  The gfx950 kernels are
// Stockham autosort passes and upload their own twiddle layout
// (bfly_table_fill); they need no permutation tables.
struct Generator {
    enum Kind : uint8_t { Radix2Dit = 0, Stockham = 1 } kind = Stockham;
    uint32_t n = 0;

    uint32_t twiddlePad() const { return n + 16; }
    // per stage h = 1..n/2 a slot of align8(h) entries, entry j = e^{-i pi j/h}
    void genTwiddles(std::vector<float>& twr, std::vector<float>& twi) const;
    // dst[j] = 2 * bitrev(j)
    void genBitReverse(std::vector<uint32_t>& dst) const;
    uint32_t planeK() const { return planeKFor(n); }
    // dst[k] = 8 * bitrev(((k % rows) << logK) | (k / rows)), rows = n / K
    void genBitReversePlane(std::vector<uint32_t>& dst) const;
    // dst[k] = 4 * ((k & (K-1)) * rows + (k >> logK))
    void genTransposePlane(std::vector<uint32_t>& dst) const;
    // dst[d*2n+2j] = 4(d n + j), dst[d*2n+2j+1] = 4(n D + d n + j)
    void genInterleave(std::vector<uint32_t>& dst, uint32_t D = 1) const;
    // The gfx950 kernel table of (kind, n) as uploaded by Plan::prepare.
    void genKernelTable(int kind, std::vector<float>& table) const;
};

// ----------------------------------------------------------------- A
struct Mapping {
    int udCore = 256;      // compute units the grid may occupy
    int td = 1;            // tile reuse (1 = one transform tile per pass)
    int tb = 1;            // 1 = one tile of transforms per workgroup (one-shot grid);
                           // 0 = a resident grid whose workgroups walk the batch
    int ub = 1;            // tiles in flight per workgroup (1, or 2 in the two-deep streaming kernel)
    int ts = 1;            // 1 = all passes of a transform stay resident in LDS
    int us = 1;            // stage services per workgroup
    int level = 0;         // 0 = one GPU; 1 = multi-GPU (not implemented)
    std::string localExchange = "shared";  // shared (LDS, padded) | swizzle (LDS, XOR) | shuffle | register

    bool validate(const Hardware& hw, std::string* why) const;
};

// ----------------------------------------------------------------- P
struct StagePlan {
    int radix = 16;                    // points per thread: largest radix of the pass plan (16 | 8)
    bool coefficientResidency = true;  // twiddles read from the device table (cached), not recomputed
    int fusionLevel = 1;               // passes fused per LDS exchange (1 = one radix pass per exchange)
    bool bitReverseInput = false;      // Stockham autosort: no bit-reversal permutation

    // LDS bytes of one 256-thread workgroup for length n.
    size_t ubBytes(uint32_t n) const;
    bool fits(uint32_t n, const Hardware& hw) const { return ubBytes(n) <= (size_t)hw.ubBytesPerCore; }
};

// ----------------------------------------------------------------- L
struct Layout {
    enum Pack : uint8_t { Interleaved = 0, Planar = 1 };
    Pack in = Interleaved;
    Pack out = Interleaved;
    bool contiguous = true;
};

// ----------------------------------------------------------------- F
struct Fusion {
    int butterflyCount = 1;
    int stageCount = 1;
    bool dftMatrix = false;  // radix-16 passes as dense DFT matmuls on MFMA (split-FP16)
};

// ----------------------------------------------------------------- Q
struct Metric {
    double etaUs = 0.0;       // model estimate (us)
    double measuredUs = 0.0;  // measured (us, HIP events)
    double maxRelErr = 0.0;   // maxRelScaled of the sampled rows
    bool measuredOk = false;
};

// Roofline cost model of the gfx950 kernels (calibrated constants in
// butterfly.cpp): launch overhead + max(HBM time, VALU time).
Metric estimate(const Hardware& hw, const Mapping& a, const StagePlan& p, uint32_t n,
                uint32_t batch);

}  // namespace bfly
