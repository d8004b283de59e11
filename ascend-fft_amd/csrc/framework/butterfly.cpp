// Plan-selection framework and Plan runtime of the C++ API
// (include/butterfly/plan.hpp) for gfx950.  Every device interaction goes
// through the C ABI of include/bfly_hip.h, so the same code links against
// libbfly_hip.so (MI355X) or libbfly_stub.so (CPU-only guard tests).
//
// Reference counterparts: src/framework/butterfly.cpp of the reference --
// Hardware::load 71-93, Generator 96-166, Mapping::validate 169-183,
// StagePlan::ubBytes 194-202, estimate 226-322, DesignSpace::load 326-350,
// enumerate 383-501, rank 503-512, Plan 515-924, Context 938-1102.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

#include "bfly_hip.h"
#include "bfly_plans.hpp"
#include "butterfly/plan.hpp"
#include "butterfly/reference.hpp"
#include "json_config.hpp"

namespace bfly {

namespace {

std::string readFile(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot read configuration: " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

uint32_t ilog2(uint32_t n) {
    uint32_t l = 0;
    while ((1u << l) < n) l++;
    return l;
}

uint32_t bitrev(uint32_t v, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; i++, v >>= 1) r = (r << 1) | (v & 1u);
    return r;
}

bool pow2(uint32_t n) { return n && !(n & (n - 1)); }

// Transform envelopes of the Plan API (reference docs/user-guide/transforms.md:46-50,
// with C2C extended to the long-FFT lengths).
constexpr uint32_t kC2CMin = 64, kC2CMax = 65536;
constexpr uint32_t kR2CMin = 128, kR2CMax = 8192;
constexpr uint32_t kC2RMin = 64, kC2RMax = 4096;

void intAxis(const json::Node& obj, const char* key, std::vector<int>* out, int lo, int hi) {
    out->clear();
    for (const auto& v : obj[key].asArray()) {
        const int x = v.asInt32();
        if (x < lo || x > hi) throw std::runtime_error(std::string(key) + " holds an out-of-range value");
        out->push_back(x);
    }
}

void boolAxis(const json::Node& obj, const char* key, std::vector<int>* out) {
    out->clear();
    for (const auto& v : obj[key].asArray()) out->push_back(v.asBool() ? 1 : 0);
}

void stringAxis(const json::Node& obj, const char* key, std::vector<std::string>* out,
                std::initializer_list<const char*> allowed) {
    out->clear();
    for (const auto& v : obj[key].asArray()) {
        const std::string& s = v.asString();
        if (std::find_if(allowed.begin(), allowed.end(), [&](const char* a) { return s == a; }) ==
            allowed.end())
            throw std::runtime_error(std::string(key) + " holds an unknown option: " + s);
        out->push_back(s);
    }
}

}  // namespace

// ==================================================================== H
Hardware Hardware::load(const std::string& profilePath) {
    const json::Node root = json::parse(readFile(profilePath));
    if (root["schema_version"].asInt() != 1) throw std::runtime_error("unsupported hardware schema_version");
    Hardware hw;
    hw.soc = root["soc"].asString();
    hw.cannVersion = root["runtime_version"].asString();
    hw.aicoreNum = root["compute_units"].asInt32();
    hw.vectorCoreNum = hw.aicoreNum;
    hw.subBlockNum = root["simds_per_cu"].asInt32();
    hw.vecLaneFp32 = root["wave_size"].asInt32();
    hw.ubBytesPerCore = root["lds_bytes_per_cu"].asInt32();
    hw.ldsBytesPerWorkgroup = root["lds_bytes_per_workgroup"].asInt32();
    hw.l2Bytes = root.has("l2_bytes") ? root["l2_bytes"].asInt32() : 0;
    hw.xcds = root.has("xcds") ? root["xcds"].asInt32() : 1;
    hw.mallBytes = root.has("mall_bytes") ? root["mall_bytes"].asInt() : 0;
    hw.hbmBytes = root["global_mem_total_bytes"].asInt();
    hw.hbmPeakGBs = root["hbm_peak_gbs"].asDouble();
    hw.hbmCopyGBs = root.has("hbm_copy_gbs") ? root["hbm_copy_gbs"].asDouble() : hw.hbmPeakGBs;
    hw.simt = root.has("cross_lane_exchange") ? root["cross_lane_exchange"].asBool() : true;
    hw.matrixUnitUsable = root.has("matrix_unit_usable") ? root["matrix_unit_usable"].asBool() : false;
    hw.ubOffsetAlignFloats = 4;
    hw.minBlockOutputBytes = 0;
    if (hw.aicoreNum <= 0 || hw.subBlockNum <= 0 || hw.vecLaneFp32 <= 0 || hw.ubBytesPerCore <= 0 ||
        hw.ldsBytesPerWorkgroup <= 0 || hw.l2Bytes < 0 || hw.xcds <= 0 || hw.mallBytes < 0 ||
        hw.hbmBytes <= 0 || !(hw.hbmPeakGBs > 0) || !(hw.hbmCopyGBs > 0))
        throw std::runtime_error("hardware profile holds nonpositive resources");
    return hw;
}

// ==================================================================== G
void Generator::genTwiddles(std::vector<float>& twr, std::vector<float>& twi) const {
    twr.assign(twiddlePad(), 0.f);
    twi.assign(twiddlePad(), 0.f);
    uint32_t slot = 0;
    for (uint32_t h = 1; 2 * h <= n; h *= 2) {
        for (uint32_t j = 0; j < h; j++) {
            const double a = -M_PI * (double)j / (double)h;
            twr[slot + j] = (float)std::cos(a);
            twi[slot + j] = (float)std::sin(a);
        }
        slot += (h + 7u) / 8u * 8u;
    }
}

void Generator::genBitReverse(std::vector<uint32_t>& dst) const {
    const uint32_t bits = ilog2(n);
    dst.resize(n);
    for (uint32_t i = 0; i < n; i++) dst[i] = 2u * bitrev(i, bits);
}

void Generator::genBitReversePlane(std::vector<uint32_t>& dst) const {
    const uint32_t bits = ilog2(n), K = planeK(), rows = n / K, logK = ilog2(K);
    dst.resize(n);
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t src = ((k % rows) << logK) | (k / rows);
        dst[k] = 8u * bitrev(src, bits);
    }
}

void Generator::genTransposePlane(std::vector<uint32_t>& dst) const {
    const uint32_t K = planeK(), rows = n / K, logK = ilog2(K);
    dst.resize(n);
    for (uint32_t k = 0; k < n; k++) dst[k] = 4u * ((k & (K - 1u)) * rows + (k >> logK));
}

void Generator::genInterleave(std::vector<uint32_t>& dst, uint32_t D) const {
    D = std::max(D, 1u);
    dst.resize((size_t)2 * n * D);
    for (uint32_t d = 0; d < D; d++)
        for (uint32_t j = 0; j < n; j++) {
            dst[(size_t)d * 2 * n + 2 * j] = 4u * (d * n + j);
            dst[(size_t)d * 2 * n + 2 * j + 1] = 4u * (n * D + d * n + j);
        }
}

void Generator::genKernelTable(int kind, std::vector<float>& table) const {
    table.assign(bfly_table_floats(kind, n), 0.f);
    if (!table.empty() && bfly_table_fill(kind, n, table.data()) != BFLY_OK) table.clear();
}

// ==================================================================== A
bool Mapping::validate(const Hardware& hw, std::string* why) const {
    auto fail = [&](const char* m) {
        if (why) *why = m;
        return false;
    };
    if (udCore <= 0) return fail("ud_core must be > 0");
    if (udCore > hw.vectorCoreNum) return fail("ud_core exceeds the compute-unit count");
    if (level != 0) return fail("level!=0 (multi-GPU) not implemented");
    if (localExchange != "shared" && localExchange != "swizzle" && localExchange != "shuffle" &&
        localExchange != "register")
        return fail("unknown local_exchange");
    if ((localExchange == "shuffle") && !hw.supportsShuffleExchange())
        return fail("local_exchange=shuffle requires cross-lane exchange");
    if (ts != 0 && ts != 1) return fail("ts must be 0 or 1");
    return true;
}

// ==================================================================== P
// LDS bytes of one 256-thread workgroup: 256/T transforms (T = n/P threads
// each) of lds_len(n) = n + n/16 complex, i.e. 256 * P * 8 * 17/16 for every
// n <= 4096; 8192 and 16384 hold one transform per workgroup (68 / 136
// KiB); longer lengths take the larger of the column kernel (512 threads
// at most) and the row kernel (256 threads), C columns of lds_len(N_i) + 1
// complex each.
size_t StagePlan::ubBytes(uint32_t n) const {
    const size_t P = (size_t)std::max(radix, 2);
    auto local = [&](uint32_t m) {
        const size_t T = std::max<size_t>(1, m / P), B = std::max<size_t>(1, 256 / T);
        return B * ((size_t)m + m / 16) * 8;
    };
    if (n <= (uint32_t)bfly_plan::kMaxSingleN) return local(n);
    uint32_t n1 = 0, n2 = 0;
    if (!bfly_plan::long_split(n, &n1, &n2)) return local(n);
    auto column = [&](uint32_t m, size_t threads) {
        const size_t T = std::max<size_t>(1, m / P), C = std::max<size_t>(1, threads / T);
        return C * ((size_t)m + m / 16 + 1) * 8;
    };
    return std::max(column(n1, 512), column(n2, 256));
}

// ==================================================================== Q
// Roofline estimate, used for kernels the space carries no fitted model for:
// a fixed launch cost plus the longer of the HBM time (the chain's traffic at
// the measured streaming-copy bandwidth, over the compute units the grid may
// occupy) and the VALU time (5 n log2 n flops per transform at the
// non-packed fp32 issue rate of those compute units).
Metric estimate(const Hardware& hw, const Mapping& a, const StagePlan& p, uint32_t n,
                uint32_t batch) {
    Metric m;
    if (n == 0 || batch == 0) return m;
    (void)p;
    const int cus = std::min(std::max(a.udCore, 1), std::max(hw.aicoreNum, 1));
    const double trips = n > (uint32_t)bfly_plan::kMaxSingleN ? 2.0 : 1.0;
    const double hbmUs = 16.0 * n * (double)batch * trips / (hw.hbmCopyGBs * 1e3) * hw.aicoreNum / cus;
    const double valuUs = 5.0 * n * ilog2(n) * (double)batch / (cus * 64.0 * 2.0 * 2.4e3);
    m.etaUs = 4.0 + std::max(hbmUs, valuUs);
    return m;
}

namespace {

struct KernelInfo {
    int variant;
    const char* name;
};
// Spelling of every C2C kernel in the cost model and the selection log.
constexpr KernelInfo kKernels[] = {
    {BFLY_VARIANT_PLAIN, "plain"},
    {BFLY_VARIANT_STREAM, "stream"},
    {BFLY_VARIANT_PLAIN_SWZ, "plain_swz"},
    {BFLY_VARIANT_STREAM_SWZ, "stream_swz"},
    {BFLY_VARIANT_STREAM_LT, "stream_lt"},
    {BFLY_VARIANT_STREAM2, "stream2"},
    {BFLY_VARIANT_FOUR_STEP, "four_step"},
    {BFLY_VARIANT_SPLIT_MFMA, "split_mfma"},
    {BFLY_VARIANT_SPLIT_MFMA_STREAM, "split_mfma_stream"},
};

// Keys a cost model may carry: every kernel name, and the one-workgroup
// forms of 8192 / 16384.
bool knownModelKey(const std::string& key) {
    for (const auto& k : kKernels)
        if (key == k.name) return true;
    return key == "plain_single" || key == "stream_lt_single";
}

// Model features of (n, batch) for a kernel: 1, MB of chain traffic,
// Gpoint-log2 of butterfly work.
void features(int variant, uint32_t n, uint32_t batch, double f[3]) {
    const bool chain = variant == BFLY_VARIANT_FOUR_STEP || n > (uint32_t)bfly_plan::kMaxSingleN;
    f[0] = 1.0;
    f[1] = 16.0 * n * (double)batch * (chain ? 2.0 : 1.0) / 1e6;
    f[2] = (double)n * batch * ilog2(n) / 1e9;
}

}  // namespace

const char* kernelName(int variant) {
    for (const auto& k : kKernels)
        if (k.variant == variant) return k.name;
    return "none";
}

std::string modelKey(int variant, uint32_t n) {
    std::string key = kernelName(variant);
    const bool single = n > (uint32_t)bfly_plan::kMaxLocalN && n <= (uint32_t)bfly_plan::kMaxSingleN;
    if (single && (variant == BFLY_VARIANT_PLAIN || variant == BFLY_VARIANT_STREAM_LT)) key += "_single";
    return key;
}

double kernelEta(const DesignSpace& sp, const Hardware& hw, int variant, uint32_t n, uint32_t batch) {
    const std::string key = modelKey(variant, n);
    for (const auto& [k, model] : sp.models) {
        if (k != key || model.points == 0) continue;
        double f[3];
        features(variant, n, batch, f);
        const double eta = model.c[0] * f[0] + model.c[1] * f[1] + model.c[2] * f[2];
        if (eta > 0.0) return eta;
    }
    Mapping a;
    a.udCore = hw.aicoreNum;
    return estimate(hw, a, StagePlan{}, n, batch).etaUs;
}

// ============================================================ design space
DesignSpace DesignSpace::load(const std::string& path) {
    DesignSpace sp;
    sp.raw = readFile(path);
    const json::Node root = json::parse(sp.raw);
    if (root["schema_version"].asInt() != 1) throw std::runtime_error("unsupported design-space schema_version");
    const json::Node& A = root["A"];
    const json::Node& P = root["P"];
    const json::Node& L = root["L"];
    intAxis(A, "ud_core", &sp.udCores, 1, 1 << 20);
    intAxis(A, "ts", &sp.tsValues, 0, 1);
    if (A.has("tb")) intAxis(A, "tb", &sp.tbValues, 0, 1);
    if (A.has("ub")) intAxis(A, "ub", &sp.ubValues, 1, 2);
    stringAxis(A, "local_exchange", &sp.localExchanges, {"shared", "swizzle", "shuffle", "register"});
    intAxis(P, "radix", &sp.radices, 2, 16);
    for (int r : sp.radices)
        if (!pow2((uint32_t)r)) throw std::runtime_error("radix must be 2, 4, 8 or 16");
    intAxis(P, "fusion_level", &sp.fusionStages, 1, 4);
    boolAxis(P, "coefficient_residency", &sp.coeffResidency);
    stringAxis(L, "in", &sp.layoutsIn, {"interleaved", "planar"});
    stringAxis(L, "out", &sp.layoutsOut, {"interleaved", "planar"});
    if (root.has("F")) boolAxis(root["F"], "dft_matrix", &sp.dftMatrix);
    if (root.has("Q") && root["Q"].has("cost_model")) {
        // the fitted model lives in its own file, named relative to this one
        const size_t slash = path.find_last_of('/');
        const std::string dir = slash == std::string::npos ? std::string() : path.substr(0, slash + 1);
        const json::Node cm = json::parse(readFile(dir + root["Q"]["cost_model"].asString()));
        for (const auto& [name, node] : cm["kernels"].asObject()) {
            if (!knownModelKey(name)) throw std::runtime_error("cost_model names an unknown kernel: " + name);
            KernelModel m;
            const auto& c = node["c"].asArray();
            if (c.size() != 3) throw std::runtime_error("cost_model coefficients come in threes");
            for (int i = 0; i < 3; i++) m.c[i] = c[i].asDouble();
            m.loo = node["loo"].asDouble();
            m.points = node["points"].asInt32();
            if (m.points < 3 || !std::isfinite(m.loo))
                throw std::runtime_error("cost_model entry needs >= 3 points and a finite loo: " + name);
            sp.models.push_back({name, m});
        }
    }
    sp.applyDefaults();
    return sp;
}

void DesignSpace::applyDefaults() {
    if (udCores.empty()) udCores = {256};
    if (radices.empty()) radices = {16};
    if (localExchanges.empty()) localExchanges = {"shared"};
    if (fusionStages.empty()) fusionStages = {1};
    if (layoutsIn.empty()) layoutsIn = {"interleaved"};
    if (layoutsOut.empty()) layoutsOut = {"interleaved"};
    if (tsValues.empty()) tsValues = {1};
    if (tbValues.empty()) tbValues = {1};
    if (ubValues.empty()) ubValues = {1};
    if (coeffResidency.empty()) coeffResidency = {1};
    if (dftMatrix.empty()) dftMatrix = {0};
}

int DesignSpace::pointSize(int radix, int fusionLevel) {
    long long R = 1;
    const long long base = std::max(2, radix);
    for (int i = 0; i < std::max(1, fusionLevel) && R <= (1LL << 24); i++) R *= base;
    return (int)R;
}

// ================================================================ lowering
const char* toString(State s) {
    static const char* const names[] = {"Infeasible", "Unverified", "Feasible", "Measured"};
    const unsigned i = static_cast<unsigned>(s);
    return i < 4 ? names[i] : "?";
}

// The kernel a candidate runs at length n.  Axis rules first (they hold at
// every n), then the kernel table of the length class:
//
//   n <= 4096, ts 1   dft 0: tb 1 -> PLAIN / PLAIN_SWZ (device-memory twiddles),
//                            tb 0 -> STREAM / STREAM_SWZ, with LDS twiddles
//                            STREAM_LT, two tiles in flight STREAM2;
//                     dft 1 (1024 / 2048 / 4096): tb 1 -> SPLIT_MFMA,
//                            tb 0 -> SPLIT_MFMA_STREAM (LDS twiddle tables)
//   8192, 16384 ts 1  tb 1 device-memory twiddles -> PLAIN (one workgroup per
//                     transform), tb 0 LDS twiddles -> STREAM_LT
//   8192 .. 65536 ts 0  -> FOUR_STEP (column and row kernels, LDS-staged
//                     twiddles, one group per workgroup)
int lowerC2C(const Candidate& c, const Hardware& hw, uint32_t n, std::string* why) {
    char buf[256];
    auto no = [&](const char* fmt, auto... args) {
        std::snprintf(buf, sizeof(buf), fmt, args...);
        if (why) *why = buf;
        return 0;
    };
    std::string v;
    if (!c.a.validate(hw, &v)) return no("%s", v.c_str());
    if (c.a.udCore != hw.aicoreNum)
        return no("ud_core %d: every grid is sized to the resident slots of all %d compute units", c.a.udCore,
                  hw.aicoreNum);
    if (c.p.radix != bfly_plan::kPoints)
        return no("radix %d has no compiled pass plan (radix-16 passes, a 2/4/8 remainder pass)", c.p.radix);
    if (c.p.fusionLevel != 1 || c.f.stageCount != 1 || c.f.butterflyCount != 1)
        return no("fusion_level %d: the kernels run one radix pass per LDS exchange", c.p.fusionLevel);
    if (c.p.bitReverseInput) return no("%s", "Stockham autosort passes read natural-order input");
    if (!c.l.contiguous || c.l.in != Layout::Interleaved || c.l.out != Layout::Interleaved)
        return no("%s", "the kernels read and write contiguous interleaved complex rows (no planar pass)");
    if (c.a.td != 1 || c.a.us != 1) return no("%s", "one tile per pass and one stage service are compiled");
    if (c.a.localExchange != "shared" && c.a.localExchange != "swizzle")
        return no("local_exchange=%s has no kernel: exchanges go through LDS", c.a.localExchange.c_str());
    if (c.a.tb != 0 && c.a.tb != 1) return no("tb %d: one-shot (1) or streaming (0) grids only", c.a.tb);
    if (c.a.ub != 1 && c.a.ub != 2) return no("ub %d: one or two tiles in flight", c.a.ub);
    if (n < 2 || !pow2(n)) return no("n=%u is not a power of two (Stockham radix plans)", n);
    if (n < kC2CMin)
        return no("n=%u below the C2C envelope: %u rows of 16 points, the envelope starts at n=%u", n, n / 16u,
                  kC2CMin);
    if (n > kC2CMax)
        return no("n=%u above the long-FFT envelope: both four-step factors must stay LDS-resident (%d KiB per CU)",
                  n, hw.ubBytesPerCore / 1024);
    const bool shared = c.a.localExchange == "shared", lds = c.p.coefficientResidency;
    const bool oneShot = c.a.tb == 1, mfma = c.f.dftMatrix;
    if (c.a.ts == 0) {
        if (n <= (uint32_t)bfly_plan::kMaxLocalN)
            return no("n=%u fits one workgroup's LDS: the four-step chain (ts=0) starts at 8192", n);
        if (mfma) return no("%s", "the split-FP16 MFMA passes exist for n = 1024 / 2048 / 4096 only");
        if (!shared || !lds || !oneShot || c.a.ub != 1)
            return no("%s", "the four-step chain is compiled with the padded exchange, LDS-staged twiddles and one "
                      "group per workgroup (shared, coefficient_residency, tb 1, ub 1)");
        return BFLY_VARIANT_FOUR_STEP;
    }
    if (n > (uint32_t)bfly_plan::kMaxSingleN)
        return no("n=%u exceeds one workgroup's LDS (%d KiB per CU): it runs as the four-step chain (ts=0)", n,
                  hw.ubBytesPerCore / 1024);
    if (!c.p.fits(n, hw))
        return no("LDS %.1f KiB > %d KiB (n=%u)", c.p.ubBytes(n) / 1024.0, hw.ubBytesPerCore / 1024, n);
    if (n > (uint32_t)bfly_plan::kMaxLocalN) {
        if (mfma) return no("%s", "the split-FP16 MFMA passes exist for n = 1024 / 2048 / 4096 only");
        if (!shared || c.a.ub != 1)
            return no("%s", "one workgroup per transform is compiled with the padded exchange and one tile in flight");
        if (oneShot && !lds) return BFLY_VARIANT_PLAIN;
        if (!oneShot && lds) return BFLY_VARIANT_STREAM_LT;
        return no("%s", "one workgroup per transform exists as a one-shot grid with device-memory twiddles (tb 1) "
                  "or a streaming grid with LDS twiddles (tb 0)");
    }
    if (mfma) {
        if (n != 1024 && n != 2048 && n != 4096)
            return no("%s", "the split-FP16 MFMA passes exist for n = 1024 / 2048 / 4096 only");
        if (!shared || !lds || c.a.ub != 1)
            return no("%s", "the MFMA kernels use the padded exchange, keep their twiddle tables in LDS and run one "
                      "tile in flight");
        return oneShot ? BFLY_VARIANT_SPLIT_MFMA : BFLY_VARIANT_SPLIT_MFMA_STREAM;
    }
    if (c.a.ub == 2) {
        if (oneShot || !shared || lds)
            return no("%s", "two tiles in flight exist as a streaming kernel (tb 0) with the padded exchange and "
                      "device-memory twiddles");
        return BFLY_VARIANT_STREAM2;
    }
    if (lds) {
        if (oneShot || !shared)
            return no("%s", "LDS-resident twiddles exist in the streaming kernel (tb 0) with the padded exchange");
        return BFLY_VARIANT_STREAM_LT;
    }
    if (oneShot) return shared ? BFLY_VARIANT_PLAIN : BFLY_VARIANT_PLAIN_SWZ;
    return shared ? BFLY_VARIANT_STREAM : BFLY_VARIANT_STREAM_SWZ;
}

// ============================================================== enumerate
namespace {

// One point of the space per odometer reading: digit i indexes axis i.
struct Odometer {
    std::vector<size_t> sizes, digit;
    explicit Odometer(std::vector<size_t> s) : sizes(std::move(s)), digit(sizes.size(), 0) {}
    bool empty() const {
        return std::any_of(sizes.begin(), sizes.end(), [](size_t z) { return z == 0; });
    }
    bool next() {  // false after the last reading
        for (size_t i = sizes.size(); i-- > 0;) {
            if (++digit[i] < sizes[i]) return true;
            digit[i] = 0;
        }
        return false;
    }
};

std::string candidateId(const Candidate& c) {
    char buf[160];
    std::snprintf(buf, sizeof(buf), "r%d_ud%d_ts%d_%s_f%d_R%d", c.p.radix, c.a.udCore, c.a.ts,
                  c.a.localExchange.c_str(), c.p.fusionLevel, DesignSpace::pointSize(c.p.radix, c.p.fusionLevel));
    std::string id = buf;
    if (c.a.tb == 0) id += "_stream";
    if (c.a.ub == 2) id += "_x2";
    if (c.l.in == Layout::Planar) id += "_inplanar";
    if (c.l.out == Layout::Planar) id += "_outplanar";
    if (!c.p.coefficientResidency) id += "_noRes";
    if (c.f.dftMatrix) id += "_mfma";
    return id;
}

}  // namespace

std::vector<Candidate> enumerate(const DesignSpace& sp, const Hardware& hw, uint32_t n,
                                 uint32_t batch) {
    DesignSpace s = sp;
    s.applyDefaults();
    Odometer odo({s.udCores.size(), s.radices.size(), s.fusionStages.size(), s.tsValues.size(),
                  s.tbValues.size(), s.ubValues.size(), s.coeffResidency.size(), s.localExchanges.size(),
                  s.layoutsIn.size(), s.layoutsOut.size(), s.dftMatrix.size()});
    std::vector<Candidate> out;
    if (odo.empty()) return out;
    const uint64_t floats = (uint64_t)batch * 2ull * n;
    do {
        const auto& d = odo.digit;
        Candidate c;
        c.a.udCore = s.udCores[d[0]];
        c.p.radix = s.radices[d[1]];
        c.p.fusionLevel = s.fusionStages[d[2]];
        c.f.stageCount = c.p.fusionLevel;
        c.a.ts = s.tsValues[d[3]];
        c.a.tb = s.tbValues[d[4]];
        c.a.ub = s.ubValues[d[5]];
        c.p.coefficientResidency = s.coeffResidency[d[6]] != 0;
        c.a.localExchange = s.localExchanges[d[7]];
        c.l.in = s.layoutsIn[d[8]] == "planar" ? Layout::Planar : Layout::Interleaved;
        c.l.out = s.layoutsOut[d[9]] == "planar" ? Layout::Planar : Layout::Interleaved;
        c.f.dftMatrix = s.dftMatrix[d[10]] != 0;
        c.id = candidateId(c);
        std::string why;
        c.kernel = lowerC2C(c, hw, n, &why);
        if (!c.kernel) {
            c.reason = why;
        } else if (batch == 0) {
            c.kernel = 0;
            c.reason = "batch must be positive";
        } else if (floats > 0xFFFFFFFFull) {
            char buf[200];
            std::snprintf(buf, sizeof(buf),
                          "batch*2*n = %llu exceeds the uint32 element-offset contract (max batch for n=%u is %llu)",
                          (unsigned long long)floats, n, 0xFFFFFFFFull / (2ull * n));
            c.kernel = 0;
            c.reason = buf;
        }
        c.state = c.kernel ? State::Unverified : State::Infeasible;
        c.q.etaUs = c.kernel ? kernelEta(s, hw, c.kernel, n, batch) : 0.0;
        out.push_back(std::move(c));
    } while (odo.next());
    return out;
}

std::vector<const Candidate*> rank(const std::vector<Candidate>& cs) {
    std::vector<const Candidate*> v;
    v.reserve(cs.size());
    for (const auto& c : cs)
        if (c.state != State::Infeasible) v.push_back(&c);
    // measured evidence outranks the model; within a state the cheaper eta first
    std::stable_sort(v.begin(), v.end(), [](const Candidate* x, const Candidate* y) {
        return std::make_pair(-static_cast<int>(x->state), x->q.etaUs) <
               std::make_pair(-static_cast<int>(y->state), y->q.etaUs);
    });
    return v;
}

// =================================================================== Plan
struct Plan::Impl {
    Hardware hardware;
    void* devIn = nullptr;
    void* devOut = nullptr;
    void* devWork = nullptr;
    size_t inCap = 0, outCap = 0, workCap = 0;
    void* devTable[3] = {nullptr, nullptr, nullptr};
    bool prepared[3] = {false, false, false};
    uint32_t preparedN[3] = {0, 0, 0};
};

Plan::Plan(Candidate c) : cand_(std::move(c)) {}

Plan::~Plan() {
    if (!impl_) return;
    bfly_rt_free(impl_->devIn);
    bfly_rt_free(impl_->devOut);
    bfly_rt_free(impl_->devWork);
    for (void* t : impl_->devTable) bfly_rt_free(t);
}

static int ensureBuf(void** p, size_t* cap, size_t bytes) {
    if (*cap >= bytes && *p) return 0;
    if (*p) bfly_rt_free(*p);
    *p = nullptr;
    *cap = 0;
    if (bfly_rt_malloc(p, bytes) != BFLY_OK) {
        *p = nullptr;
        return -1;
    }
    *cap = bytes;
    return 0;
}

// C2C runs the kernel the candidate lowers to at n (none: the shape guard);
// R2C and C2R run the library's real kernels for every candidate.
bool Plan::validShape(int kind, uint32_t n, uint32_t batch, bool checkOffset) const {
    if (!impl_ || batch == 0 || !pow2(n)) return false;
    switch (kind) {
        case BFLY_KIND_C2C:
            if (!lowerC2C(cand_, impl_->hardware, n, nullptr)) return false;
            break;
        case BFLY_KIND_R2C:
            if (n < kR2CMin || n > kR2CMax) return false;
            break;
        case BFLY_KIND_C2R:
            if (n < kC2RMin || n > kC2RMax) return false;
            break;
        default:
            return false;
    }
    if (!bfly_supported(kind, n)) return false;
    return !checkOffset || (uint64_t)batch * 2ull * n <= 0xFFFFFFFFull;
}

int Plan::prepare(uint32_t n, uint32_t batch) { return prepareKind(BFLY_KIND_C2C, n, batch); }

int Plan::prepareKind(int kind, uint32_t n, uint32_t batch) {
    // prepare's default batch is a sentinel, not an execution shape.
    if (!validShape(kind, n, batch, false)) return -8;
    if (impl_->prepared[kind] && impl_->preparedN[kind] == n) return 0;
    // Invalidate before touching the old table: a failed rebuild must never
    // leave the previous key looking valid.
    impl_->prepared[kind] = false;
    impl_->preparedN[kind] = 0;
    Generator g;
    g.n = n;
    std::vector<float> table;
    g.genKernelTable(kind, table);
    if (table.empty()) return -2;
    if (impl_->devTable[kind]) bfly_rt_free(impl_->devTable[kind]);
    impl_->devTable[kind] = nullptr;
    if (bfly_rt_malloc(&impl_->devTable[kind], table.size() * sizeof(float)) != BFLY_OK) {
        impl_->devTable[kind] = nullptr;
        return -3;
    }
    if (bfly_rt_memcpy_h2d(impl_->devTable[kind], table.data(), table.size() * sizeof(float),
                           ctxStream_) != BFLY_OK)
        return -4;
    impl_->prepared[kind] = true;
    impl_->preparedN[kind] = n;
    return 0;
}

// Launch one transform on device buffers with the kernel this plan runs at n.
static int launchKind(const Candidate& c, const Hardware& hw, int kind, int sign, const float* in, float* out,
                      uint32_t n, uint32_t batch, const float* table, void* work, void* stream) {
    if (kind == BFLY_KIND_C2C)
        return bfly_c2c_ex(in, out, n, batch, sign, lowerC2C(c, hw, n, nullptr), table, work, stream);
    if (kind == BFLY_KIND_R2C) return bfly_r2c(in, out, n, batch, table, stream);
    return bfly_c2r(in, out, n, batch, table, stream);
}

static size_t workBytes(const Candidate& c, const Hardware& hw, int kind, uint32_t n, uint32_t batch) {
    if (kind != BFLY_KIND_C2C) return 0;
    return bfly_c2c_launches(n, batch, lowerC2C(c, hw, n, nullptr)) == 2 ? (size_t)8 * n * batch : 0;
}

int Plan::execute(int kind, int sign, const float* in, float* out, uint32_t n, uint32_t batch) {
    if (!in || !out || !validShape(kind, n, batch)) return -8;
    if (prepareKind(kind, n, batch)) return -2;
    const size_t rowIn = kind == BFLY_KIND_C2C ? 2 * (size_t)n : kind == BFLY_KIND_R2C ? n : n + 2;
    const size_t rowOut = kind == BFLY_KIND_C2C ? 2 * (size_t)n : kind == BFLY_KIND_R2C ? n + 2 : n;
    const size_t inB = rowIn * batch * sizeof(float), outB = rowOut * batch * sizeof(float);
    const int variant =
        kind == BFLY_KIND_C2C ? lowerC2C(cand_, impl_->hardware, n, nullptr) : BFLY_VARIANT_AUTO;
    // AUTO picks a long C2C's kernel per chunk, which may be the four-step chain
    size_t ws = workBytes(cand_, impl_->hardware, kind, n, batch);
    if (kind == BFLY_KIND_C2C && variant == BFLY_VARIANT_AUTO && n > (uint32_t)bfly_plan::kMaxLocalN)
        ws = (size_t)8 * n * batch;
    if (ensureBuf(&impl_->devIn, &impl_->inCap, inB) || ensureBuf(&impl_->devOut, &impl_->outCap, outB) ||
        (ws && ensureBuf(&impl_->devWork, &impl_->workCap, ws)))
        return -3;
    // upload, transform and download pipelined in chunks over the link (the
    // reference does the three in turn); the library's status codes are the
    // reference's -1..-8
    return bfly_host_pipelined(kind, sign, variant, in, out, n, batch,
                               static_cast<const float*>(impl_->devTable[kind]), impl_->devIn, impl_->devOut,
                               impl_->devWork, 0, ctxStream_);
}

int Plan::run(const float* in, float* out, uint32_t n, uint32_t batch) {
    return execute(BFLY_KIND_C2C, -1, in, out, n, batch);
}

int Plan::runInverse(const float* in, float* out, uint32_t n, uint32_t batch) {
    return execute(BFLY_KIND_C2C, 1, in, out, n, batch);
}

int Plan::runR2C(const float* in, float* out, uint32_t n, uint32_t batch) {
    return execute(BFLY_KIND_R2C, -1, in, out, n, batch);
}

int Plan::runC2R(const float* in, float* out, uint32_t n, uint32_t batch) {
    return execute(BFLY_KIND_C2R, 1, in, out, n, batch);
}

int Plan::runDevice(const float* devIn, float* devOut, uint32_t n, uint32_t batch, int sign) {
    if (!devIn || !devOut || (sign != 1 && sign != -1) || !validShape(BFLY_KIND_C2C, n, batch)) return -8;
    if (prepareKind(BFLY_KIND_C2C, n, batch)) return -2;
    const size_t ws = workBytes(cand_, impl_->hardware, BFLY_KIND_C2C, n, batch);
    if (ws && ensureBuf(&impl_->devWork, &impl_->workCap, ws)) return -3;
    const int rc = launchKind(cand_, impl_->hardware, BFLY_KIND_C2C, sign, devIn, devOut, n, batch,
                              static_cast<const float*>(impl_->devTable[BFLY_KIND_C2C]), impl_->devWork,
                              ctxStream_);
    return rc == BFLY_OK ? 0 : rc == BFLY_ERR_SHAPE ? -8 : -5;
}

// Per-launch microseconds of this plan's kernel on the resident buffers:
// HIP events around `reps` back-to-back launches, best of `rounds`.
static int timeKernel(const Candidate& c, const Hardware& hw, const float* in, float* out, uint32_t n,
                      uint32_t batch, const float* table, void* work, void* stream, int reps, int rounds,
                      double* us) {
    void *e0 = nullptr, *e1 = nullptr;
    if (bfly_rt_event_create(&e0) != BFLY_OK || bfly_rt_event_create(&e1) != BFLY_OK) {
        if (e0) bfly_rt_event_destroy(e0);
        return -5;
    }
    int rc = BFLY_OK;
    double best = 0.0;
    for (int r = 0; r < rounds && rc == BFLY_OK; r++) {
        rc = bfly_rt_event_record(e0, stream);
        for (int i = 0; i < reps && rc == BFLY_OK; i++)
            rc = launchKind(c, hw, BFLY_KIND_C2C, -1, in, out, n, batch, table, work, stream);
        if (rc == BFLY_OK) rc = bfly_rt_event_record(e1, stream);
        float ms = 0.f;
        if (rc == BFLY_OK) rc = bfly_rt_event_elapsed(e0, e1, &ms);
        const double t = 1000.0 * ms / reps;
        if (rc == BFLY_OK && (r == 0 || t < best)) best = t;
    }
    bfly_rt_event_destroy(e0);
    bfly_rt_event_destroy(e1);
    *us = best;
    return rc == BFLY_OK ? 0 : -5;
}

int Plan::measure(uint32_t n, uint32_t batch, Metric* m) {
    if (!m) return -1;
    if (!validShape(BFLY_KIND_C2C, n, batch)) return -8;
    // the reference's pattern input, one checked run, then the timing
    std::vector<float> in((size_t)2 * n * batch), got(in.size());
    for (size_t i = 0; i < in.size(); i++) in[i] = patternAt(i);
    if (const int rc = run(in.data(), got.data(), n, batch)) return rc;
    Metric q;
    q.etaUs = cand_.q.etaUs;
    uint32_t rows[8];
    const uint32_t k = sampleBatches(batch, 8, rows, 8);
    std::vector<float> want(2 * (size_t)n);
    for (uint32_t i = 0; i < k; i++) {
        refFftF32(in.data() + (size_t)rows[i] * 2 * n, want.data(), n);
        q.maxRelErr = std::max(q.maxRelErr, maxRelScaled(got.data() + (size_t)rows[i] * 2 * n, want.data(), 2 * n));
    }
    q.measuredOk = std::isfinite(q.maxRelErr) && q.maxRelErr <= 1e-4;
    if (!q.measuredOk) {
        *m = q;
        cand_.q = q;
        cand_.state = State::Unverified;
        cand_.reason = "correctness check failed (maxRelScaled > 1e-4 against the FP64 reference)";
        return -7;
    }
    const int rc = timeKernel(cand_, impl_->hardware, static_cast<const float*>(impl_->devIn),
                              static_cast<float*>(impl_->devOut), n, batch,
                              static_cast<const float*>(impl_->devTable[BFLY_KIND_C2C]), impl_->devWork,
                              ctxStream_, 10, 3, &q.measuredUs);
    if (rc) return rc;
    *m = q;
    cand_.q = q;
    cand_.reason.clear();
    cand_.state = State::Measured;
    return 0;
}

// ================================================================ Context
Context::Context() = default;

Context::~Context() {
    if (stream_) bfly_rt_stream_destroy(stream_);
}

int Context::init(const std::string& profilePath, const std::string& spacePath,
                  const std::string& kernelPath) {
    if (initialized_) return -9;
    try {
        Hardware hw = Hardware::load(profilePath);
        DesignSpace sp = DesignSpace::load(spacePath);
        hw_ = std::move(hw);
        sp_ = std::move(sp);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Context::init configuration error: %s\n", e.what());
        return -4;
    }
    kernelPath_ = kernelPath;
    switch (bfly_rt_init(0)) {
        case BFLY_OK: break;
        case BFLY_ERR_NOT_LOADED: return -1;
        default: return -2;
    }
    if (bfly_rt_stream_create(&stream_) != BFLY_OK) return -3;
    initialized_ = true;
    return 0;
}

std::vector<Candidate> Context::enumerate(uint32_t n, uint32_t batch) const {
    return bfly::enumerate(sp_, hw_, n, batch);
}

// A plan for any candidate that lowers to some kernel at some C2C length of
// the envelope; the kernel itself is chosen per call from the length.
std::unique_ptr<Plan> Context::makePlan(const Candidate& c) {
    if (!initialized_ || c.state == State::Infeasible) return nullptr;
    bool lowers = false;
    for (uint32_t n = kC2CMin; n <= kC2CMax && !lowers; n *= 2) lowers = lowerC2C(c, hw_, n, nullptr) != 0;
    if (!lowers) return nullptr;
    auto p = std::unique_ptr<Plan>(new Plan(c));
    p->impl_.reset(new Plan::Impl());
    p->impl_->hardware = hw_;
    p->ctxStream_ = stream_;
    p->tf_.push_back({Transform::GenTwiddles, "FP64 twiddle tables in the Stockham pass layout"});
    p->tf_.push_back({Transform::GenBitReverse, "none: autosort passes need no permutation table"});
    p->tf_.push_back({Transform::LaunchKernel, "the candidate's kernel on the context stream"});
    p->tf_.push_back({Transform::Measure, "HIP events around back-to-back launches"});
    return p;
}

// Selection: the model ranks every lowered candidate; the first topK
// distinct kernels in that order are built and measured (candidates that
// lower to the same kernel share its measurement); the fastest kernel that
// passes the correctness gate is returned.
std::unique_ptr<Plan> Context::select(uint32_t n, uint32_t batch, int topK, std::string* log) {
    const std::vector<Candidate> all = enumerate(n, batch);
    const std::vector<const Candidate*> order = rank(all);
    const size_t quota = (size_t)std::max(1, topK);
    std::string text;
    char line[400];
    auto say = [&](const char* fmt, auto... args) {
        std::snprintf(line, sizeof(line), fmt, args...);
        text += line;
    };
    std::vector<int> kernels;
    for (const Candidate* c : order)
        if (std::find(kernels.begin(), kernels.end(), c->kernel) == kernels.end()) kernels.push_back(c->kernel);
    say("select n=%u batch=%u: %zu candidates, %zu infeasible, %zu lowered onto %zu kernels\n", n, batch,
        all.size(), all.size() - order.size(), order.size(), kernels.size());
    std::map<std::string, int> why;
    for (const auto& c : all)
        if (c.state == State::Infeasible) why[c.reason]++;
    std::vector<std::pair<int, std::string>> common;
    for (const auto& [r, k] : why) common.push_back({-k, r});
    std::sort(common.begin(), common.end());
    for (size_t i = 0; i < common.size() && i < 5; i++)
        say("  infeasible x%-4d %s\n", -common[i].first, common[i].second.c_str());
    say("  model order (eta, first %zu kernels are measured):\n", quota);
    for (size_t i = 0; i < kernels.size(); i++) {
        const Candidate* first = *std::find_if(order.begin(), order.end(),
                                               [&](const Candidate* c) { return c->kernel == kernels[i]; });
        const auto aliases = std::count_if(order.begin(), order.end(),
                                           [&](const Candidate* c) { return c->kernel == kernels[i]; });
        say("    %2zu. %-18s eta=%8.1f us  %ld candidates, first %s\n", i + 1, kernelName(kernels[i]),
            first->q.etaUs, (long)aliases, first->id.c_str());
    }
    say("%s", "  measured:\n");
    std::unique_ptr<Plan> best;
    for (size_t i = 0; i < kernels.size() && i < quota; i++) {
        const Candidate* c = *std::find_if(order.begin(), order.end(),
                                           [&](const Candidate* x) { return x->kernel == kernels[i]; });
        auto plan = makePlan(*c);
        Metric m;
        const int rc = plan ? plan->measure(n, batch, &m) : -1;
        if (rc) {
            say("    %-18s rc=%d maxRel=%.2e  rejected (%s)\n", kernelName(c->kernel), rc, m.maxRelErr,
                plan ? plan->candidate().reason.c_str() : "no plan");
            continue;
        }
        say("    %-18s eta=%8.1f us  measured=%8.2f us  maxRel=%.1e  [%s] %s\n", kernelName(c->kernel), m.etaUs,
            m.measuredUs, m.maxRelErr, toString(plan->candidate().state), c->id.c_str());
        if (!best || m.measuredUs < best->candidate().q.measuredUs) best = std::move(plan);
    }
    if (best)
        say("  selected: %s  kernel %s  (%.2f us)\n", best->candidate().id.c_str(),
            kernelName(best->candidate().kernel), best->candidate().q.measuredUs);
    else
        say("%s", "  selected: none (no measured kernel passed the gate)\n");
    if (log) *log = text;
    return best;
}

}  // namespace bfly
