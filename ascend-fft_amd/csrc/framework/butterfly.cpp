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

// Radices with compiled kernels: the points-per-thread of the pass plans.
bool radixLowered(int radix) { return radix == bfly_plan::kPoints; }

bool canonicalLowering(const Candidate& c, const Hardware& hw) {
    return c.a.validate(hw, nullptr) && c.a.localExchange == "shared" && c.a.td == 1 &&
           c.a.tb == 1 && c.a.ub == 1 && c.a.us == 1 && c.a.ts == 1 && c.a.level == 0 &&
           radixLowered(c.p.radix) && c.p.fusionLevel == 1 && c.p.coefficientResidency &&
           !c.p.bitReverseInput && c.l.contiguous && c.l.in == Layout::Interleaved &&
           c.l.out == Layout::Interleaved && c.f.butterflyCount == 1 && c.f.stageCount == 1 &&
           !c.f.dftMatrix;
}

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
// eta = launch + max(HBM time, VALU time).  HBM time = payload bytes at the
// measured streaming-copy bandwidth; VALU time = 5 n log2 n flops per
// transform at the non-packed fp32 rate (64 lanes x 2 flops x CUs x clock)
// derated by the measured FFT issue efficiency.  Constants fitted to the
// C2C 4096x4096 profile (63.7 us, 13.4 M VALU instructions per launch).
Metric estimate(const Hardware& hw, const Mapping& a, const StagePlan& p, uint32_t n,
                uint32_t batch) {
    Metric m;
    if (n == 0 || batch == 0) return m;
    const double launchUs = 4.0;
    const double clockGHz = 2.4, valuEff = 0.38;
    const int cus = std::min(std::max(a.udCore, 1), std::max(hw.aicoreNum, 1));
    const double roundTrips = n > (uint32_t)bfly_plan::kMaxSingleN ? 2.0 : 1.0;
    const double bytes = 16.0 * n * (double)batch * roundTrips;
    const double memUs = bytes / (hw.hbmCopyGBs * 1e3) * ((double)hw.aicoreNum / cus);
    const double flops = 5.0 * n * ilog2(n) * (double)batch;
    const double valuRate = cus * 64.0 * 2.0 * clockGHz * 1e3 * valuEff;  // flop per us
    const int passes = (int)std::ceil(ilog2(n) / std::log2((double)std::max(p.radix, 2)));
    const double ldsUs = 0.6 * std::max(0, passes - 1) * 16.0 * n * (double)batch / (hw.hbmCopyGBs * 1e3 * 8.0);
    m.etaUs = launchUs + std::max(memUs, flops / valuRate + ldsUs);
    return m;
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
    stringAxis(A, "local_exchange", &sp.localExchanges, {"shared", "swizzle", "shuffle", "register"});
    intAxis(P, "radix", &sp.radices, 2, 16);
    for (int r : sp.radices)
        if (!pow2((uint32_t)r)) throw std::runtime_error("radix must be 2, 4, 8 or 16");
    intAxis(P, "fusion_level", &sp.fusionStages, 1, 4);
    boolAxis(P, "coefficient_residency", &sp.coeffResidency);
    stringAxis(L, "in", &sp.layoutsIn, {"interleaved", "planar"});
    stringAxis(L, "out", &sp.layoutsOut, {"interleaved", "planar"});
    if (root.has("F")) boolAxis(root["F"], "dft_matrix", &sp.dftMatrix);
    else sp.dftMatrix = {0};
    return sp;
}

void DesignSpace::applyDefaults() {
    if (udCores.empty()) udCores = {256};
    if (radices.empty()) radices = {16, 8};
    if (localExchanges.empty()) localExchanges = {"shared"};
    if (fusionStages.empty()) fusionStages = {1};
    if (layoutsIn.empty()) layoutsIn = {"interleaved"};
    if (layoutsOut.empty()) layoutsOut = {"interleaved"};
    if (tsValues.empty()) tsValues = {1};
    if (coeffResidency.empty()) coeffResidency = {1};
    if (dftMatrix.empty()) dftMatrix = {0};
}

int DesignSpace::pointSize(int radix, int fusionLevel) {
    long long R = 1;
    const long long base = std::max(2, radix);
    for (int i = 0; i < std::max(1, fusionLevel) && R <= (1LL << 24); i++) R *= base;
    return (int)R;
}

// ============================================================== enumerate
const char* toString(State s) {
    switch (s) {
        case State::Infeasible: return "Infeasible";
        case State::Unverified: return "Unverified";
        case State::Feasible: return "Feasible";
        case State::Measured: return "Measured";
    }
    return "?";
}

std::vector<Candidate> enumerate(const DesignSpace& sp, const Hardware& hw, uint32_t n,
                                 uint32_t batch) {
    std::vector<Candidate> out;
    char buf[256];
    for (int ud : sp.udCores)
        for (int radix : sp.radices)
            for (int fs : sp.fusionStages)
                for (int ts : sp.tsValues)
                    for (int cr : sp.coeffResidency)
                        for (const auto& lx : sp.localExchanges)
                            for (const auto& li : sp.layoutsIn)
                                for (const auto& lo : sp.layoutsOut)
                                    for (int dm : sp.dftMatrix) {
        Candidate c;
        c.a.udCore = ud;
        c.a.ts = ts;
        c.a.localExchange = lx;
        c.p.radix = radix;
        c.p.fusionLevel = fs;
        c.p.coefficientResidency = cr != 0;
        c.p.bitReverseInput = false;
        c.l.in = li == "planar" ? Layout::Planar : Layout::Interleaved;
        c.l.out = lo == "planar" ? Layout::Planar : Layout::Interleaved;
        c.f.stageCount = fs;
        c.f.dftMatrix = dm != 0;
        const int R = DesignSpace::pointSize(radix, fs);
        std::snprintf(buf, sizeof(buf), "r%d_ud%d_ts%d_%s_f%d_R%d", radix, ud, ts, lx.c_str(), fs, R);
        c.id = buf;
        if (c.l.in == Layout::Planar) c.id += "_inplanar";
        if (c.l.out == Layout::Planar) c.id += "_outplanar";
        if (!c.p.coefficientResidency) c.id += "_noRes";
        if (c.f.dftMatrix) c.id += "_mfma";

        std::string why;
        const uint64_t elems = (uint64_t)batch * 2ull * n;
        if (!c.a.validate(hw, &why)) {
            c.state = State::Infeasible;
            c.reason = why;
        } else if (!c.p.coefficientResidency) {
            c.state = State::Infeasible;
            c.reason = "kernel requires the twiddle table resident in cache "
                       "(coefficient_residency=false would recompute twiddles every pass)";
        } else if (ts != 1) {
            c.state = State::Infeasible;
            c.reason = "kernel keeps every pass of a transform resident in LDS "
                       "(ts=0 needs an HBM round trip per pass)";
        } else if (n < 2 || !pow2(n)) {
            c.state = State::Infeasible;
            c.reason = "n must be a power of two (Stockham radix plans)";
        } else if (n < kC2CMin) {
            c.state = State::Infeasible;
            std::snprintf(buf, sizeof(buf),
                          "n below the C2C envelope: rows=n/16=%u threads per transform, "
                          "the envelope starts at n=%u",
                          n / 16u, kC2CMin);
            c.reason = buf;
        } else if (n > kC2CMax) {
            c.state = State::Infeasible;
            std::snprintf(buf, sizeof(buf),
                          "n=%u above the long-FFT envelope: two-factor split needs both "
                          "factors <= %d to stay LDS-resident (LDS per CU %d KiB)",
                          n, bfly_plan::kMaxLocalN, hw.ubBytesPerCore / 1024);
            c.reason = buf;
        } else if (!c.p.fits(n, hw)) {
            c.state = State::Infeasible;
            std::snprintf(buf, sizeof(buf), "LDS %.1f KiB > %d KiB (radix=%d, n=%u)",
                          c.p.ubBytes(n) / 1024.0, hw.ubBytesPerCore / 1024, radix, n);
            c.reason = buf;
        } else if (R < 2 || !pow2((uint32_t)R)) {
            c.state = State::Infeasible;
            c.reason = "butterfly point size R=radix^fusionLevel is not a power of two";
        } else if (n % (uint32_t)R != 0) {
            c.state = State::Infeasible;
            std::snprintf(buf, sizeof(buf), "R=%d does not divide n=%u", R, n);
            c.reason = buf;
        } else if (batch == 0) {
            c.state = State::Infeasible;
            c.reason = "batch must be positive";
        } else if (elems > 0xFFFFFFFFull) {
            c.state = State::Infeasible;
            std::snprintf(buf, sizeof(buf),
                          "batch*2*n = %llu exceeds the uint32 element-offset (goff) contract "
                          "(max batch for n=%u is %llu)",
                          (unsigned long long)elems, n, 0xFFFFFFFFull / (2ull * n));
            c.reason = buf;
        } else if (c.l.in != Layout::Interleaved || c.l.out != Layout::Interleaved) {
            c.state = State::Infeasible;
            c.reason = "kernel consumes and produces interleaved complex "
                       "(planar layout would need an extra pass)";
        } else {
            c.state = State::Unverified;
        }
        if (c.state != State::Infeasible && !canonicalLowering(c, hw)) {
            c.state = State::Infeasible;
            c.reason = "no lowering for this candidate: the gfx950 kernels are radix-16 "
                       "Stockham passes with a padded-LDS exchange, one pass per exchange";
        }
        c.q = estimate(hw, c.a, c.p, n, batch);
        out.push_back(std::move(c));
    }
    return out;
}

std::vector<const Candidate*> rank(const std::vector<Candidate>& cs) {
    std::vector<const Candidate*> v;
    for (const auto& c : cs)
        if (c.state != State::Infeasible) v.push_back(&c);
    std::stable_sort(v.begin(), v.end(), [](const Candidate* x, const Candidate* y) {
        if (x->state != y->state) return x->state > y->state;
        return x->q.etaUs < y->q.etaUs;
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

bool Plan::validShape(int kind, uint32_t n, uint32_t batch, bool checkOffset) const {
    if (!impl_ || batch == 0 || !pow2(n)) return false;
    switch (kind) {
        case BFLY_KIND_C2C:
            if (n < kC2CMin || n > kC2CMax) return false;
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
    if (!canonicalLowering(cand_, impl_->hardware)) return false;
    const uint32_t local = kind == BFLY_KIND_C2C ? n : n / 2;
    if (!cand_.p.fits(local, impl_->hardware)) return false;
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
    tf_.clear();
    tf_.push_back({Transform::GenTwiddles, "gen FP64 twiddle table (Stockham pass layout)"});
    tf_.push_back({Transform::GenBitReverse, "no permutation tables (autosort passes)"});
    tf_.push_back({Transform::LaunchKernel, "bfly kernel on the context hipStream"});
    return 0;
}

int Plan::execute(int kind, int sign, const float* in, float* out, uint32_t n, uint32_t batch) {
    if (!in || !out || !validShape(kind, n, batch)) return -8;
    if (!impl_) return -1;
    if (prepareKind(kind, n, batch)) return -2;
    const size_t inF = kind == BFLY_KIND_C2C ? (size_t)2 * n * batch
                       : kind == BFLY_KIND_R2C ? (size_t)n * batch
                                               : (size_t)(n + 2) * batch;
    const size_t outF = kind == BFLY_KIND_C2C ? (size_t)2 * n * batch
                        : kind == BFLY_KIND_R2C ? (size_t)(n + 2) * batch
                                                : (size_t)n * batch;
    if (ensureBuf(&impl_->devIn, &impl_->inCap, inF * 4)) return -3;
    if (ensureBuf(&impl_->devOut, &impl_->outCap, outF * 4)) return -3;
    const size_t ws = bfly_workspace_bytes(kind, n, batch);
    if (ws && ensureBuf(&impl_->devWork, &impl_->workCap, ws)) return -3;
    if (bfly_rt_memcpy_h2d(impl_->devIn, in, inF * 4, ctxStream_) != BFLY_OK) return -4;
    const float* dIn = static_cast<const float*>(impl_->devIn);
    float* dOut = static_cast<float*>(impl_->devOut);
    const float* tab = static_cast<const float*>(impl_->devTable[kind]);
    int rc = BFLY_OK;
    if (kind == BFLY_KIND_C2C) rc = bfly_c2c(dIn, dOut, n, batch, sign, tab, impl_->devWork, ctxStream_);
    else if (kind == BFLY_KIND_R2C) rc = bfly_r2c(dIn, dOut, n, batch, tab, ctxStream_);
    else rc = bfly_c2r(dIn, dOut, n, batch, tab, ctxStream_);
    if (rc == BFLY_ERR_SHAPE) return -8;
    if (rc != BFLY_OK) return -5;
    if (bfly_rt_stream_sync(ctxStream_) != BFLY_OK) return -6;
    if (bfly_rt_memcpy_d2h(out, impl_->devOut, outF * 4, ctxStream_) != BFLY_OK) return -7;
    return 0;
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
    const size_t ws = bfly_workspace_bytes(BFLY_KIND_C2C, n, batch);
    if (ws && ensureBuf(&impl_->devWork, &impl_->workCap, ws)) return -3;
    const int rc = bfly_c2c(devIn, devOut, n, batch, sign,
                            static_cast<const float*>(impl_->devTable[BFLY_KIND_C2C]),
                            impl_->devWork, ctxStream_);
    return rc == BFLY_OK ? 0 : rc == BFLY_ERR_SHAPE ? -8 : -5;
}

int Plan::measure(uint32_t n, uint32_t batch, Metric* m) {
    if (!m) return -1;
    if (!validShape(BFLY_KIND_C2C, n, batch)) return -8;
    const size_t nf = (size_t)2 * n * batch;
    std::vector<float> in(nf), got(nf);
    for (size_t i = 0; i < nf; i++) in[i] = patternAt(i);
    if (run(in.data(), got.data(), n, batch)) return -2;
    // Time launches only, on the resident buffers (HIP events, best of 3 x 10).
    float best = 0.f;
    for (int round = 0; round < 3; round++) {
        float ms = 0.f;
        if (bfly_rt_time_transform(BFLY_KIND_C2C, -1, static_cast<const float*>(impl_->devIn),
                                   static_cast<float*>(impl_->devOut), n, batch,
                                   static_cast<const float*>(impl_->devTable[BFLY_KIND_C2C]),
                                   impl_->devWork, 10, ctxStream_, &ms) != BFLY_OK)
            return -5;
        if (round == 0 || ms < best) best = ms;
    }
    m->measuredUs = 1000.0 * best / 10.0;
    m->etaUs = cand_.q.etaUs;
    uint32_t rows[8];
    const uint32_t nrows = sampleBatches(batch, 8, rows, 8);
    std::vector<float> ref(2 * n);
    double worst = 0.0;
    for (uint32_t i = 0; i < nrows; i++) {
        refFftF32(in.data() + (size_t)rows[i] * 2 * n, ref.data(), n);
        worst = std::max(worst, maxRelScaled(got.data() + (size_t)rows[i] * 2 * n, ref.data(), 2 * n));
    }
    m->maxRelErr = worst;
    if (!std::isfinite(worst) || worst > 1e-4) {
        m->measuredOk = false;
        cand_.q = *m;
        cand_.state = State::Unverified;
        cand_.reason = "correctness check failed (maxRel > 1e-4)";
        return -7;
    }
    m->measuredOk = true;
    cand_.state = State::Feasible;
    cand_.q = *m;
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
    Hardware hw;
    DesignSpace sp;
    try {
        hw = Hardware::load(profilePath);
        sp = DesignSpace::load(spacePath);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Context::init configuration error: %s\n", e.what());
        return -4;
    }
    hw_ = std::move(hw);
    sp_ = std::move(sp);
    kernelPath_ = kernelPath;
    const int rc = bfly_rt_init(0);
    if (rc == BFLY_ERR_NOT_LOADED) return -1;
    if (rc != BFLY_OK) return -2;
    if (bfly_rt_stream_create(&stream_) != BFLY_OK) return -3;
    initialized_ = true;
    return 0;
}

std::vector<Candidate> Context::enumerate(uint32_t n, uint32_t batch) const {
    return bfly::enumerate(sp_, hw_, n, batch);
}

std::unique_ptr<Plan> Context::makePlan(const Candidate& c) {
    if (!initialized_ || c.state == State::Infeasible || !canonicalLowering(c, hw_)) return nullptr;
    auto p = std::unique_ptr<Plan>(new Plan(c));
    p->impl_.reset(new Plan::Impl());
    p->impl_->hardware = hw_;
    p->ctxStream_ = stream_;
    p->tf_.push_back({Transform::GenTwiddles, "gen FP64 twiddle table (Stockham pass layout)"});
    p->tf_.push_back({Transform::GenBitReverse, "no permutation tables (autosort passes)"});
    p->tf_.push_back({Transform::LaunchKernel, "bfly kernel on the context hipStream"});
    return p;
}

std::unique_ptr<Plan> Context::select(uint32_t n, uint32_t batch, int topK, std::string* log) {
    auto all = enumerate(n, batch);
    auto order = rank(all);
    const int k = std::max(1, topK);
    char line[512];
    if (log) {
        log->clear();
        int infeasible = 0;
        for (const auto& c : all) infeasible += c.state == State::Infeasible;
        std::snprintf(line, sizeof(line), "enumerated=%zu infeasible=%d feasible=%zu (n=%u batch=%u)\n",
                      all.size(), infeasible, order.size(), n, batch);
        *log += line;
        std::vector<std::pair<int, std::string>> reasons;
        for (const auto& c : all) {
            if (c.state != State::Infeasible) continue;
            auto it = std::find_if(reasons.begin(), reasons.end(),
                                   [&](const auto& r) { return r.second == c.reason; });
            if (it == reasons.end()) reasons.push_back({1, c.reason});
            else it->first++;
        }
        std::stable_sort(reasons.begin(), reasons.end(),
                         [](const auto& x, const auto& y) { return x.first > y.first; });
        for (size_t i = 0; i < reasons.size() && i < 5; i++) {
            std::snprintf(line, sizeof(line), "  infeasible x%-4d %s\n", reasons[i].first,
                          reasons[i].second.c_str());
            *log += line;
        }
        std::snprintf(line, sizeof(line), "  eta rank (top %d):\n", k);
        *log += line;
        for (size_t i = 0; i < order.size() && i < (size_t)k; i++) {
            std::snprintf(line, sizeof(line), "    %2zu. %-44s eta=%7.1f us  [%s]\n", i + 1,
                          order[i]->id.c_str(), order[i]->q.etaUs, toString(order[i]->state));
            *log += line;
        }
        *log += "  measured selection:\n";
    }
    std::unique_ptr<Plan> best;
    double bestUs = 0.0;
    int accepted = 0;
    size_t examined = 0;
    for (size_t i = 0; i < order.size() && accepted < k; i++) {
        examined++;
        const Candidate& c = *order[i];
        auto p = makePlan(c);
        if (!p) {
            if (log) {
                std::snprintf(line, sizeof(line), "    fail   %-44s makePlan failed\n", c.id.c_str());
                *log += line;
            }
            continue;
        }
        Metric m;
        const int rc = p->measure(n, batch, &m);
        if (rc) {
            if (log) {
                std::snprintf(line, sizeof(line), "    reject %-44s rc=%d maxRel=%.2e [%s]\n",
                              c.id.c_str(), rc, m.maxRelErr, toString(p->candidate().state));
                *log += line;
            }
            continue;
        }
        accepted++;
        if (log) {
            std::snprintf(line, sizeof(line),
                          "    accept %-44s eta=%7.1f us  measured=%7.1f us  maxRel=%.1e  [%s]\n",
                          p->candidate().id.c_str(), m.etaUs, m.measuredUs, m.maxRelErr,
                          toString(p->candidate().state));
            *log += line;
        }
        if (!best || m.measuredUs < bestUs) {
            bestUs = m.measuredUs;
            best = std::move(p);
        }
    }
    if (log) {
        if (best)
            std::snprintf(line, sizeof(line), "  selected: %s  (%.1f us, %d accepted after %zu examined, quota %d)\n",
                          best->candidate().id.c_str(), bestUs, accepted, examined, k);
        else
            std::snprintf(line, sizeof(line), "  selected: none (%d accepted after %zu examined, quota %d)\n",
                          accepted, examined, k);
        *log += line;
    }
    return best;
}

}  // namespace bfly
