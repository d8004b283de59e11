// Host-only checks of the C++ Context / Plan API on the CPU stub of the C ABI
// (libbfly_stub.so, which runs every argument guard, counts device traffic
// and computes nothing).  The contract under test is the reference's
// (src/framework/butterfly.cpp:515-1102, docs/reference/api.md:38-55):
// configuration errors give -4 and never touch the device, re-init gives -9,
// shape / pointer / offset guards give -8 with zero device traffic, a failed
// table rebuild never leaves the old key valid, and launch / copy failures
// map to -5 / -4 / -7.  On top of it, the gfx950 selection: which kernel every
// candidate lowers to at every length, and that a plan runs exactly that kernel.
//
// Usage: plan_host_checks <config dir> <scratch dir>; prints one line per
// failed check and a summary, exit status 0 when every check held.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <functional>
#include <limits>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "bfly_hip.h"
#include "butterfly/plan.hpp"
#include "butterfly/reference.hpp"

extern "C" {
void bfly_stub_reset(void);
void bfly_stub_set_failures(int allocation, int copy, int launch);
void bfly_stub_counters(int* out);
int bfly_stub_last_c2c_variant(void);
}

namespace {

int g_checks = 0, g_failed = 0;
std::string g_section;

void expect(bool ok, const std::string& what) {
    g_checks++;
    if (!ok) {
        g_failed++;
        std::printf("FAILED [%s] %s\n", g_section.c_str(), what.c_str());
    }
}

struct Traffic {
    int allocations, frees, copies, launches, streams, streamsGone, inits, syncs;
    static Traffic now() {
        int v[8];
        bfly_stub_counters(v);
        return {v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]};
    }
    bool none_since(const Traffic& t) const {
        const Traffic n = now();
        return n.allocations == t.allocations && n.copies == t.copies && n.launches == t.launches;
    }
};

std::string slurp(const std::string& path) {
    std::ifstream f(path);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// `text` with its first `from` replaced; empty when `from` is absent (the
// caller then reports the fixture as stale).
std::string edit(std::string text, const std::string& from, const std::string& to) {
    const size_t at = text.find(from);
    if (at == std::string::npos) return {};
    return text.replace(at, from.size(), to);
}

void put(const std::string& path, const std::string& text) { std::ofstream(path) << text; }

bool loadFails(const std::function<void()>& load) {
    try {
        load();
    } catch (const std::exception&) {
        return true;
    }
    return false;
}

// A candidate by its axes (everything else at the lowered defaults).
bfly::Candidate cand(int ts, int tb, int ub, bool ldsTw, const char* exchange, bool mfma, int radix = 16,
                     int ud = 256) {
    bfly::Candidate c;
    c.a.ts = ts;
    c.a.tb = tb;
    c.a.ub = ub;
    c.a.udCore = ud;
    c.a.localExchange = exchange;
    c.p.coefficientResidency = ldsTw;
    c.p.radix = radix;
    c.f.dftMatrix = mfma;
    c.state = bfly::State::Unverified;
    return c;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::printf("usage: %s <config dir> <scratch dir>\n", argv[0]);
        return 2;
    }
    const std::string cfg = argv[1], scratch = argv[2];
    const std::string hwPath = cfg + "/gfx950_profile.json", spPath = cfg + "/butterfly_space.json";
    const std::string hwText = slurp(hwPath), spText = slurp(spPath);
    bfly_stub_reset();

    g_section = "configuration";
    {
        struct Bad {
            const std::string* base;
            const char* from;
            const char* to;
        };
        const Bad bad[] = {
            {&hwText, "\"compute_units\": 256", "\"compute_units\": 0"},
            {&hwText, "\"lds_bytes_per_cu\": 163840", "\"lds_bytes_per_cu\": \"163840\""},
            {&hwText, "\"wave_size\": 64", "\"wave_size\": 64.5"},
            {&hwText, "\"schema_version\": 1", "\"schema_version\": 2"},
            {&hwText, "\"soc\": \"gfx950\",", "\"soc\": \"gfx950\", \"soc\": \"again\","},
            {&hwText, "\"hbm_peak_gbs\": 8000", "\"hbm_peak_gbs\": 01"},
            {&spText, "\"ud_core\": [64, 128, 256]", "\"ud_core\": []"},
            {&spText, "\"radix\": [16, 8, 4, 2]", "\"radix\": [\"16\"]"},
            {&spText, "\"radix\": [16, 8, 4, 2]", "\"radix\": [12]"},
            {&spText, "\"ts\": [1, 0]", "\"ts\": [2]"},
            {&spText, "\"ub\": [1, 2]", "\"ub\": [3]"},
            {&spText, "\"out\": [\"interleaved\", \"planar\"]", "\"out\": [\"bogus\"]"},
        };
        for (const Bad& b : bad) {
            const std::string text = edit(*b.base, b.from, b.to);
            expect(!text.empty(), std::string("fixture text present: ") + b.from);
            const bool isHw = b.base == &hwText;
            const std::string path = scratch + (isHw ? "/hw.json" : "/space.json");
            put(path, text);
            expect(loadFails([&] { isHw ? (void)bfly::Hardware::load(path) : (void)bfly::DesignSpace::load(path); }),
                   std::string("malformed configuration rejected: ") + b.to);
            bfly::Context ctx;
            expect(ctx.init(isHw ? path : hwPath, isHw ? spPath : path, "libbfly_hip.so") == -4,
                   std::string("init -4 on ") + b.to);
        }
        for (const std::string& junk : {std::string("{"), std::string("{}"), hwText + " x"}) {
            put(scratch + "/hw.json", junk);
            expect(loadFails([&] { (void)bfly::Hardware::load(scratch + "/hw.json"); }), "junk profile rejected");
        }
        expect(loadFails([] { (void)bfly::Hardware::load("/no/such/profile.json"); }), "missing profile");
        expect(loadFails([] { (void)bfly::DesignSpace::load("/no/such/space.json"); }), "missing space");
        expect(Traffic::now().streams == 0 && Traffic::now().inits == 0, "no device call on a bad configuration");
    }

    g_section = "maxRelScaled";
    {
        // values held by the reference's tests/test_runtime_guards.cpp:63-85
        const float inf = std::numeric_limits<float>::infinity(), nan = std::numeric_limits<float>::quiet_NaN();
        struct Row {
            float got, ref;
            double lo, hi;  // expected value within [lo, hi]
        };
        const Row rows[] = {{1.f, 0.f, 1.0, 1.0},     {1e-6f, 0.f, 1e-12, 1e-5}, {1e-6f, 1e-8f, 0.0, 1e-4},
                            {101.f, 100.f, .01, .01}, {nan, 1e-8f, inf, inf},   {inf, 0.f, inf, inf},
                            {-inf, 0.f, inf, inf},    {0.f, nan, inf, inf},     {0.f, -inf, inf, inf}};
        for (const Row& r : rows) {
            const float got[2] = {r.got, 0.f}, ref[2] = {r.ref, 0.f};
            const double v = bfly::maxRelScaled(got, ref, 2);
            expect(v >= r.lo && v <= r.hi, "maxRelScaled(" + std::to_string(r.got) + ", " + std::to_string(r.ref) +
                                               ") = " + std::to_string(v));
        }
        const float zero[2] = {0.f, 0.f};
        expect(std::isinf(bfly::maxRelScaled(nullptr, zero, 2)), "null output is an infinite error");
        expect(bfly::maxRelScaled(nullptr, nullptr, 0) == 0.0, "empty comparison is exact");
    }

    bfly::Context ctx;
    g_section = "init";
    expect(ctx.init(hwPath, spPath, "libbfly_hip.so") == 0, "init");
    expect(Traffic::now().streams == 1, "one stream");
    expect(ctx.init(hwPath, spPath, "libbfly_hip.so") == -9, "re-init -9");
    expect(Traffic::now().streams == 1, "re-init creates nothing");
    const bfly::Hardware& hw = ctx.hardware();
    expect(hw.aicoreNum == 256 && hw.vecLaneFp32 == 64 && hw.ubBytesPerCore == 163840, "gfx950 profile");

    g_section = "lowering";
    {
        struct Row {
            bfly::Candidate c;
            uint32_t n;
            int kernel;         // expected BFLY_VARIANT_*, 0 = no lowering
            const char* words;  // in the reason when kernel == 0
        };
        const Row rows[] = {
            {cand(1, 0, 1, true, "shared", false), 4096, BFLY_VARIANT_STREAM_LT, ""},
            {cand(1, 0, 1, true, "shared", false), 64, BFLY_VARIANT_STREAM_LT, ""},
            {cand(1, 0, 1, true, "shared", false), 8192, BFLY_VARIANT_STREAM_LT, ""},
            {cand(1, 0, 1, true, "shared", false), 16384, BFLY_VARIANT_STREAM_LT, ""},
            {cand(1, 0, 1, true, "shared", false), 32768, 0, "four-step chain (ts=0)"},
            {cand(1, 1, 1, false, "shared", false), 1024, BFLY_VARIANT_PLAIN, ""},
            {cand(1, 1, 1, false, "shared", false), 16384, BFLY_VARIANT_PLAIN, ""},
            {cand(1, 1, 1, false, "swizzle", false), 256, BFLY_VARIANT_PLAIN_SWZ, ""},
            {cand(1, 0, 1, false, "shared", false), 512, BFLY_VARIANT_STREAM, ""},
            {cand(1, 0, 1, false, "swizzle", false), 512, BFLY_VARIANT_STREAM_SWZ, ""},
            {cand(1, 0, 1, false, "swizzle", false), 8192, 0, "padded exchange"},
            {cand(1, 0, 2, false, "shared", false), 2048, BFLY_VARIANT_STREAM2, ""},
            {cand(1, 0, 2, true, "shared", false), 2048, 0, "two tiles in flight"},
            {cand(1, 1, 1, true, "shared", false), 2048, 0, "LDS-resident twiddles"},
            {cand(1, 1, 1, true, "shared", true), 4096, BFLY_VARIANT_SPLIT_MFMA, ""},
            {cand(1, 0, 1, true, "shared", true), 1024, BFLY_VARIANT_SPLIT_MFMA_STREAM, ""},
            {cand(1, 1, 1, true, "shared", true), 512, 0, "1024 / 2048 / 4096"},
            {cand(1, 1, 1, false, "shared", true), 4096, 0, "twiddle tables in LDS"},
            {cand(0, 1, 1, true, "shared", false), 65536, BFLY_VARIANT_FOUR_STEP, ""},
            {cand(0, 1, 1, true, "shared", false), 8192, BFLY_VARIANT_FOUR_STEP, ""},
            {cand(0, 1, 1, true, "shared", false), 4096, 0, "starts at 8192"},
            {cand(0, 0, 1, true, "shared", false), 65536, 0, "four-step chain is compiled"},
            {cand(1, 1, 1, false, "shared", false, 8), 1024, 0, "radix 8"},
            {cand(1, 1, 1, false, "shared", false, 16, 128), 1024, 0, "ud_core 128"},
            {cand(1, 1, 1, false, "shuffle", false), 1024, 0, "local_exchange=shuffle"},
            {cand(1, 1, 1, false, "shared", false), 32, 0, "rows"},
            {cand(1, 1, 1, false, "shared", false), 16, 0, "rows"},
            {cand(1, 1, 1, false, "shared", false), 8, 0, "rows"},
            {cand(1, 1, 1, false, "shared", false), 100, 0, "power of two"},
            {cand(1, 1, 1, false, "shared", false), 192, 0, "power of two"},
            {cand(0, 1, 1, true, "shared", false), 131072, 0, "long-FFT envelope"},
        };
        for (const Row& r : rows) {
            std::string why;
            const int k = bfly::lowerC2C(r.c, hw, r.n, &why);
            const std::string tag = r.c.a.localExchange + " ts" + std::to_string(r.c.a.ts) + " tb" +
                                    std::to_string(r.c.a.tb) + " ub" + std::to_string(r.c.a.ub) + " n=" +
                                    std::to_string(r.n);
            expect(k == r.kernel, tag + ": kernel " + bfly::kernelName(k) + ", want " + bfly::kernelName(r.kernel) +
                                      (k ? "" : " (" + why + ")"));
            if (!r.kernel) expect(why.find(r.words) != std::string::npos, tag + ": reason '" + why + "' lacks '" +
                                                                                r.words + "'");
        }
        // mutations the kernels do not have
        const std::vector<std::function<void(bfly::Candidate&)>> foreign = {
            [](auto& c) { c.p.bitReverseInput = true; }, [](auto& c) { c.p.fusionLevel = 2; },
            [](auto& c) { c.l.contiguous = false; },     [](auto& c) { c.l.in = bfly::Layout::Planar; },
            [](auto& c) { c.l.out = bfly::Layout::Planar; }, [](auto& c) { c.a.td = 2; },
            [](auto& c) { c.a.us = 2; },                 [](auto& c) { c.a.level = 1; },
            [](auto& c) { c.a.udCore = 0; },             [](auto& c) { c.a.udCore = 512; },
            [](auto& c) { c.a.tb = 2; },                 [](auto& c) { c.a.ub = 3; },
            [](auto& c) { c.a.localExchange = "register"; }, [](auto& c) { c.a.localExchange = "bogus"; },
            [](auto& c) { c.f.butterflyCount = 2; },     [](auto& c) { c.f.stageCount = 2; },
        };
        const Traffic t = Traffic::now();
        for (size_t i = 0; i < foreign.size(); i++) {
            auto c = cand(1, 0, 1, true, "shared", false);
            foreign[i](c);
            for (uint32_t n = 64; n <= 65536; n *= 2)
                expect(!bfly::lowerC2C(c, hw, n, nullptr), "mutation " + std::to_string(i) + " lowers at " +
                                                               std::to_string(n));
            expect(!ctx.makePlan(c), "makePlan refuses mutation " + std::to_string(i));
        }
        expect(t.none_since(t), "refused plans allocate nothing");
    }

    g_section = "cost model";
    {
        // eta of the fitted model as the C++ side reads it, for the Python
        // test to compare with the calibration's own prediction
        struct Point {
            int kernel;
            uint32_t n, batch;
        };
        const Point pts[] = {{BFLY_VARIANT_STREAM_LT, 4096, 4096}, {BFLY_VARIANT_PLAIN, 256, 64},
                             {BFLY_VARIANT_SPLIT_MFMA_STREAM, 2048, 8192}, {BFLY_VARIANT_STREAM_LT, 16384, 1024},
                             {BFLY_VARIANT_PLAIN, 8192, 64}, {BFLY_VARIANT_FOUR_STEP, 65536, 256},
                             {BFLY_VARIANT_FOUR_STEP, 8192, 4}, {BFLY_VARIANT_STREAM2, 1024, 3}};
        for (const Point& p : pts) {
            const double eta = bfly::kernelEta(ctx.space(), hw, p.kernel, p.n, p.batch);
            expect(eta > 0.0, "positive eta");
            std::printf("eta %s %u %u %.6f\n", bfly::kernelName(p.kernel), p.n, p.batch, eta);
        }
        expect(bfly::modelKey(BFLY_VARIANT_PLAIN, 16384) == "plain_single" &&
                   bfly::modelKey(BFLY_VARIANT_STREAM_LT, 4096) == "stream_lt" &&
                   bfly::modelKey(BFLY_VARIANT_FOUR_STEP, 8192) == "four_step",
               "model keys");
    }

    g_section = "enumerate";
    {
        auto kernelsAt = [&](uint32_t n, uint32_t b) {
            std::set<std::string> ks;
            std::set<std::string> ids;
            for (const auto& c : ctx.enumerate(n, b)) {
                expect(ids.insert(c.id).second, "unique id " + c.id);
                if (c.state == bfly::State::Infeasible) {
                    expect(!c.reason.empty() && c.kernel == 0, "infeasible " + c.id + " carries a reason");
                } else {
                    expect(c.kernel == bfly::lowerC2C(c, hw, n, nullptr) && c.q.etaUs > 0.0,
                           "lowered " + c.id + " has its kernel and an eta");
                    ks.insert(bfly::kernelName(c.kernel));
                }
            }
            return ks;
        };
        expect(kernelsAt(4096, 4096) == std::set<std::string>{"plain", "plain_swz", "split_mfma",
                                                              "split_mfma_stream", "stream", "stream2",
                                                              "stream_lt", "stream_swz"},
               "eight kernels at 4096");
        expect(kernelsAt(256, 64) == std::set<std::string>{"plain", "plain_swz", "stream", "stream2", "stream_lt",
                                                           "stream_swz"},
               "no MFMA kernel at 256");
        expect(kernelsAt(8192, 64) == std::set<std::string>{"four_step", "plain", "stream_lt"}, "8192");
        expect(kernelsAt(65536, 4) == std::set<std::string>{"four_step"}, "65536");
        expect(kernelsAt(32, 4).empty() && kernelsAt(64, 0).empty(), "nothing lowers below 64 or at batch 0");
        expect(kernelsAt(4096, 524288).empty(), "batch*2n beyond 2^32-1");
        // the ranking puts lower eta first
        const auto all = ctx.enumerate(4096, 4096);
        const auto order = bfly::rank(all);
        for (size_t i = 1; i < order.size(); i++)
            expect(order[i - 1]->q.etaUs <= order[i]->q.etaUs, "rank by eta");
    }

    g_section = "plans";
    auto byKernel = [&](uint32_t n, uint32_t b, int k) -> std::unique_ptr<bfly::Plan> {
        for (const auto& c : ctx.enumerate(n, b))
            if (c.kernel == k) return ctx.makePlan(c);
        return nullptr;
    };
    {
        bfly::Context idle;
        auto c = cand(1, 0, 1, true, "shared", false);
        expect(!idle.makePlan(c), "an uninitialised context makes no plan");
        c.state = bfly::State::Infeasible;
        expect(!ctx.makePlan(c), "an Infeasible candidate makes no plan");
        // every lowered kernel runs as itself, at every length it lowers to
        std::vector<float> in(2 * 8192 * 3, 0.25f), out(in.size());
        const int kernels[] = {BFLY_VARIANT_PLAIN,     BFLY_VARIANT_STREAM,     BFLY_VARIANT_PLAIN_SWZ,
                               BFLY_VARIANT_STREAM_SWZ, BFLY_VARIANT_STREAM_LT, BFLY_VARIANT_STREAM2,
                               BFLY_VARIANT_SPLIT_MFMA, BFLY_VARIANT_SPLIT_MFMA_STREAM};
        for (int k : kernels) {
            auto plan = byKernel(4096, 3, k);
            expect(plan != nullptr, std::string("a plan for ") + bfly::kernelName(k));
            if (!plan) continue;
            expect(plan->run(in.data(), out.data(), 4096, 3) == 0 && bfly_stub_last_c2c_variant() == k,
                   std::string("runs ") + bfly::kernelName(k));
        }
        auto lt = byKernel(4096, 3, BFLY_VARIANT_STREAM_LT);
        expect(lt->run(in.data(), out.data(), 8192, 3) == 0 && bfly_stub_last_c2c_variant() == BFLY_VARIANT_STREAM_LT,
               "the LDS-twiddle candidate runs its 8192 form");
        const Traffic t = Traffic::now();
        expect(lt->run(in.data(), out.data(), 32768, 1) == -8 && t.none_since(t),
               "the LDS-twiddle candidate has no 32768 form: -8, no traffic");
        auto chain = byKernel(65536, 1, BFLY_VARIANT_FOUR_STEP);
        expect(chain && chain->run(in.data(), out.data(), 8192, 3) == 0 &&
                   bfly_stub_last_c2c_variant() == BFLY_VARIANT_FOUR_STEP,
               "the four-step candidate runs 8192 as the chain");
        expect(chain && chain->run(in.data(), out.data(), 4096, 3) == -8, "and has no 4096 form");
        auto mfma = byKernel(4096, 3, BFLY_VARIANT_SPLIT_MFMA);
        expect(mfma && mfma->run(in.data(), out.data(), 256, 3) == -8, "split MFMA has no 256 form");
        // the real transforms run for every plan
        expect(mfma && mfma->runR2C(in.data(), out.data(), 256, 3) == 0 && mfma->runC2R(in.data(), out.data(), 256, 2) == 0,
               "real transforms on an MFMA plan");
    }

    g_section = "guards";
    auto plan = byKernel(64, 1, BFLY_VARIANT_STREAM_LT);
    expect(plan != nullptr, "plan at 64");
    if (!plan) {
        std::printf("plan host checks: %d of %d failed\n", g_failed, g_checks);
        return 1;
    }
    {
        const Traffic t0 = Traffic::now();
        float x = 0.f;
        for (uint32_t n : {0u, 1u, 32u, 192u, 131072u}) {
            expect(plan->run(&x, &x, n, 1) == -8 && plan->runInverse(&x, &x, n, 1) == -8, "C2C n " + std::to_string(n));
            expect(plan->runR2C(&x, &x, n, 1) == -8 && plan->runC2R(&x, &x, n, 1) == -8, "real n " + std::to_string(n));
            expect(plan->prepare(n, 1) == -8, "prepare n " + std::to_string(n));
        }
        expect(plan->runC2R(&x, &x, 8192, 1) == -8, "C2R envelope ends at 4096");
        expect(plan->runR2C(&x, &x, 64, 1) == -8 && plan->runR2C(&x, &x, 16384, 1) == -8, "R2C envelope 128..8192");
        using Run = int (bfly::Plan::*)(const float*, float*, uint32_t, uint32_t);
        for (Run r : {&bfly::Plan::run, &bfly::Plan::runInverse, &bfly::Plan::runR2C, &bfly::Plan::runC2R}) {
            expect((plan.get()->*r)(&x, &x, 128, 0) == -8, "batch 0");
            expect((plan.get()->*r)(nullptr, &x, 128, 1) == -8 && (plan.get()->*r)(&x, nullptr, 128, 1) == -8,
                   "null pointers");
            expect((plan.get()->*r)(&x, &x, 128, 0xffffffffu) == -8, "offset guard");
        }
        expect(plan->run(&x, &x, 4096, 524288) == -8, "batch*2n = 2^32");
        expect(plan->runDevice(&x, &x, 4096, 1, 0) == -8, "sign 0");
        bfly::Metric m;
        expect(plan->measure(1, 1, &m) == -8 && plan->prepare(64, 0) == -8, "measure / prepare guards");
        expect(t0.none_since(t0), "guards touch no device");
    }

    g_section = "invalidation";
    {
        auto fail_next = [](int which) {  // 0 allocation, 1 copy
            const Traffic t = Traffic::now();
            bfly_stub_set_failures(which == 0 ? t.allocations + 1 : 0, which == 1 ? t.copies + 1 : 0, 0);
        };
        expect(plan->prepare(64, 1) == 0, "prepare 64");
        for (int which : {0, 1}) {
            fail_next(which);
            expect(plan->prepare(128, 1) != 0, "failed rebuild reported");
            bfly_stub_set_failures(0, 0, 0);
            const int before = Traffic::now().allocations;
            expect(plan->prepare(64, 1) == 0 && Traffic::now().allocations > before,
                   "the old key was invalidated, so 64 is rebuilt");
        }
        const int before = Traffic::now().allocations;
        expect(plan->prepare(64, 1) == 0 && Traffic::now().allocations == before, "a valid key is reused");
    }

    g_section = "execute errors";
    {
        std::vector<float> in(2 * 256 * 3, 0.25f), out(in.size());
        const Traffic t0 = Traffic::now();
        expect(plan->run(in.data(), out.data(), 256, 3) == 0, "run");
        const Traffic t1 = Traffic::now();
        expect(t1.launches == t0.launches + 1 && t1.copies >= t0.copies + 2, "one launch, copies in and out");
        // the Nth call from now fails: (launch 1) -> -5, (copy 1, the upload) -> -4,
        // (copy 2, the download) -> -7
        struct Inject {
            int launch, copy, rc;
        };
        for (const Inject& f : {Inject{1, 0, -5}, Inject{0, 1, -4}, Inject{0, 2, -7}}) {
            const Traffic t = Traffic::now();
            bfly_stub_set_failures(0, f.copy ? t.copies + f.copy : 0, f.launch ? t.launches + f.launch : 0);
            expect(plan->run(in.data(), out.data(), 256, 3) == f.rc, "injected failure -> " + std::to_string(f.rc));
            bfly_stub_set_failures(0, 0, 0);
        }
        expect(plan->runR2C(in.data(), out.data(), 256, 3) == 0 && plan->runC2R(in.data(), out.data(), 256, 2) == 0,
               "real transforms");
    }

    std::printf("plan host checks: %d of %d failed\n", g_failed, g_checks);
    return g_failed ? 1 : 0;
}
