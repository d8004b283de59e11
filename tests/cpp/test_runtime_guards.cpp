// Host-only guard tests of the C++ API, linked against the CPU stub of the C
// ABI (libbfly_stub.so): configuration rejection, re-init, non-canonical
// candidates, shape/pointer/offset guards with zero device traffic, and
// cache invalidation after a failed prepare.  Mirrors the reference's
// tests/test_runtime_guards.cpp:36-179 on the gfx950 configuration.
#include <cassert>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>

#include "bfly_hip.h"
#include "butterfly/plan.hpp"
#include "butterfly/reference.hpp"

extern "C" {
void bfly_stub_reset(void);
void bfly_stub_set_failures(int allocation, int copy, int launch);
void bfly_stub_counters(int* out);
}

namespace {
struct Counts {
    int allocations, frees, copies, launches, streamsCreated, streamsDestroyed, inits, syncs;
};
Counts counts() {
    int v[8];
    bfly_stub_counters(v);
    return {v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]};
}
std::string read(const std::string& path) {
    std::ifstream f(path);
    return {std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>()};
}
std::string replaced(std::string text, const std::string& from, const std::string& to) {
    const size_t at = text.find(from);
    if (at == std::string::npos) throw std::runtime_error("fixture text not found: " + from);
    text.replace(at, from.size(), to);
    return text;
}
void write(const std::string& path, const std::string& text) {
    std::ofstream f(path);
    f << text;
}
bool throws(const std::function<void()>& op) {
    try {
        op();
    } catch (const std::exception&) {
        return true;
    }
    return false;
}
int checks = 0;
#define CHECK(x)                                                                   \
    do {                                                                           \
        if (!(x)) {                                                                \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #x);  \
            return 1;                                                              \
        }                                                                          \
        checks++;                                                                  \
    } while (0)
}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    const std::string configDir = argv[1], tmp = argv[2];
    const std::string hwPath = configDir + "/gfx950_profile.json";
    const std::string spPath = configDir + "/butterfly_space.json";
    const auto hwText = read(hwPath), spText = read(spPath);
    const std::string badHw = tmp + "/hardware.json", badSp = tmp + "/space.json";
    bfly_stub_reset();

    CHECK(throws([] { bfly::Hardware::load("/nonexistent-hardware-config"); }));
    CHECK(throws([] { bfly::DesignSpace::load("/nonexistent-space-config"); }));
    for (const auto& bad : {std::string("{}"), hwText + " trailing",
                            replaced(hwText, "\"compute_units\": 256", "\"compute_units\": 0"),
                            replaced(hwText, "\"lds_bytes_per_cu\": 163840", "\"lds_bytes_per_cu\": \"163840\""),
                            replaced(hwText, "\"wave_size\": 64", "\"wave_size\": 64.5"),
                            replaced(hwText, "\"schema_version\": 1", "\"schema_version\": 2"),
                            replaced(hwText, "\"soc\": \"gfx950\",", "\"soc\": \"gfx950\", \"soc\": \"x\","),
                            replaced(hwText, "\"hbm_peak_gbs\": 8000", "\"hbm_peak_gbs\": 01")}) {
        write(badHw, bad);
        CHECK(throws([&] { bfly::Hardware::load(badHw); }));
        bfly::Context invalid;
        CHECK(invalid.init(badHw, spPath, "libbfly_hip.so") == -4);
    }
    for (const auto& bad : {std::string("{"), std::string("{}"), spText + " trailing",
                            replaced(spText, "\"ud_core\": [64, 128, 256]", "\"ud_core\": []"),
                            replaced(spText, "\"radix\": [16, 8, 4, 2]", "\"radix\": [\"16\"]"),
                            replaced(spText, "\"radix\": [16, 8, 4, 2]", "\"radix\": [12]"),
                            replaced(spText, "\"ts\": [1, 0]", "\"ts\": [2]"),
                            replaced(spText, "\"out\": [\"interleaved\", \"planar\"]", "\"out\": [\"bogus\"]")}) {
        write(badSp, bad);
        CHECK(throws([&] { bfly::DesignSpace::load(badSp); }));
        bfly::Context invalid;
        CHECK(invalid.init(hwPath, badSp, "libbfly_hip.so") == -4);
    }
    CHECK(counts().streamsCreated == 0);

    // maxRelScaled semantics (reference tests/test_runtime_guards.cpp:63-85)
    float reference[] = {0.f, 0.f}, output[] = {1.f, 0.f};
    CHECK(bfly::maxRelScaled(output, reference, 2) == 1.0);
    output[0] = 1e-6f;
    CHECK(bfly::maxRelScaled(output, reference, 2) > 0);
    reference[0] = 1e-8f;
    CHECK(bfly::maxRelScaled(output, reference, 2) < 1e-4);
    for (float v : {std::numeric_limits<float>::quiet_NaN(), std::numeric_limits<float>::infinity(),
                    -std::numeric_limits<float>::infinity()}) {
        output[0] = v;
        CHECK(std::isinf(bfly::maxRelScaled(output, reference, 2)));
        output[0] = 0;
        reference[0] = v;
        CHECK(std::isinf(bfly::maxRelScaled(output, reference, 2)));
        reference[0] = 0;
    }
    CHECK(std::isinf(bfly::maxRelScaled(nullptr, reference, 2)));
    CHECK(bfly::maxRelScaled(nullptr, nullptr, 0) == 0);
    reference[0] = 100.f;
    output[0] = 101.f;
    CHECK(bfly::maxRelScaled(output, reference, 2) == .01);

    bfly::Context context;
    CHECK(context.init(hwPath, spPath, "libbfly_hip.so") == 0);
    CHECK(counts().streamsCreated == 1);
    CHECK(context.init(hwPath, spPath, "libbfly_hip.so") == -9);
    CHECK(counts().streamsCreated == 1);
    CHECK(context.hardware().aicoreNum == 256 && context.hardware().vecLaneFp32 == 64);

    auto candidates = context.enumerate(64, 1);
    std::unique_ptr<bfly::Plan> plan;
    for (const auto& c : candidates)
        if (c.state != bfly::State::Infeasible) {
            plan = context.makePlan(c);
            break;
        }
    CHECK(plan != nullptr);
    const auto canonical = plan->candidate();
    const std::vector<std::function<void(bfly::Candidate&)>> mutations = {
        [](auto& c) { c.p.bitReverseInput = true; },
        [](auto& c) { c.p.fusionLevel = 2; },
        [](auto& c) { c.p.coefficientResidency = false; },
        [](auto& c) { c.p.radix = 4; },
        [](auto& c) { c.p.radix = 2; },
        [](auto& c) { c.l.contiguous = false; },
        [](auto& c) { c.l.in = bfly::Layout::Planar; },
        [](auto& c) { c.l.out = bfly::Layout::Planar; },
        [](auto& c) { c.a.td = 2; },
        [](auto& c) { c.a.tb = 2; },
        [](auto& c) { c.a.ub = 2; },
        [](auto& c) { c.a.us = 2; },
        [](auto& c) { c.a.ts = 0; },
        [](auto& c) { c.a.level = 1; },
        [](auto& c) { c.a.udCore = 0; },
        [](auto& c) { c.a.udCore = 512; },
        [](auto& c) { c.a.localExchange = "register"; },
        [](auto& c) { c.a.localExchange = "bogus"; },
        [](auto& c) { c.f.butterflyCount = 2; },
        [](auto& c) { c.f.stageCount = 2; },
        [](auto& c) { c.f.dftMatrix = true; },
        [](auto& c) { c.state = bfly::State::Infeasible; },
    };
    const int allocsBefore = counts().allocations;
    for (const auto& mutate : mutations) {
        auto c = canonical;
        mutate(c);
        CHECK(!context.makePlan(c));
    }
    CHECK(counts().allocations == allocsBefore);
    bfly::Context uninitialized;
    CHECK(!uninitialized.makePlan(canonical));
    for (const auto& c : candidates) {
        if (c.state != bfly::State::Infeasible) CHECK(context.makePlan(c) != nullptr);
        if (c.a.localExchange != "shared") CHECK(c.state == bfly::State::Infeasible);
    }

    // Shape guards: -8 with no allocation, copy or launch.
    float dummy = 0;
    for (uint32_t n : {0u, 1u, 32u, 192u, 131072u}) {
        CHECK(plan->run(&dummy, &dummy, n, 1) == -8);
        CHECK(plan->runInverse(&dummy, &dummy, n, 1) == -8);
        CHECK(plan->runR2C(&dummy, &dummy, n, 1) == -8);
        CHECK(plan->runC2R(&dummy, &dummy, n, 1) == -8);
        CHECK(plan->prepare(n, 1) == -8);
    }
    CHECK(plan->runC2R(&dummy, &dummy, 8192, 1) == -8);  // C2R envelope 64..4096
    CHECK(plan->runR2C(&dummy, &dummy, 64, 1) == -8);    // R2C envelope 128..8192
    CHECK(plan->runR2C(&dummy, &dummy, 16384, 1) == -8);
    for (auto run : {&bfly::Plan::run, &bfly::Plan::runR2C, &bfly::Plan::runC2R, &bfly::Plan::runInverse}) {
        CHECK((plan.get()->*run)(&dummy, &dummy, 128, 0) == -8);
        CHECK((plan.get()->*run)(nullptr, &dummy, 128, 1) == -8);
        CHECK((plan.get()->*run)(&dummy, nullptr, 128, 1) == -8);
        CHECK((plan.get()->*run)(&dummy, &dummy, 128, 0xffffffffu) == -8);
    }
    CHECK(plan->run(&dummy, &dummy, 4096, 524288) == -8);  // batch*2n = 2^32 > 2^32-1
    CHECK(plan->runDevice(&dummy, &dummy, 4096, 1, 0) == -8);
    bfly::Metric metric;
    CHECK(plan->measure(1, 1, &metric) == -8);
    CHECK(plan->prepare(64, 0) == -8);
    {
        const Counts c = counts();
        CHECK(c.allocations == 0 && c.copies == 0 && c.launches == 0);
    }
    for (const auto& c : context.enumerate(64, 0)) CHECK(c.state == bfly::State::Infeasible);

    // A failed allocation or copy while changing length must invalidate the old key.
    CHECK(plan->prepare(64, 1) == 0);
    bfly_stub_set_failures(counts().allocations + 1, 0, 0);
    CHECK(plan->prepare(128, 1) != 0);
    bfly_stub_set_failures(0, 0, 0);
    int before = counts().allocations;
    CHECK(plan->prepare(64, 1) == 0);
    CHECK(counts().allocations > before);
    bfly_stub_set_failures(0, counts().copies + 1, 0);
    CHECK(plan->prepare(128, 1) != 0);
    bfly_stub_set_failures(0, 0, 0);
    before = counts().allocations;
    CHECK(plan->prepare(64, 1) == 0);
    CHECK(counts().allocations > before);
    before = counts().allocations;
    CHECK(plan->prepare(64, 1) == 0);
    CHECK(counts().allocations == before);

    // Execute path error mapping through the stub: run copies in, launches,
    // syncs and copies out; an injected launch failure maps to -5, a copy
    // failure to -4 (in) / -7 (out).
    std::vector<float> in(2 * 256 * 3, 0.25f), out(2 * 256 * 3, 0.f);
    Counts c0 = counts();
    CHECK(plan->run(in.data(), out.data(), 256, 3) == 0);
    Counts c1 = counts();
    CHECK(c1.launches == c0.launches + 1 && c1.copies >= c0.copies + 2);
    bfly_stub_set_failures(0, 0, counts().launches + 1);
    CHECK(plan->run(in.data(), out.data(), 256, 3) == -5);
    bfly_stub_set_failures(0, counts().copies + 1, 0);
    CHECK(plan->run(in.data(), out.data(), 256, 3) == -4);
    bfly_stub_set_failures(0, counts().copies + 2, 0);
    CHECK(plan->run(in.data(), out.data(), 256, 3) == -7);
    bfly_stub_set_failures(0, 0, 0);
    std::vector<float> big(2 * 8192 * 3, 0.f), bigOut(2 * 8192 * 3, 0.f);
    CHECK(plan->run(big.data(), bigOut.data(), 8192, 3) == 0);  // long FFT lengths accepted
    CHECK(plan->runR2C(in.data(), out.data(), 256, 3) == 0);
    CHECK(plan->runC2R(in.data(), out.data(), 256, 2) == 0);

    std::printf("runtime guards: %d checks passed\n", checks);
    return 0;
}
