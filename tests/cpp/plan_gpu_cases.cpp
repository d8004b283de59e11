// Drives the drop-in C++ API (bfly::Context / bfly::Plan of
// include/butterfly/plan.hpp) on the GPU, the way a caller of the reference's
// API does: Context::init, Context::select (enumerate -> rank -> measure),
// Plan::run / runInverse / runR2C / runC2R with host buffers, runDevice with
// device buffers, measure, and the shape guard.
//
// Usage: bfly_plan_cases <config dir> <output dir>
// Writes every case's input and output as raw float32 files
// (<case>_in.f32, <case>_out.f32) into <output dir> and prints one line per
// case, "case <name> n <n> batch <b> rc <rc> maxrel <e>", with the error
// against the library's FP64 reference (bfly::refFftF32 / refR2CF32 /
// refC2RF32, the reference's maxRelScaled metric).  Exit status 0 when every
// call returned its expected code.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "bfly_hip.h"
#include "butterfly/plan.hpp"
#include "butterfly/reference.hpp"

namespace {

std::string g_dir;
int g_fail = 0;

void dump(const std::string& name, const std::vector<float>& v) {
    const std::string path = g_dir + "/" + name;
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(float), v.size(), f) != v.size()) {
        std::printf("cannot write %s\n", path.c_str());
        g_fail++;
    }
    if (f) std::fclose(f);
}

// uniform [-1, 1) from a fixed xorshift stream
struct Rng {
    uint64_t s;
    float next() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return (float)((double)(s >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0);
    }
};

std::vector<float> uniform(Rng& r, size_t len) {
    std::vector<float> v(len);
    for (auto& x : v) x = r.next();
    return v;
}

void report(const char* name, uint32_t n, uint32_t b, int rc, int want, double err) {
    std::printf("case %s n %u batch %u rc %d maxrel %.3e\n", name, n, b, rc, err);
    if (rc != want) {
        std::printf("  unexpected rc (want %d)\n", want);
        g_fail++;
    }
}

// worst maxRelScaled over the rows, against ref(row_in, row_out, n)
template <class Ref>
double rows_err(const std::vector<float>& in, const std::vector<float>& out, uint32_t n, uint32_t b,
                size_t in_row, size_t out_row, Ref ref) {
    std::vector<float> want(out_row);
    double worst = 0.0;
    for (uint32_t r = 0; r < b; r++) {
        ref(in.data() + r * in_row, want.data(), n);
        worst = std::fmax(worst, bfly::maxRelScaled(out.data() + r * out_row, want.data(), out_row));
    }
    return worst;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("usage: %s <config dir> <output dir>\n", argv[0]);
        return 2;
    }
    const std::string cfg = argv[1];
    g_dir = argv[2];
    bfly::Context ctx;
    int rc = ctx.init(cfg + "/gfx950_profile.json", cfg + "/butterfly_space.json", "libbfly_hip.so");
    std::printf("init rc %d\n", rc);
    if (rc != 0) return 1;
    Rng rng{0x9E3779B97F4A7C15ull};

    // C2C forward through the measured selection, local and long lengths
    const uint32_t shapes[][2] = {{64, 5}, {1024, 64}, {4096, 16}, {8192, 3}, {16384, 2}, {65536, 2}};
    for (const auto& s : shapes) {
        const uint32_t n = s[0], b = s[1];
        std::string log;
        auto plan = ctx.select(n, b, 3, &log);
        if (!plan) {
            std::printf("case c2c n %u batch %u select failed\n%s", n, b, log.c_str());
            g_fail++;
            continue;
        }
        auto in = uniform(rng, 2ull * n * b);
        std::vector<float> out(in.size());
        rc = plan->run(in.data(), out.data(), n, b);
        report("c2c", n, b, rc, 0, rows_err(in, out, n, b, 2 * n, 2 * n, bfly::refFftF32));
        const std::string tag = "c2c_n" + std::to_string(n) + "_b" + std::to_string(b);
        dump(tag + "_in.f32", in);
        dump(tag + "_out.f32", out);
        // unnormalised inverse of the forward output gives n x the input
        std::vector<float> back(in.size());
        rc = plan->runInverse(out.data(), back.data(), n, b);
        double worst = 0.0;
        for (size_t i = 0; i < in.size(); i++) worst = std::fmax(worst, std::fabs(back[i] / n - in[i]));
        report("c2c_inverse_round_trip", n, b, rc, 0, worst);
    }

    // R2C / C2R with one plan; the C2R input is the R2C output
    {
        auto plan = ctx.select(1024, 8, 3, nullptr);
        for (uint32_t n : {128u, 1024u, 4096u, 8192u}) {
            const uint32_t b = 7;
            auto x = uniform(rng, (size_t)n * b);
            std::vector<float> h((size_t)(n + 2) * b);
            rc = plan->runR2C(x.data(), h.data(), n, b);
            report("r2c", n, b, rc, 0, rows_err(x, h, n, b, n, n + 2, bfly::refR2CF32));
            const std::string tag = "r2c_n" + std::to_string(n) + "_b" + std::to_string(b);
            dump(tag + "_in.f32", x);
            dump(tag + "_out.f32", h);
            if (n > 4096) continue;
            std::vector<float> y(x.size());
            rc = plan->runC2R(h.data(), y.data(), n, b);
            report("c2r", n, b, rc, 0, rows_err(h, y, n, b, n + 2, n, bfly::refC2RF32));
            const std::string t2 = "c2r_n" + std::to_string(n) + "_b" + std::to_string(b);
            dump(t2 + "_in.f32", h);
            dump(t2 + "_out.f32", y);
        }
        // shape guards: not a power of two, C2R above 4096, R2C below 128
        std::vector<float> buf(1 << 16);
        report("guard_c2c_n100", 100, 1, plan->run(buf.data(), buf.data() + 8192, 100, 1), -8, 0.0);
        report("guard_c2r_n8192", 8192, 1, plan->runC2R(buf.data(), buf.data() + 32768, 8192, 1), -8, 0.0);
        report("guard_r2c_n64", 64, 1, plan->runR2C(buf.data(), buf.data() + 8192, 64, 1), -8, 0.0);
    }

    // device pointers on the context stream
    {
        const uint32_t n = 4096, b = 33;
        auto plan = ctx.select(n, b, 3, nullptr);
        auto in = uniform(rng, 2ull * n * b);
        std::vector<float> out(in.size());
        const size_t bytes = in.size() * sizeof(float);
        void *din = nullptr, *dout = nullptr;
        if (bfly_rt_malloc(&din, bytes) != BFLY_OK || bfly_rt_malloc(&dout, bytes) != BFLY_OK) {
            std::printf("device allocation failed\n");
            return 1;
        }
        rc = bfly_rt_memcpy_h2d(din, in.data(), bytes, ctx.stream());
        if (rc == BFLY_OK) rc = plan->runDevice(static_cast<const float*>(din), static_cast<float*>(dout), n, b);
        if (rc == BFLY_OK) rc = bfly_rt_memcpy_d2h(out.data(), dout, bytes, ctx.stream());
        if (rc == BFLY_OK) rc = bfly_rt_stream_sync(ctx.stream());
        report("c2c_device", n, b, rc, 0, rows_err(in, out, n, b, 2 * n, 2 * n, bfly::refFftF32));
        // overlapping device buffers are refused
        report("guard_device_in_place", n, b,
               plan->runDevice(static_cast<const float*>(din), static_cast<float*>(din), n, b), -8, 0.0);
        (void)bfly_rt_free(din);
        (void)bfly_rt_free(dout);
        // the measured figure of the selected plan
        bfly::Metric m;
        rc = plan->measure(n, b, &m);
        std::printf("case measure n %u batch %u rc %d us %.2f maxrel %.3e ok %d\n", n, b, rc, m.measuredUs,
                    m.maxRelErr, (int)m.measuredOk);
        if (rc != 0 || !m.measuredOk || !(m.measuredUs > 0.0)) g_fail++;
    }

    // the measured selection at the headline shape: every lowered kernel is
    // measured (topK = 9) and its log printed line by line
    {
        std::string log;
        auto plan = ctx.select(4096, 4096, 9, &log);
        std::istringstream lines(log);
        for (std::string l; std::getline(lines, l);) std::printf("selectlog %s\n", l.c_str());
        if (!plan) g_fail++;
        else std::printf("case select n 4096 batch 4096 kernel %s us %.2f\n", bfly::kernelName(plan->candidate().kernel),
                         plan->candidate().q.measuredUs);
    }

    std::printf("plan gpu cases: %s (%d failures)\n", g_fail ? "FAILED" : "passed", g_fail);
    return g_fail ? 1 : 0;
}
