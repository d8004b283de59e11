// bfly_check -- correctness and timing check of one transform shape through
// the C ABI of libbfly_hip.so: the gfx950 counterpart of the reference's
// research driver build/fft_check (src/host/fft_check.cpp:213-1315), with
// the same arguments, environment switches, report lines and exit codes.
//
//   bfly_check <n> <batch> [reps]
//
// Environment (names and meanings of the reference's):
//   AB_DIR=c2c|r2c|c2r      transform (default c2c)
//   AB_INPUT=<mode>         input generator (default sin); C2R always takes
//                           the seeded half spectrum
//   AB_SEED=<u32>           seed of the random / random-seeded mode
//   AB_SGN=+1|-1            C2C twiddle sign (default -1; +1 is the
//                           unnormalised inverse, checked against the FP64
//                           inverse DFT).  R2C / C2R have fixed signs.
//   AB_INPUT_FILE=<path>    raw float32 input replacing the generated one
//   AB_INPUT_SEQ=a,b[,..]   C2C: run the same plan on each input (a mode
//                           name or a float32 file), check each against FP64
//                           and flag an output that did not change when its
//                           input did (stale buffer)
//   AB_E2E=<reps>           end-to-end loop: H2D, transform, D2H per rep;
//                           AB_E2E_MODE=async|sync|xfer|pipe, AB_E2E_HOST=pinned|pageable;
//                           pipe = bfly_host_pipelined in chunks of AB_E2E_CHUNK
//                           transforms (0: about 8 MiB of input), output checked
//   AB_DUMP=<prefix>        write <prefix>.in.bin and <prefix>.out.bin
//   AB_DEVICE_ID=<id>       device index
//   AB_BOUNDARY=device      the long chain always runs on the device here;
//                           host is refused
//   AB_LONG_BOUNDARY_IMPL   only "fused" is accepted (the chain is fused)
//   AB_FOLD_D, AB_PLANE_K   the Ascend batch-fold / K-plane overrides:
//                           validated as in the reference, no effect on gfx950
//   AB_VARIANT=<name|id>    gfx950 kernel variant (auto, plain, stream,
//                           plain_swz, stream_swz, stream_lt, stream2,
//                           four_step, split_mfma, split_mfma_stream)
//   AB_DESC=1               long chains: print the lowering descriptor of the
//                           chain this call runs (include/butterfly/descriptors.hpp)
//
// Lengths: C2C n a power of two in [16, 65536]; R2C / C2R in [32, 8192] --
// a superset of the reference's envelope (C2C 64..65536, R2C 128..8192,
// C2R 64..4096).  The C2R input rows are n+2 floats (the reference's host
// buffer uses n+16).  Long chains (C2C n > 4096) upload the input and read
// the output back inside every launch, as the reference's device chain
// does, so their scopes line reports h2d and d2h spans.
//
// Exit status: 0 PASS, 1 FAIL or runtime error, 2 bad arguments.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "bfly_hip.h"
#include "butterfly/descriptors.hpp"
#include "butterfly/reference.hpp"
#include "inputs.hpp"

namespace {

using Clock = std::chrono::steady_clock;

double us_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::micro>(Clock::now() - t0).count();
}

const char* env(const char* name) {
    const char* v = std::getenv(name);
    return v && *v ? v : nullptr;
}

int parse_variant(const char* s) {
    static const char* const kNames[] = {"auto",      "plain",   "stream",    "plain_swz",  "stream_swz",
                                         "stream_lt", "stream2", "four_step", "split_mfma", "split_mfma_stream"};
    for (int i = 0; i < (int)(sizeof(kNames) / sizeof(kNames[0])); i++)
        if (std::strcmp(s, kNames[i]) == 0) return i;
    char* end = nullptr;
    const long v = std::strtol(s, &end, 10);
    if (end && *end == 0 && v >= BFLY_VARIANT_AUTO && v <= BFLY_VARIANT_LAST) return (int)v;
    return -1;
}

bool read_floats(const char* path, std::vector<float>& v) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    const bool ok = std::fread(v.data(), sizeof(float), v.size(), f) == v.size();
    std::fclose(f);
    return ok;
}

void write_floats(const std::string& path, const std::vector<float>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::printf("dump open failed: %s\n", path.c_str());
        return;
    }
    std::fwrite(v.data(), sizeof(float), v.size(), f);
    std::fclose(f);
    std::printf("dumped %s (%zu floats)\n", path.c_str(), v.size());
}

std::string span(double us) {
    char b[40];
    if (us >= 0.0) std::snprintf(b, sizeof b, "%.1f us", us);
    else std::snprintf(b, sizeof b, "NA");
    return b;
}

struct Errors {
    double max_abs = 0.0, max_rel = 0.0;
    uint64_t worst = 0;
};

// Every row against the FP64 reference of the library (bfly::refFftF32 /
// refR2CF32 / refC2RF32): maxRelScaled per row (the PASS criterion, as in
// src/host/fft_check.cpp:1186-1248) and the largest absolute error with its
// flat index (diagnostic only).
Errors verify(int kind, int sign, const std::vector<float>& in, const std::vector<float>& out, uint32_t n,
              uint32_t batch) {
    Errors e;
    const size_t in_row = kind == BFLY_KIND_C2C ? 2 * (size_t)n : kind == BFLY_KIND_R2C ? n : n + 2;
    const size_t out_row = kind == BFLY_KIND_C2C ? 2 * (size_t)n : kind == BFLY_KIND_R2C ? n + 2 : n;
    std::vector<float> x(in_row), ref(out_row);
    for (uint32_t b = 0; b < batch; b++) {
        const float* row = in.data() + b * in_row;
        const float* got = out.data() + b * out_row;
        if (kind == BFLY_KIND_C2C && sign > 0) {
            // unnormalised inverse = conj(forward(conj(x)))
            for (size_t i = 0; i < in_row; i++) x[i] = (i & 1) ? -row[i] : row[i];
            bfly::refFftF32(x.data(), ref.data(), n);
            for (size_t i = 1; i < out_row; i += 2) ref[i] = -ref[i];
        } else if (kind == BFLY_KIND_C2C) {
            bfly::refFftF32(row, ref.data(), n);
        } else if (kind == BFLY_KIND_R2C) {
            bfly::refR2CF32(row, ref.data(), n);
        } else {
            bfly::refC2RF32(row, ref.data(), n);
        }
        e.max_rel = std::max(e.max_rel, bfly::maxRelScaled(got, ref.data(), out_row));
        const size_t step = kind == BFLY_KIND_C2R ? 1 : 2;
        for (size_t j = 0; j < out_row; j += step) {
            const double dr = (double)got[j] - ref[j];
            const double di = step == 2 ? (double)got[j + 1] - ref[j + 1] : 0.0;
            const double a = std::sqrt(dr * dr + di * di);
            if (a > e.max_abs || !(a == a)) {
                e.max_abs = a == a ? a : INFINITY;
                e.worst = (uint64_t)b * n + j / step;
            }
        }
    }
    return e;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("usage: %s <n> <batch> [reps]   (AB_DIR=c2c|r2c|c2r)\n", argv[0]);
        return 2;
    }
    const auto t_init0 = Clock::now();
    const uint32_t n = (uint32_t)std::strtoul(argv[1], nullptr, 10);
    const uint32_t batch = (uint32_t)std::strtoul(argv[2], nullptr, 10);
    const int reps = argc > 3 ? std::atoi(argv[3]) : 10;
    if (n < 8 || (n & (n - 1))) {
        std::printf("n must be power of two >= 8\n");
        return 2;
    }
    if (batch == 0 || reps <= 0) {
        std::printf("batch and reps must be >= 1\n");
        return 2;
    }
    const char* dir = env("AB_DIR") ? env("AB_DIR") : "c2c";
    int kind = -1;
    if (!std::strcmp(dir, "c2c")) kind = BFLY_KIND_C2C;
    else if (!std::strcmp(dir, "r2c")) kind = BFLY_KIND_R2C;
    else if (!std::strcmp(dir, "c2r")) kind = BFLY_KIND_C2R;
    if (kind < 0) {
        std::printf("AB_DIR must be c2c | r2c | c2r (got %s)\n", dir);
        return 2;
    }
    if (!bfly_supported(kind, n)) {
        std::printf("%s n=%u is outside the library's envelope (C2C 16..65536, R2C/C2R 32..8192, powers of two)\n",
                    dir, n);
        return 2;
    }
    const bool is_long = kind == BFLY_KIND_C2C && n > 4096;
    if (const char* b = env("AB_BOUNDARY")) {
        if (std::strcmp(b, "device")) {
            std::printf("AB_BOUNDARY must be device: the long chain runs on the device (got %s)\n", b);
            return 2;
        }
    }
    if (const char* impl = env("AB_LONG_BOUNDARY_IMPL")) {
        if (std::strcmp(impl, "fused")) {
            std::printf("AB_LONG_BOUNDARY_IMPL must be fused: the chain's boundary is fused into its kernels (got %s)\n",
                        impl);
            return 2;
        }
    }
    if (const char* k = env("AB_PLANE_K")) {
        const int v = std::atoi(k);
        if (v != 8 && v != 16 && v != 32) {
            std::printf("AB_PLANE_K must be one of 8|16|32, got %s\n", k);
            return 2;
        }
    }
    if (env("AB_PLANE_K") || env("AB_FOLD_D"))
        std::printf("note: AB_FOLD_D / AB_PLANE_K shape the Ascend kernel only; no effect on gfx950\n");
    int variant = BFLY_VARIANT_AUTO;
    if (const char* v = env("AB_VARIANT")) {
        variant = parse_variant(v);
        if (variant < 0) {
            std::printf("unknown AB_VARIANT %s\n", v);
            return 2;
        }
    }
    int sign = kind == BFLY_KIND_C2R ? 1 : -1;
    if (const char* s = env("AB_SGN")) {
        const int want = std::atof(s) > 0 ? 1 : -1;
        if (kind != BFLY_KIND_C2C && want != sign) {
            std::printf("AB_SGN=%s: %s has a fixed sign (%+d)\n", s, dir, sign);
            return 2;
        }
        sign = want;
    }
    const uint32_t seed = env("AB_SEED") ? (uint32_t)std::strtoul(env("AB_SEED"), nullptr, 0) : 0u;
    const char* mode = env("AB_INPUT") ? env("AB_INPUT") : "sin";
    if (!bfly_check::is_input_mode(mode)) {
        std::printf("unsupported AB_INPUT mode: %s\n", mode);
        return 2;
    }
    // AB_INPUT_SEQ tokens (checked before any device work)
    std::vector<std::string> seq;
    if (const char* sq = env("AB_INPUT_SEQ")) {
        if (kind != BFLY_KIND_C2C) {
            std::printf("AB_INPUT_SEQ only supported for c2c (set AB_DIR=c2c)\n");
            return 2;
        }
        for (const char* p = sq; *p;) {
            const char* c = std::strchr(p, ',');
            const size_t len = c ? (size_t)(c - p) : std::strlen(p);
            if (len) seq.emplace_back(p, len);
            if (!c) break;
            p = c + 1;
        }
        if (seq.size() < 2) {
            std::printf("AB_INPUT_SEQ needs at least two inputs (A,B[,A])\n");
            return 2;
        }
    }
    if (kind != BFLY_KIND_C2C && variant != BFLY_VARIANT_AUTO && variant != BFLY_VARIANT_PLAIN &&
        variant != BFLY_VARIANT_STREAM && variant != BFLY_VARIANT_STREAM_LT) {
        std::printf("AB_VARIANT %d does not apply to %s (auto, plain, stream, stream_lt)\n", variant, dir);
        return 2;
    }
    const int launches = kind == BFLY_KIND_C2C ? bfly_c2c_launches(n, batch, variant) : 1;
    if (launches < 0) {
        std::printf("AB_VARIANT %d does not apply to c2c n=%u batch=%u\n", variant, n, batch);
        return 2;
    }

    // long chains: the lowering of the chain this call runs, and its split
    // (src/host/fft_check.cpp:318-362)
    if (is_long) {
        const auto& hw = butterfly::builtin_hardware_profile();
        const butterfly::TransformSpec spec{n, batch, butterfly::Precision::FP32, butterfly::Direction::C2C_FWD};
        uint32_t n1 = n, n2 = 1;
        if (launches == 2) {
            uint32_t lg = 0;
            while ((1u << lg) < n) lg++;
            n1 = 1u << (lg / 2);
            n2 = n / n1;
        }
        const auto low = launches == 2
                             ? butterfly::query_lowering(spec, butterfly::default_long_mapping(n1, n2, hw),
                                                         butterfly::local_fft_capability(), hw)
                             : butterfly::query_lowering(spec, butterfly::single_pass_mapping(n, hw),
                                                         butterfly::single_fft_capability(), hw);
        if (!low.supported) {
            std::printf("lowering rejected: %s\n", low.reason.c_str());
            return 2;
        }
        if (env("AB_DESC")) {
            std::printf("lowering: launches=%d gm_boundaries=%d modeled_payload_gm=%llu impl=%s resident=%s "
                        "on_chip=%d host_assisted=%d abstract=%d kinds=",
                        low.visible_launches, low.materialized_gm_boundaries,
                        (unsigned long long)low.modeled_payload_gm_rw_bytes, launches == 2 ? "fused2" : "single",
                        butterfly::residence_name(low.resident_subgraph), (int)low.whole_transform_on_chip,
                        (int)low.host_assisted, (int)low.abstract_feasible);
            for (size_t i = 0; i < low.launch_manifest.size(); i++)
                std::printf("%s%s", i ? "|" : "", butterfly::launch_kind_name(low.launch_manifest[i].kind));
            std::printf("\n");
        }
        const uint64_t gm = (16ull + (launches == 2 ? 8ull : 0ull)) * n * batch;
        std::printf("long backend: n=%u split N1=%u N2=%u rows1=%llu rows2=%llu gm_bytes=%llu\n", n, n1, n2,
                    (unsigned long long)n2 * batch, launches == 2 ? (unsigned long long)n1 * batch : 0ull,
                    (unsigned long long)gm);
    }

    const size_t in_elems = kind == BFLY_KIND_C2C ? 2 * (size_t)n * batch
                            : kind == BFLY_KIND_R2C ? (size_t)n * batch
                                                    : (size_t)(n + 2) * batch;
    const size_t out_elems = kind == BFLY_KIND_C2C ? 2 * (size_t)n * batch
                             : kind == BFLY_KIND_R2C ? (size_t)(n + 2) * batch
                                                     : (size_t)n * batch;
    const size_t in_bytes = in_elems * sizeof(float), out_bytes = out_elems * sizeof(float);

    int dev = 0;
    if (const char* d = env("AB_DEVICE_ID")) dev = std::atoi(d);
    int rc = bfly_rt_init(dev);
    if (rc != BFLY_OK) {
        std::printf("device init failed: %d\n", rc);
        return 1;
    }
    void* stream = nullptr;
    if ((rc = bfly_rt_stream_create(&stream)) != BFLY_OK) {
        std::printf("stream create failed: %d\n", rc);
        return 1;
    }
    const double boot_us = us_since(t_init0);

    std::vector<float> h_in(in_elems, 0.0f), h_out(out_elems, 0.0f);
    if (kind == BFLY_KIND_C2R) bfly_check::half_input(h_in.data(), n, batch, n + 2);
    else if (kind == BFLY_KIND_R2C) bfly_check::r2c_input(h_in.data(), n, batch, mode, seed);
    else bfly_check::c2c_input(h_in.data(), n, batch, mode, seed);
    if (const char* f = env("AB_INPUT_FILE")) {
        if (!read_floats(f, h_in)) {
            std::printf("AB_INPUT_FILE read failed: %s\n", f);
            return 2;
        }
    }

    // plan setup: the input-independent table, device buffers, first upload
    const auto t_plan0 = Clock::now();
    std::vector<float> table(bfly_table_floats(kind, n));
    if ((rc = bfly_table_fill(kind, n, table.data())) != BFLY_OK) {
        std::printf("table fill failed: %d\n", rc);
        return 1;
    }
    const size_t work_bytes = launches == 2 ? 2 * (size_t)n * batch * sizeof(float) : 0;
    void *d_in = nullptr, *d_out = nullptr, *d_table = nullptr, *d_work = nullptr;
    if (bfly_rt_malloc(&d_in, in_bytes) != BFLY_OK || bfly_rt_malloc(&d_out, out_bytes) != BFLY_OK ||
        bfly_rt_malloc(&d_table, table.size() * sizeof(float)) != BFLY_OK ||
        (work_bytes && bfly_rt_malloc(&d_work, work_bytes) != BFLY_OK)) {
        std::printf("device allocation failed\n");
        return 1;
    }
    if (bfly_rt_memcpy_h2d(d_table, table.data(), table.size() * sizeof(float), stream) != BFLY_OK ||
        (!is_long && bfly_rt_memcpy_h2d(d_in, h_in.data(), in_bytes, stream) != BFLY_OK)) {
        std::printf("upload failed\n");
        return 1;
    }
    // events: h2d start, chain markers (launches + 1), d2h end
    std::vector<void*> ev_chain(launches + 1, nullptr);
    void *ev_in = nullptr, *ev_out = nullptr;
    bool ev_ok = bfly_rt_event_create(&ev_in) == BFLY_OK && bfly_rt_event_create(&ev_out) == BFLY_OK;
    for (auto& e : ev_chain) ev_ok = ev_ok && bfly_rt_event_create(&e) == BFLY_OK;
    if (!ev_ok) {
        std::printf("event create failed\n");
        return 1;
    }
    const double plan_us = us_since(t_plan0);

    // one execution: the kernels (short path) or H2D + chain + D2H (long
    // chain); fills the spans of the launch in ms, -1 where they do not apply
    const float* h2d_src = h_in.data();
    float* d2h_dst = h_out.data();
    double chain_ms = -1, h2d_ms = -1, d2h_ms = -1;
    std::vector<double> seg_ms(launches, -1.0);
    const float* dtab = static_cast<const float*>(d_table);
    auto launch = [&]() -> bool {
        int e = BFLY_OK;
        if (is_long) {
            e = bfly_rt_event_record(ev_in, stream);
            if (!e) e = bfly_rt_memcpy_h2d_async(d_in, h2d_src, in_bytes, stream);
        }
        if (!e) {
            if (kind == BFLY_KIND_C2C) {
                e = bfly_c2c_traced(static_cast<const float*>(d_in), static_cast<float*>(d_out), n, batch, sign,
                                    variant, dtab, d_work, stream, ev_chain.data(), (int)ev_chain.size());
            } else {
                e = bfly_rt_event_record(ev_chain[0], stream);
                if (!e)
                    e = bfly_real_ex(kind, static_cast<const float*>(d_in), static_cast<float*>(d_out), n, batch,
                                     n / 2 + 1, n + 2, variant, 0, dtab, stream);
                if (!e) e = bfly_rt_event_record(ev_chain[1], stream);
            }
        }
        if (is_long) {
            if (!e) e = bfly_rt_memcpy_d2h_async(d2h_dst, d_out, out_bytes, stream);
            if (!e) e = bfly_rt_event_record(ev_out, stream);
        }
        if (!e) e = bfly_rt_stream_sync(stream);
        if (e) {
            std::printf("launch=%d\n", e);
            return false;
        }
        float ms = -1.0f;
        chain_ms = bfly_rt_event_elapsed(ev_chain.front(), ev_chain.back(), &ms) == BFLY_OK ? ms : -1.0;
        for (int k = 0; k < launches; k++)
            seg_ms[k] = bfly_rt_event_elapsed(ev_chain[k], ev_chain[k + 1], &ms) == BFLY_OK ? ms : -1.0;
        if (is_long) {
            h2d_ms = bfly_rt_event_elapsed(ev_in, ev_chain.front(), &ms) == BFLY_OK ? ms : -1.0;
            d2h_ms = bfly_rt_event_elapsed(ev_chain.back(), ev_out, &ms) == BFLY_OK ? ms : -1.0;
        }
        return true;
    };

    const auto t_first = Clock::now();
    if (!launch()) return 1;
    const double first_us = us_since(t_first);

    // end-to-end loop (off by default)
    bool e2e_ok = true;
    if (const char* ee = env("AB_E2E")) {
        int ereps = std::atoi(ee);
        if (ereps <= 0) ereps = reps;
        const char* emode = env("AB_E2E_MODE");
        const bool use_async = !(emode && !std::strcmp(emode, "sync"));
        const bool xfer_only = emode && !std::strcmp(emode, "xfer");
        const bool pipe = emode && !std::strcmp(emode, "pipe");
        const uint32_t chunk = env("AB_E2E_CHUNK") ? (uint32_t)std::strtoul(env("AB_E2E_CHUNK"), nullptr, 10) : 0;
        // a long C2C under AUTO may run a chunk as the four-step chain
        if (pipe && kind == BFLY_KIND_C2C && !d_work && n > 4096 &&
            bfly_rt_malloc(&d_work, 2 * (size_t)n * batch * sizeof(float)) != BFLY_OK) {
            std::printf("E2E workspace allocation failed\n");
            return 1;
        }
        const char* hmode = env("AB_E2E_HOST");
        const bool want_pinned = !(hmode && !std::strcmp(hmode, "pageable"));
        float *p_in = nullptr, *p_out = nullptr;
        if (want_pinned) {
            void *a = nullptr, *b = nullptr;
            const int ea = bfly_rt_malloc_host(&a, in_bytes), eb = bfly_rt_malloc_host(&b, out_bytes);
            if (ea == BFLY_OK && eb == BFLY_OK) {
                p_in = static_cast<float*>(a);
                p_out = static_cast<float*>(b);
                std::memcpy(p_in, h_in.data(), in_bytes);
            } else {
                (void)bfly_rt_free_host(a);
                (void)bfly_rt_free_host(b);
                std::printf("E2E host=pinned alloc failed (%d/%d) -> pageable\n", ea, eb);
            }
        }
        const float* src = p_in ? p_in : h_in.data();
        float* dst = p_out ? p_out : h_out.data();
        h2d_src = src;
        d2h_dst = dst;
        double sum = 0.0, mn = 0.0, first = 0.0;
        const auto t0e = Clock::now();
        for (int i = 0; i < ereps; i++) {
            const auto a0 = Clock::now();
            if (pipe) {
                const int e = bfly_host_pipelined(kind, sign, variant, src, dst, n, batch, dtab, d_in, d_out, d_work,
                                                  chunk, stream);
                if (e) {
                    std::printf("E2E pipelined=%d\n", e);
                    return 1;
                }
            } else if (is_long && !xfer_only) {
                if (!launch()) return 1;
            } else {
                // async: queued copies with an explicit synchronisation; sync: the blocking copies
                int e = use_async ? bfly_rt_memcpy_h2d_async(d_in, src, in_bytes, stream)
                                  : bfly_rt_memcpy_h2d(d_in, src, in_bytes, stream);
                if (!e && use_async) e = bfly_rt_stream_sync(stream);
                if (e) {
                    std::printf("E2E h2d=%d\n", e);
                    return 1;
                }
                if (!xfer_only && !launch()) return 1;
                e = use_async ? bfly_rt_memcpy_d2h_async(dst, d_out, out_bytes, stream)
                              : bfly_rt_memcpy_d2h(dst, d_out, out_bytes, stream);
                if (!e && use_async) e = bfly_rt_stream_sync(stream);
                if (e) {
                    std::printf("E2E d2h=%d\n", e);
                    return 1;
                }
            }
            const double u = us_since(a0);
            sum += u;
            if (i == 0) first = mn = u;
            else mn = std::min(mn, u);
        }
        const double warm_us = std::chrono::duration<double, std::micro>(t0e - t_first).count();
        std::printf("E2E n=%u batch=%u reps=%d e2e_us=%.1f e2e_min_us=%.1f first_us=%.1f boot_us=%.1f plan_us=%.1f "
                    "warmup_us=%.1f mode=%s host=%s input=%s\n",
                    n, batch, ereps, sum / ereps, mn, first, boot_us, plan_us, warm_us,
                    pipe ? "pipe" : xfer_only ? "xfer" : use_async ? "async" : "sync", p_in ? "pinned" : "pageable",
                    mode);
        if (pipe) {
            // the pipelined call's last output against the FP64 reference
            const std::vector<float> got(dst, dst + out_elems);
            const Errors pe = verify(kind, sign, h_in, got, n, batch);
            e2e_ok = pe.max_rel <= 1e-4;
            std::printf("E2E pipe maxRel=%.3e %s\n", pe.max_rel, e2e_ok ? "PASS" : "FAIL");
        }
        // one logical input and one logical output per execution; the
        // chain's boundary never leaves the device
        std::printf("E2E transfers: in=1 out=1 boundary=0 per_execution\n");
        std::fflush(stdout);
        h2d_src = h_in.data();
        d2h_dst = h_out.data();
        if (p_in) {
            (void)bfly_rt_free_host(p_in);
            (void)bfly_rt_free_host(p_out);
        }
    }

    // timed launches: wall mean / min / median; the device spans of the
    // launch with the shortest chain, so its segments add up to device_chain
    double sum_us = 0.0, min_us = 0.0, chain_min = -1, h2d_min = -1, d2h_min = -1;
    std::vector<double> seg_min(launches, -1.0), samples;
    samples.reserve(reps);
    for (int i = 0; i < reps; i++) {
        const auto a0 = Clock::now();
        if (!launch()) return 1;
        const double u = us_since(a0);
        sum_us += u;
        if (i == 0 || u < min_us) min_us = u;
        samples.push_back(u);
        if (chain_ms >= 0 && (chain_min < 0 || chain_ms * 1000.0 < chain_min)) {
            chain_min = chain_ms * 1000.0;
            for (int k = 0; k < launches; k++) seg_min[k] = seg_ms[k] >= 0 ? seg_ms[k] * 1000.0 : -1.0;
        }
        if (h2d_ms >= 0 && (h2d_min < 0 || h2d_ms * 1000.0 < h2d_min)) h2d_min = h2d_ms * 1000.0;
        if (d2h_ms >= 0 && (d2h_min < 0 || d2h_ms * 1000.0 < d2h_min)) d2h_min = d2h_ms * 1000.0;
    }
    std::sort(samples.begin(), samples.end());
    const size_t m = samples.size() / 2;
    const double med_us = samples.size() % 2 ? samples[m] : 0.5 * (samples[m - 1] + samples[m]);
    std::printf("scopes: plan_setup=%.1f us first_use=%.1f us host_end_to_end mean=%.1f min=%.1f us h2d=%s "
                "device_chain=%s d2h=%s reps=%d\n",
                plan_us, first_us, sum_us / reps, min_us, span(h2d_min).c_str(), span(chain_min).c_str(),
                span(d2h_min).c_str(), reps);
    if (is_long) {
        // one fused launch, or the two fused launches of the four-step chain
        std::printf("boundary_impl: %s\n", launches == 2 ? "fused2" : "single");
        std::printf("segments: fft1=%.1f us", seg_min[0]);
        if (launches == 2) std::printf(" fft2=%.1f us", seg_min[1]);
        std::printf("\n");
    } else {
        std::printf("segments: NA\n");
    }

    if (!is_long && bfly_rt_memcpy_d2h(h_out.data(), d_out, out_bytes, stream) != BFLY_OK) {
        std::printf("d2h failed\n");
        return 1;
    }
    if (const char* dp = env("AB_DUMP")) {
        write_floats(std::string(dp) + ".in.bin", h_in);
        write_floats(std::string(dp) + ".out.bin", h_out);
    }
    const Errors err = verify(kind, sign, h_in, h_out, n, batch);

    // the same plan on a sequence of inputs, each checked, stale outputs flagged
    bool seq_ok = true;
    if (!seq.empty()) {
        std::vector<float> prev_in, prev_out;
        for (size_t t = 0; t < seq.size(); t++) {
            const std::string& tok = seq[t];
            if (bfly_check::is_input_mode(tok.c_str())) {
                bfly_check::c2c_input(h_in.data(), n, batch, tok.c_str(), seed);
            } else if (!read_floats(tok.c_str(), h_in)) {
                std::printf("AB_INPUT_SEQ[%zu] read failed: %s\n", t, tok.c_str());
                return 2;
            }
            if (!is_long && bfly_rt_memcpy_h2d(d_in, h_in.data(), in_bytes, stream) != BFLY_OK) return 1;
            if (!launch()) return 1;
            if (!is_long && bfly_rt_memcpy_d2h(h_out.data(), d_out, out_bytes, stream) != BFLY_OK) return 1;
            const Errors se = verify(kind, sign, h_in, h_out, n, batch);
            const bool ok = se.max_rel <= 1e-4;
            bool stale = false;
            if (!prev_in.empty()) stale = prev_in != h_in && prev_out == h_out;
            std::printf("seq[%zu]=%.80s maxRel=%.3e %s%s\n", t, tok.c_str(), se.max_rel, ok ? "PASS" : "FAIL",
                        stale ? " STALE-OUTPUT" : "");
            if (!ok || stale) seq_ok = false;
            prev_in = h_in;
            prev_out = h_out;
        }
        std::printf("seq: %zu inputs re-executed under one plan (no rebuild), stale-check on -> %s\n", seq.size(),
                    seq_ok ? "PASS" : "FAIL");
    }

    std::printf("n=%u batch=%u launches=%d  maxAbs=%.3e maxRel=%.3e (worst idx %llu)\n", n, batch, launches,
                err.max_abs, err.max_rel, (unsigned long long)err.worst);
    const char* kname = kind == BFLY_KIND_R2C ? "bfly_r2c" : kind == BFLY_KIND_C2R ? "bfly_c2r" : "bfly_c2c";
    std::printf("%s: %.1f us/call over %d reps (min %.1f us) med %.1f us\n", kname, sum_us / reps, reps, min_us,
                med_us);
    const bool pass = err.max_rel <= 1e-4 && seq_ok && e2e_ok;
    std::printf("%s\n", pass ? "PASS" : "FAIL");

    for (auto e : ev_chain) (void)bfly_rt_event_destroy(e);
    (void)bfly_rt_event_destroy(ev_in);
    (void)bfly_rt_event_destroy(ev_out);
    (void)bfly_rt_free(d_in);
    (void)bfly_rt_free(d_out);
    (void)bfly_rt_free(d_table);
    (void)bfly_rt_free(d_work);
    (void)bfly_rt_stream_destroy(stream);
    return pass ? 0 : 1;
}
