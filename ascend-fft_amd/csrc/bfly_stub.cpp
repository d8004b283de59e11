// bfly_stub.cpp -- CPU stub of the C ABI (include/bfly_hip.h) for host-only
// tests: the runtime calls are malloc/memcpy with counters and failure
// injection, the execute calls run the real argument guards and count a launch
// without computing anything.  Same role as the reference's fake AscendCL
// (tests/stubs/acl/acl_rt.h:15-53 of the reference): it lets the C++
// Context/Plan guard and cache-invalidation tests run without a GPU.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <new>

#include "bfly_hip.h"
#include "bfly_tables.hpp"

namespace {
struct Counters {
    int allocations = 0, frees = 0, copies = 0, launches = 0;
    int streams_created = 0, streams_destroyed = 0, inits = 0, syncs = 0;
    int fail_allocation = 0;  // the Nth allocation fails (0 = never)
    int fail_copy = 0;        // the Nth copy fails (0 = never)
    int fail_launch = 0;      // the Nth launch fails (0 = never)
    int last_c2c_variant = 0; // kernel of the last accepted C2C call (AUTO resolved)
} g;
int launch() {
    ++g.launches;
    return g.launches == g.fail_launch ? BFLY_ERR_LAUNCH : BFLY_OK;
}
}  // namespace

extern "C" {

// ----------------------------------------------------- stub-only controls
BFLY_API void bfly_stub_reset(void) { g = Counters{}; }
BFLY_API void bfly_stub_set_failures(int allocation, int copy, int launch_) {
    g.fail_allocation = allocation;
    g.fail_copy = copy;
    g.fail_launch = launch_;
}
// out[0..7] = allocations, frees, copies, launches, streams_created,
// streams_destroyed, inits, syncs
BFLY_API void bfly_stub_counters(int* out) {
    const int v[8] = {g.allocations, g.frees, g.copies, g.launches,
                      g.streams_created, g.streams_destroyed, g.inits, g.syncs};
    std::memcpy(out, v, sizeof(v));
}

// BFLY_VARIANT_* the last accepted C2C execute call ran (0 before any)
BFLY_API int bfly_stub_last_c2c_variant(void) { return g.last_c2c_variant; }

// ------------------------------------------------------------- runtime
int bfly_rt_init(int device) {
    ++g.inits;
    return device == 0 ? BFLY_OK : BFLY_ERR_TABLE;
}
int bfly_rt_teardown(void) {
    ++g.syncs;
    return BFLY_OK;
}
int bfly_rt_stream_create(void** stream) {
    if (!stream) return BFLY_ERR_SHAPE;
    ++g.streams_created;
    *stream = reinterpret_cast<void*>(static_cast<uintptr_t>(0x1000 + g.streams_created));
    return BFLY_OK;
}
int bfly_rt_stream_destroy(void* stream) {
    if (stream) ++g.streams_destroyed;
    return BFLY_OK;
}
int bfly_rt_stream_sync(void*) {
    ++g.syncs;
    return BFLY_OK;
}
// the stub's device: 16 GiB, reported free in full; larger requests fail
constexpr size_t kStubDeviceBytes = size_t(16) << 30;
int bfly_rt_mem_info(size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return BFLY_ERR_SHAPE;
    *free_bytes = *total_bytes = kStubDeviceBytes;
    return BFLY_OK;
}
int bfly_rt_malloc(void** ptr, size_t bytes) {
    if (!ptr) return BFLY_ERR_SHAPE;
    ++g.allocations;
    if (g.allocations == g.fail_allocation || bytes > kStubDeviceBytes) { *ptr = nullptr; return BFLY_ERR_ALLOC; }
    *ptr = std::malloc(bytes ? bytes : 16);
    return *ptr ? BFLY_OK : BFLY_ERR_ALLOC;
}
int bfly_rt_free(void* ptr) {
    if (ptr) { ++g.frees; std::free(ptr); }
    return BFLY_OK;
}
int bfly_rt_memcpy_h2d(void* dst, const void* src, size_t bytes, void*) {
    ++g.copies;
    if (g.copies == g.fail_copy || !dst || !src) return BFLY_ERR_H2D;
    std::memcpy(dst, src, bytes);
    return BFLY_OK;
}
int bfly_rt_memcpy_d2h(void* dst, const void* src, size_t bytes, void*) {
    ++g.copies;
    if (g.copies == g.fail_copy || !dst || !src) return BFLY_ERR_D2H;
    std::memcpy(dst, src, bytes);
    return BFLY_OK;
}

int bfly_rt_memcpy_h2d_async(void* dst, const void* src, size_t bytes, void* stream) {
    return bfly_rt_memcpy_h2d(dst, src, bytes, stream);
}
int bfly_rt_memcpy_d2h_async(void* dst, const void* src, size_t bytes, void* stream) {
    return bfly_rt_memcpy_d2h(dst, src, bytes, stream);
}
int bfly_rt_malloc_host(void** ptr, size_t bytes) { return bfly_rt_malloc(ptr, bytes); }
int bfly_rt_free_host(void* ptr) { return bfly_rt_free(ptr); }
// graphs: begin/end bracket nothing; a launch counts as one
int bfly_graph_begin(void* stream) { return stream ? BFLY_OK : BFLY_ERR_SHAPE; }
int bfly_graph_end(void* stream, void** graph) {
    if (!stream || !graph) return BFLY_ERR_SHAPE;
    *graph = new int(0);
    return BFLY_OK;
}
int bfly_graph_launch(void* graph, void*) { return graph ? launch() : BFLY_ERR_SHAPE; }
int bfly_graph_destroy(void* graph) {
    delete static_cast<int*>(graph);
    return BFLY_OK;
}
// events hold the host clock at their last record (ms)
int bfly_rt_event_create(void** event) {
    if (!event) return BFLY_ERR_SHAPE;
    *event = new double(0.0);
    return BFLY_OK;
}
int bfly_rt_event_destroy(void* event) {
    delete static_cast<double*>(event);
    return BFLY_OK;
}
int bfly_rt_event_record(void* event, void*) {
    if (!event) return BFLY_ERR_SHAPE;
    *static_cast<double*>(event) =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
    return BFLY_OK;
}
int bfly_rt_event_elapsed(void* start, void* end, float* ms) {
    if (!start || !end || !ms) return BFLY_ERR_SHAPE;
    *ms = (float)(*static_cast<double*>(end) - *static_cast<double*>(start));
    return BFLY_OK;
}

// ------------------------------------------------------------- execute
size_t bfly_c2c_columns_workspace_bytes(uint32_t n, uint32_t ncols, uint32_t batch) {
    return bfly_host::fft2d_layout(n).q > 1 ? 8 * (size_t)n * ncols * batch : 0;
}
int bfly_c2c_columns(const float* in, float* out, uint32_t n, uint32_t ncols, uint32_t batch, int sign,
                     const float* table_cols, void* work, void*) {
    if (int g = bfly_host::check_columns(sign, in, out, n, ncols, batch, table_cols, work)) return g;
    return launch();
}
int bfly_c2c_rows(const float* in, uint64_t idist, float* out, uint64_t odist, uint32_t n, uint32_t batch,
                  int sign, const float* table, void* work, void*) {
    if (int g = bfly_host::check_rows(sign, in, idist, out, odist, n, batch, table, work)) return g;
    return launch();
}
size_t bfly_fft2d_table_floats(uint32_t n1) { return 2 * bfly_host::fft2d_layout(n1).complex_len; }
int bfly_fft2d_table_fill(uint32_t n1, float* host_table) { return bfly_host::fft2d_fill(n1, host_table); }
size_t bfly_fft2d_workspace_bytes(uint32_t n1, uint32_t n2, uint32_t batch) {
    return bfly_host::fft2d_ok(n1, n2) ? 8 * (size_t)n1 * n2 * batch : 0;
}
int bfly_fft2d(const float* in, float* out, uint32_t n1, uint32_t n2, uint32_t batch, int sign,
               const float* table_rows, const float* table_cols, void* work, void*) {
    if (int g = bfly_host::check_fft2d(sign, in, out, n1, n2, batch, table_rows, table_cols, work)) return g;
    return launch();
}
int bfly_c2c_launches(uint32_t n, uint32_t batch, int variant) {
    if (!bfly_supported(BFLY_KIND_C2C, n) || batch == 0) return BFLY_ERR_SHAPE;
    if (variant < BFLY_VARIANT_AUTO || variant > BFLY_VARIANT_LAST) return BFLY_ERR_SHAPE;
    if ((variant == BFLY_VARIANT_SPLIT_MFMA || variant == BFLY_VARIANT_SPLIT_MFMA_STREAM) &&
        !bfly_host::split_mfma_ok(n))
        return BFLY_ERR_SHAPE;
    return bfly_host::needs_work(BFLY_KIND_C2C, n, batch, variant) ? 2 : 1;
}
int bfly_c2c_traced(const float* in, float* out, uint32_t n, uint32_t batch, int sign, int variant,
                    const float* table, void* work, void* stream, void* const* events, int nevents) {
    const int launches = bfly_c2c_launches(n, batch, variant);
    if (launches < 0) return launches;
    if (!events || nevents != launches + 1) return BFLY_ERR_SHAPE;
    for (int i = 0; i < nevents; i++)
        if (!events[i]) return BFLY_ERR_SHAPE;
    // events change only when the call is accepted
    double start = 0.0;
    (void)bfly_rt_event_record(&start, stream);
    const int rc = bfly_c2c_ex(in, out, n, batch, sign, variant, table, work, stream);
    if (rc != BFLY_OK) return rc;
    *static_cast<double*>(events[0]) = start;
    for (int i = 1; i < nevents; i++) (void)bfly_rt_event_record(events[i], stream);
    return BFLY_OK;
}
int bfly_c2c_ex(const float* in, float* out, uint32_t n, uint32_t batch, int sign, int variant,
                const float* table, void* work, void* stream) {
    (void)stream;
    if (int e = bfly_host::check_execute(BFLY_KIND_C2C, sign, in, out, n, batch, table, work, variant))
        return e;
    if (variant < BFLY_VARIANT_AUTO || variant > BFLY_VARIANT_LAST) return BFLY_ERR_SHAPE;
    if (variant == BFLY_VARIANT_AUTO) variant = bfly_host::auto_variant(n, batch);
    const bool uses_work = bfly_host::needs_work(BFLY_KIND_C2C, n, batch, variant);
    if (!bfly_host::disjoint(BFLY_KIND_C2C, in, out, uses_work ? work : nullptr, n, batch, 0, 0))
        return BFLY_ERR_SHAPE;
    g.last_c2c_variant = variant;
    return launch();
}
int bfly_noalloc_prepare(void) { return BFLY_OK; }
int bfly_c2c_noalloc(const float* in, float* out, uint32_t n, uint32_t batch, int sign, const float* table,
                     void* stream) {
    static unsigned char scratch[BFLY_NOALLOC_SCRATCH_BYTES];  // host stand-in: the stub computes nothing
    const bool known = bfly_supported(BFLY_KIND_C2C, n) && batch;
    const int variant = known ? bfly_host::auto_variant(n, batch) : BFLY_VARIANT_AUTO;
    const bool chain = known && bfly_host::needs_work(BFLY_KIND_C2C, n, batch, variant);
    if (int g = bfly_host::check_execute(BFLY_KIND_C2C, sign, in, out, n, batch, table, chain ? scratch : nullptr,
                                         variant))
        return g;
    if (chain && (uint64_t)n * batch * 8ull > sizeof(scratch)) return BFLY_ERR_SCRATCH;
    return bfly_c2c_ex(in, out, n, batch, sign, variant, table, chain ? scratch : nullptr, stream);
}
int bfly_c2c(const float* in, float* out, uint32_t n, uint32_t batch, int sign,
             const float* table, void* work, void* stream) {
    return bfly_c2c_ex(in, out, n, batch, sign, BFLY_VARIANT_AUTO, table, work, stream);
}
int bfly_c2c_scaled(const float* in, float* out, uint32_t n, uint32_t batch, int sign, float scale,
                    const float* table, void* work, void* stream) {
    if (!std::isfinite(scale)) return BFLY_ERR_SHAPE;
    if (!bfly_supported(BFLY_KIND_C2C, n)) return BFLY_ERR_SHAPE;
    return bfly_c2c_ex(in, out, n, batch, sign, bfly_host::scaled_variant(n, batch), table, work, stream);
}
int bfly_c2c_fwd(const float* in, float* out, uint32_t n, uint32_t batch, const float* table, void* work,
                 void* stream) {
    return bfly_c2c_ex(in, out, n, batch, -1, BFLY_VARIANT_AUTO, table, work, stream);
}
int bfly_c2c_inv(const float* in, float* out, uint32_t n, uint32_t batch, const float* table, void* work,
                 void* stream) {
    return bfly_c2c_ex(in, out, n, batch, 1, BFLY_VARIANT_AUTO, table, work, stream);
}
int bfly_r2c(const float* in, float* out, uint32_t n, uint32_t batch, const float* table, void* stream) {
    return bfly_real_ex(BFLY_KIND_R2C, in, out, n, batch, n / 2 + 1, n + 2, BFLY_VARIANT_AUTO, 0, table, stream);
}
int bfly_c2r(const float* in, float* out, uint32_t n, uint32_t batch, const float* table, void* stream) {
    return bfly_real_ex(BFLY_KIND_C2R, in, out, n, batch, n / 2 + 1, n + 2, BFLY_VARIANT_AUTO, 0, table, stream);
}
int bfly_real_ex(int kind, const float* in, float* out, uint32_t n, uint32_t batch, uint32_t bins,
                 uint32_t stride, int variant, uint32_t flags, const float* table, void*) {
    if (kind != BFLY_KIND_R2C && kind != BFLY_KIND_C2R) return BFLY_ERR_SHAPE;
    if (flags & ~BFLY_REAL_ADJOINT) return BFLY_ERR_SHAPE;
    if (int e = bfly_host::check_execute(kind, kind == BFLY_KIND_R2C ? -1 : 1, in, out, n, batch, table,
                                         nullptr))
        return e;
    if (!bfly_host::bins_ok(n, bins, stride)) return BFLY_ERR_SHAPE;
    if (!bfly_host::disjoint(kind, in, out, nullptr, n, batch, bins, stride)) return BFLY_ERR_SHAPE;
    if (variant != BFLY_VARIANT_AUTO && variant != BFLY_VARIANT_PLAIN && variant != BFLY_VARIANT_STREAM &&
        variant != BFLY_VARIANT_STREAM_LT)
        return BFLY_ERR_SHAPE;
    return launch();
}
int bfly_r2c_ex(const float* in, float* out, uint32_t n, uint32_t batch, uint32_t bins,
                uint32_t stride, const float* table, void* stream) {
    return bfly_real_ex(BFLY_KIND_R2C, in, out, n, batch, bins, stride, BFLY_VARIANT_AUTO, 0, table, stream);
}
int bfly_c2r_ex(const float* in, float* out, uint32_t n, uint32_t batch, uint32_t bins,
                uint32_t stride, const float* table, void* stream) {
    return bfly_real_ex(BFLY_KIND_C2R, in, out, n, batch, bins, stride, BFLY_VARIANT_AUTO, 0, table, stream);
}
// Same guards and chunking as the HIP library; the steps run one after the
// other (copy in, launch, copy out per chunk), each counted and failable.
int bfly_host_pipelined(int kind, int sign, int variant, const float* host_in, float* host_out, uint32_t n,
                        uint32_t batch, const float* table, void* dev_in, void* dev_out, void* work,
                        uint32_t chunk_rows, void* stream) {
    bfly_host::PipeShape sh;
    if (int e = bfly_host::check_pipelined(kind, sign, variant, host_in, host_out, n, batch, table, dev_in, dev_out,
                                           work, chunk_rows, &sh))
        return e;
    for (uint64_t r0 = 0; r0 < batch; r0 += sh.chunk) {
        const uint32_t rows = (uint32_t)(batch - r0 < sh.chunk ? batch - r0 : sh.chunk);
        float* din = static_cast<float*>(dev_in) + r0 * sh.row_in;
        float* dout = static_cast<float*>(dev_out) + r0 * sh.row_out;
        if (int e = bfly_rt_memcpy_h2d(din, host_in + r0 * sh.row_in, rows * sh.row_in * sizeof(float), stream))
            return e;
        const int e = kind == BFLY_KIND_C2C
                          ? bfly_c2c_ex(din, dout, n, rows, sh.sign, sh.variant, table, work, stream)
                          : bfly_real_ex(kind, din, dout, n, rows, n / 2 + 1, n + 2, sh.variant, 0, table, stream);
        if (e) return e;
        if (int e2 = bfly_rt_memcpy_d2h(host_out + r0 * sh.row_out, dout, rows * sh.row_out * sizeof(float), stream))
            return e2;
    }
    return BFLY_OK;
}

int bfly_rt_time_transform(int kind, int sign, const float* in, float* out, uint32_t n,
                           uint32_t batch, const float* table, void* work, int reps, void* stream,
                           float* ms) {
    if (!ms || reps <= 0) return BFLY_ERR_SHAPE;
    for (int i = 0; i < reps; i++) {
        int rc = kind == BFLY_KIND_C2C ? bfly_c2c(in, out, n, batch, sign, table, work, stream)
                 : kind == BFLY_KIND_R2C ? bfly_r2c(in, out, n, batch, table, stream)
                 : kind == BFLY_KIND_C2R ? bfly_c2r(in, out, n, batch, table, stream)
                                         : BFLY_ERR_SHAPE;
        if (rc) return rc;
    }
    *ms = 0.001f * (float)reps;
    return BFLY_OK;
}

int bfly_mode_mix(int op, const float* a, const float* b, float* c, uint32_t batch, uint32_t c_in,
                  uint32_t c_out, uint32_t modes, void*) {
    if (int e = bfly_host::check_mode_mix(op, a, b, c, batch, c_in, c_out, modes)) return e;
    return launch();
}

// --------------------------------------------------------------- probes
int bfly_probe_copy(const void* src, void* dst, size_t bytes, void*) {
    if (!src || !dst || bytes == 0 || (bytes & 15)) return BFLY_ERR_SHAPE;
    return launch();
}
int bfly_probe_device(int device, char* json, size_t cap) {
    if (!json || cap == 0 || device != 0) return BFLY_ERR_SHAPE;
    const int w = std::snprintf(json, cap, "{\"name\": \"cpu-stub\", \"gcn_arch\": \"none\"}");
    return (w > 0 && (size_t)w < cap) ? BFLY_OK : BFLY_ERR_SHAPE;
}


// ------------------------------------------------------- grouped launch
// The host plan (validation, unit list) is the library's own; the device
// part is a host allocation standing in for the request table.
namespace {
struct StubGroup {
    uint32_t nreq = 0, units = 0, grid = 0;
    size_t bytes = 0;
};
}  // namespace

int bfly_group_supported(int kind, uint32_t n, int workgroup) {
    return bfly_host::group_ok(kind, n, workgroup) ? 1 : 0;
}
int bfly_group_create(const bfly_group_request* reqs, int nreq, const bfly_group_config* cfg,
                      void** group) {
    if (!group) return BFLY_ERR_SHAPE;
    *group = nullptr;
    bfly_host::GroupPlan P;
    if (int e = bfly_host::group_plan(reqs, nreq, cfg, &P)) return e;
    ++g.allocations;
    if (g.allocations == g.fail_allocation) return BFLY_ERR_ALLOC;
    StubGroup* G = new (std::nothrow) StubGroup();
    if (!G) return BFLY_ERR_ALLOC;
    G->nreq = (uint32_t)nreq;
    G->units = (uint32_t)P.total;
    const uint64_t cap = P.cfg.grid > 0 ? (uint64_t)P.cfg.grid : 512u;  // 256 CUs x 2 workgroups
    G->grid = (uint32_t)(cap < P.total ? cap : P.total);
    G->bytes = 48u * (size_t)nreq + 8u * (size_t)P.total;
    ++g.copies;
    if (g.copies == g.fail_copy) {
        delete G;
        ++g.frees;
        return BFLY_ERR_H2D;
    }
    *group = G;
    return BFLY_OK;
}
int bfly_group_launch(void* group, unsigned long long*, void*) {
    if (!group) return BFLY_ERR_SHAPE;
    return launch();
}
int bfly_group_query(void* group, uint32_t* units, uint32_t* grid, size_t* bytes) {
    const StubGroup* G = static_cast<const StubGroup*>(group);
    if (!G) return BFLY_ERR_SHAPE;
    if (units) *units = G->units;
    if (grid) *grid = G->grid;
    if (bytes) *bytes = G->bytes;
    return BFLY_OK;
}
int bfly_group_destroy(void* group) {
    if (group) {
        ++g.frees;
        delete static_cast<StubGroup*>(group);
    }
    return BFLY_OK;
}

}  // extern "C"
