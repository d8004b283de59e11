// bfly_capi.cpp -- the exported C ABI of libbfly_hip.so (include/bfly_hip.h)
// on the HIP runtime: argument guards, runtime wrappers and kernel dispatch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "bfly_hip.h"
#include "bfly_kernels.hpp"
#include "bfly_plans.hpp"
#include "bfly_tables.hpp"

namespace {

inline hipStream_t S(void* stream) { return reinterpret_cast<hipStream_t>(stream); }

int launched(hipError_t e) { return e == hipSuccess ? BFLY_OK : BFLY_ERR_LAUNCH; }

// The transform and download streams of bfly_host_pipelined, and a ring of
// events ordering each chunk's steps, created once per device (default
// priority) and kept until the process ends.  An event is recorded again
// only after the wait that used its previous recording was issued.
constexpr int kPipeRing = 32;
constexpr int kPipeDevices = 64;
struct PipeStreams {
    hipStream_t comp = nullptr, down = nullptr;
    hipEvent_t ev[2 * kPipeRing] = {};
    bool ready = false;
};
PipeStreams g_pipe[kPipeDevices];
std::mutex g_pipe_lock[kPipeDevices];

// Creates the device's streams and events on first use (lock held).
int pipe_ready(PipeStreams& P) {
    if (P.ready) return BFLY_OK;
    if (!P.comp && hipStreamCreateWithFlags(&P.comp, hipStreamNonBlocking) != hipSuccess) {
        P.comp = nullptr;
        (void)hipGetLastError();
        return BFLY_ERR_ALLOC;
    }
    if (!P.down && hipStreamCreateWithFlags(&P.down, hipStreamNonBlocking) != hipSuccess) {
        P.down = nullptr;
        (void)hipGetLastError();
        return BFLY_ERR_ALLOC;
    }
    for (hipEvent_t& e : P.ev)
        if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            e = nullptr;
            (void)hipGetLastError();
            return BFLY_ERR_ALLOC;
        }
    P.ready = true;
    // a first kernel on the transform stream, a first recording of every
    // event and a first download on the download stream (large enough to go
    // through the copy engine: small copies run as kernels), so that their
    // set-up is paid here rather than in the first call (best effort: a
    // failure leaves that set-up to the first call)
    constexpr size_t kWarm = 4u << 20;
    void *d = nullptr, *h = nullptr;
    if (hipMalloc(&d, kWarm) == hipSuccess && hipHostMalloc(&h, kWarm, hipHostMallocDefault) == hipSuccess) {
        (void)hipMemsetAsync(d, 0, kWarm, P.comp);
        for (int i = 0; i < 2 * kPipeRing; i++) {
            (void)hipEventRecord(P.ev[i], i & 1 ? P.down : P.comp);
            (void)hipStreamWaitEvent(i & 1 ? P.comp : P.down, P.ev[i], 0);
        }
        (void)hipMemcpyAsync(h, d, kWarm, hipMemcpyDeviceToHost, P.down);
        (void)hipStreamSynchronize(P.down);
        (void)hipStreamSynchronize(P.comp);
    }
    if (h) (void)hipHostFree(h);
    if (d) (void)hipFree(d);
    (void)hipGetLastError();
    return BFLY_OK;
}

}  // namespace

extern "C" {

int bfly_rt_init(int device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return BFLY_ERR_NOT_LOADED;
    if (device < 0 || device >= count) return BFLY_ERR_TABLE;
    if (hipSetDevice(device) != hipSuccess) return BFLY_ERR_TABLE;
    // the pipelined host call's streams and events, so that its first call
    // does not create them (a failure here is retried by that call)
    if (device < kPipeDevices) {
        std::lock_guard<std::mutex> lock(g_pipe_lock[device]);
        (void)pipe_ready(g_pipe[device]);
    }
    // the allocation-free entry point's scratch address, once per device
    return bfly_noalloc_prepare() == BFLY_OK ? BFLY_OK : BFLY_ERR_TABLE;
}

int bfly_rt_teardown(void) { return hipDeviceSynchronize() == hipSuccess ? BFLY_OK : BFLY_ERR_SYNC; }

int bfly_rt_stream_create(void** stream) {
    if (!stream) return BFLY_ERR_SHAPE;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return BFLY_ERR_ALLOC;
    *stream = s;
    return BFLY_OK;
}

int bfly_rt_stream_destroy(void* stream) {
    if (!stream) return BFLY_OK;
    return hipStreamDestroy(S(stream)) == hipSuccess ? BFLY_OK : BFLY_ERR_SYNC;
}

int bfly_rt_stream_sync(void* stream) {
    return hipStreamSynchronize(S(stream)) == hipSuccess ? BFLY_OK : BFLY_ERR_SYNC;
}

int bfly_rt_malloc(void** ptr, size_t bytes) {
    if (!ptr) return BFLY_ERR_SHAPE;
    *ptr = nullptr;
    if (hipMalloc(ptr, bytes ? bytes : 16) != hipSuccess) {
        // consume the runtime's last-error record: a refused allocation must
        // not surface as the failure of the next kernel launch
        (void)hipGetLastError();
        *ptr = nullptr;
        return BFLY_ERR_ALLOC;
    }
    return BFLY_OK;
}

int bfly_rt_mem_info(size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return BFLY_ERR_SHAPE;
    return hipMemGetInfo(free_bytes, total_bytes) == hipSuccess ? BFLY_OK : BFLY_ERR_NOT_LOADED;
}

int bfly_rt_free(void* ptr) {
    if (!ptr) return BFLY_OK;
    return hipFree(ptr) == hipSuccess ? BFLY_OK : BFLY_ERR_ALLOC;
}

int bfly_rt_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
    if (!dst || !src) return BFLY_ERR_H2D;
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream)) != hipSuccess)
        return BFLY_ERR_H2D;
    return hipStreamSynchronize(S(stream)) == hipSuccess ? BFLY_OK : BFLY_ERR_H2D;
}

int bfly_rt_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
    if (!dst || !src) return BFLY_ERR_D2H;
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)) != hipSuccess)
        return BFLY_ERR_D2H;
    return hipStreamSynchronize(S(stream)) == hipSuccess ? BFLY_OK : BFLY_ERR_D2H;
}

int bfly_rt_memcpy_h2d_async(void* dst, const void* src, size_t bytes, void* stream) {
    if (!dst || !src) return BFLY_ERR_H2D;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream)) == hipSuccess ? BFLY_OK
                                                                                          : BFLY_ERR_H2D;
}

int bfly_rt_memcpy_d2h_async(void* dst, const void* src, size_t bytes, void* stream) {
    if (!dst || !src) return BFLY_ERR_D2H;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)) == hipSuccess ? BFLY_OK
                                                                                          : BFLY_ERR_D2H;
}

int bfly_rt_malloc_host(void** ptr, size_t bytes) {
    if (!ptr) return BFLY_ERR_SHAPE;
    *ptr = nullptr;
    if (hipHostMalloc(ptr, bytes ? bytes : 16, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return BFLY_ERR_ALLOC;
    }
    return BFLY_OK;
}

int bfly_rt_free_host(void* ptr) {
    if (!ptr) return BFLY_OK;
    return hipHostFree(ptr) == hipSuccess ? BFLY_OK : BFLY_ERR_ALLOC;
}

int bfly_graph_begin(void* stream) {
    if (!stream) return BFLY_ERR_SHAPE;
    return hipStreamBeginCapture(S(stream), hipStreamCaptureModeThreadLocal) == hipSuccess ? BFLY_OK
                                                                                          : BFLY_ERR_LAUNCH;
}

int bfly_graph_end(void* stream, void** graph) {
    if (!stream || !graph) return BFLY_ERR_SHAPE;
    *graph = nullptr;
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(S(stream), &g) != hipSuccess || !g) {
        (void)hipGetLastError();
        if (g) (void)hipGraphDestroy(g);
        return BFLY_ERR_LAUNCH;
    }
    hipGraphExec_t x = nullptr;
    const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return BFLY_ERR_ALLOC;
    }
    *graph = x;
    return BFLY_OK;
}

int bfly_graph_launch(void* graph, void* stream) {
    if (!graph) return BFLY_ERR_SHAPE;
    return hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(graph), S(stream)) == hipSuccess ? BFLY_OK
                                                                                             : BFLY_ERR_LAUNCH;
}

int bfly_graph_destroy(void* graph) {
    if (!graph) return BFLY_OK;
    return hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(graph)) == hipSuccess ? BFLY_OK : BFLY_ERR_SYNC;
}

int bfly_rt_event_create(void** event) {
    if (!event) return BFLY_ERR_SHAPE;
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return BFLY_ERR_ALLOC;
    *event = e;
    return BFLY_OK;
}

int bfly_rt_event_destroy(void* event) {
    if (!event) return BFLY_OK;
    return hipEventDestroy(reinterpret_cast<hipEvent_t>(event)) == hipSuccess ? BFLY_OK : BFLY_ERR_SYNC;
}

int bfly_rt_event_record(void* event, void* stream) {
    if (!event) return BFLY_ERR_SHAPE;
    return hipEventRecord(reinterpret_cast<hipEvent_t>(event), S(stream)) == hipSuccess ? BFLY_OK
                                                                                         : BFLY_ERR_LAUNCH;
}

int bfly_rt_event_elapsed(void* start, void* end, float* ms) {
    if (!start || !end || !ms) return BFLY_ERR_SHAPE;
    hipEvent_t a = reinterpret_cast<hipEvent_t>(start), b = reinterpret_cast<hipEvent_t>(end);
    if (hipEventSynchronize(b) != hipSuccess) return BFLY_ERR_SYNC;
    return hipEventElapsedTime(ms, a, b) == hipSuccess ? BFLY_OK : BFLY_ERR_SYNC;
}

size_t bfly_c2c_columns_workspace_bytes(uint32_t n, uint32_t ncols, uint32_t batch) {
    return bfly_host::fft2d_layout(n).q > 1 ? 8 * (size_t)n * ncols * batch : 0;
}

int bfly_c2c_columns(const float* in, float* out, uint32_t n, uint32_t ncols, uint32_t batch, int sign,
                     const float* table_cols, void* work, void* stream) {
    if (int g = bfly_host::check_columns(sign, in, out, n, ncols, batch, table_cols, work)) return g;
    return launched(bfly_kern::columns(in, out, n, ncols, batch, sign, table_cols, work, S(stream)));
}

int bfly_c2c_rows(const float* in, uint64_t idist, float* out, uint64_t odist, uint32_t n, uint32_t batch,
                  int sign, const float* table, void* work, void* stream) {
    if (int g = bfly_host::check_rows(sign, in, idist, out, odist, n, batch, table, work)) return g;
    return launched(bfly_kern::c2c_rows(in, idist, out, odist, n, batch, sign, table, work, S(stream)));
}

size_t bfly_fft2d_table_floats(uint32_t n1) { return 2 * bfly_host::fft2d_layout(n1).complex_len; }

int bfly_fft2d_table_fill(uint32_t n1, float* host_table) { return bfly_host::fft2d_fill(n1, host_table); }

size_t bfly_fft2d_workspace_bytes(uint32_t n1, uint32_t n2, uint32_t batch) {
    return bfly_host::fft2d_ok(n1, n2) ? 8 * (size_t)n1 * n2 * batch : 0;
}

int bfly_fft2d(const float* in, float* out, uint32_t n1, uint32_t n2, uint32_t batch, int sign,
               const float* table_rows, const float* table_cols, void* work, void* stream) {
    if (int g = bfly_host::check_fft2d(sign, in, out, n1, n2, batch, table_rows, table_cols, work)) return g;
    return launched(bfly_kern::fft2d(in, out, n1, n2, batch, sign, table_rows, table_cols, work, S(stream)));
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
    if (int g = bfly_host::check_execute(BFLY_KIND_C2C, sign, in, out, n, batch, table, work, variant))
        return g;
    if (variant == BFLY_VARIANT_AUTO) variant = bfly_host::auto_variant(n, batch);
    const bool uses_work = launches == 2;
    if (!bfly_host::disjoint(BFLY_KIND_C2C, in, out, uses_work ? work : nullptr, n, batch, 0, 0))
        return BFLY_ERR_SHAPE;
    auto ev = [&](int i) { return reinterpret_cast<hipEvent_t>(events[i]); };
    if (hipEventRecord(ev(0), S(stream)) != hipSuccess) return BFLY_ERR_LAUNCH;
    const int rc = launched(bfly_kern::c2c(in, out, n, batch, sign, variant, table, work, S(stream),
                                           launches == 2 ? ev(1) : nullptr));
    if (rc != BFLY_OK) return rc;
    return hipEventRecord(ev(launches), S(stream)) == hipSuccess ? BFLY_OK : BFLY_ERR_LAUNCH;
}

static int c2c_checked(const float* in, float* out, uint32_t n, uint32_t batch, int sign, int variant,
                       const float* table, void* work, void* stream, float scale) {
    if (int g = bfly_host::check_execute(BFLY_KIND_C2C, sign, in, out, n, batch, table, work, variant))
        return g;
    if (variant < BFLY_VARIANT_AUTO || variant > BFLY_VARIANT_LAST) return BFLY_ERR_SHAPE;
    if (variant == BFLY_VARIANT_AUTO) variant = bfly_host::auto_variant(n, batch);
    const bool uses_work = bfly_host::needs_work(BFLY_KIND_C2C, n, batch, variant);
    if (!bfly_host::disjoint(BFLY_KIND_C2C, in, out, uses_work ? work : nullptr, n, batch, 0, 0))
        return BFLY_ERR_SHAPE;
    return launched(bfly_kern::c2c(in, out, n, batch, sign, variant, table, work, S(stream), nullptr, scale));
}

int bfly_c2c_ex(const float* in, float* out, uint32_t n, uint32_t batch, int sign, int variant,
                const float* table, void* work, void* stream) {
    return c2c_checked(in, out, n, batch, sign, variant, table, work, stream, 1.0f);
}

int bfly_c2c_scaled(const float* in, float* out, uint32_t n, uint32_t batch, int sign, float scale,
                    const float* table, void* work, void* stream) {
    if (!std::isfinite(scale)) return BFLY_ERR_SHAPE;
    // validate n before it picks a variant; check_execute repeats it
    if (!bfly_supported(BFLY_KIND_C2C, n)) return BFLY_ERR_SHAPE;
    return c2c_checked(in, out, n, batch, sign, bfly_host::scaled_variant(n, batch), table, work, stream, scale);
}

int bfly_c2c(const float* in, float* out, uint32_t n, uint32_t batch, int sign,
             const float* table, void* work, void* stream) {
    return bfly_c2c_ex(in, out, n, batch, sign, BFLY_VARIANT_AUTO, table, work, stream);
}

int bfly_c2c_fwd(const float* in, float* out, uint32_t n, uint32_t batch, const float* table, void* work,
                 void* stream) {
    return bfly_c2c_ex(in, out, n, batch, -1, BFLY_VARIANT_AUTO, table, work, stream);
}

int bfly_c2c_inv(const float* in, float* out, uint32_t n, uint32_t batch, const float* table, void* work,
                 void* stream) {
    return bfly_c2c_ex(in, out, n, batch, 1, BFLY_VARIANT_AUTO, table, work, stream);
}

int bfly_real_ex(int kind, const float* in, float* out, uint32_t n, uint32_t batch, uint32_t bins,
                 uint32_t stride, int variant, uint32_t flags, const float* table, void* stream) {
    if (kind != BFLY_KIND_R2C && kind != BFLY_KIND_C2R) return BFLY_ERR_SHAPE;
    if (flags & ~BFLY_REAL_ADJOINT) return BFLY_ERR_SHAPE;
    const int sign = kind == BFLY_KIND_R2C ? -1 : 1;
    if (int g = bfly_host::check_execute(kind, sign, in, out, n, batch, table, nullptr)) return g;
    if (!bfly_host::bins_ok(n, bins, stride)) return BFLY_ERR_SHAPE;
    if (!bfly_host::disjoint(kind, in, out, nullptr, n, batch, bins, stride)) return BFLY_ERR_SHAPE;
    if (variant == BFLY_VARIANT_AUTO) variant = bfly_host::auto_real_variant(n, batch);
    if (variant != BFLY_VARIANT_PLAIN && variant != BFLY_VARIANT_STREAM && variant != BFLY_VARIANT_STREAM_LT)
        return BFLY_ERR_SHAPE;
    return launched(bfly_kern::real(kind, in, out, n, batch, bins, stride, variant,
                                    (flags & BFLY_REAL_ADJOINT) != 0, table, S(stream)));
}

int bfly_r2c_ex(const float* in, float* out, uint32_t n, uint32_t batch, uint32_t bins,
                uint32_t out_stride, const float* table, void* stream) {
    return bfly_real_ex(BFLY_KIND_R2C, in, out, n, batch, bins, out_stride, BFLY_VARIANT_AUTO, 0, table, stream);
}

int bfly_r2c(const float* in, float* out, uint32_t n, uint32_t batch, const float* table,
             void* stream) {
    return bfly_r2c_ex(in, out, n, batch, n / 2 + 1, n + 2, table, stream);
}

int bfly_c2r_ex(const float* in, float* out, uint32_t n, uint32_t batch, uint32_t bins,
                uint32_t in_stride, const float* table, void* stream) {
    return bfly_real_ex(BFLY_KIND_C2R, in, out, n, batch, bins, in_stride, BFLY_VARIANT_AUTO, 0, table, stream);
}

int bfly_c2r(const float* in, float* out, uint32_t n, uint32_t batch, const float* table,
             void* stream) {
    return bfly_c2r_ex(in, out, n, batch, n / 2 + 1, n + 2, table, stream);
}

int bfly_host_pipelined(int kind, int sign, int variant, const float* host_in, float* host_out, uint32_t n,
                        uint32_t batch, const float* table, void* dev_in, void* dev_out, void* work,
                        uint32_t chunk_rows, void* stream) {
    bfly_host::PipeShape sh;
    if (int g = bfly_host::check_pipelined(kind, sign, variant, host_in, host_out, n, batch, table, dev_in, dev_out,
                                           work, chunk_rows, &sh))
        return g;
    sign = sh.sign;
    variant = sh.variant;
    const size_t row_in = sh.row_in, row_out = sh.row_out;
    const uint64_t cr = sh.chunk;
    const hipStream_t up = S(stream);
    // Pageable host memory goes through the runtime's own staging buffers,
    // which serialise both directions: chunks measured slower there (4096 x
    // 4096: 13.3-14.8 ms against 12.4 ms in turn), so only pinned buffers
    // are pipelined.
    auto pinned = [](const void* p) {
        hipPointerAttribute_t a;
        if (hipPointerGetAttributes(&a, p) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        return a.type == hipMemoryTypeHost;
    };
    if (cr >= batch || !pinned(host_in) || !pinned(host_out)) {
        // one chunk, or pageable memory: the three steps run in turn on the
        // caller's stream, for the whole batch at once unless the kernel needs
        // the workspace, which holds one chunk
        const uint64_t step =
            bfly_host::needs_work(kind, n, batch, variant) ? cr : (uint64_t)batch;
        int rc = BFLY_OK;
        for (uint64_t r0 = 0; r0 < batch && rc == BFLY_OK; r0 += step) {
            const uint32_t rows = (uint32_t)(batch - r0 < step ? batch - r0 : step);
            const float* din = static_cast<const float*>(dev_in) + r0 * row_in;
            float* dout = static_cast<float*>(dev_out) + r0 * row_out;
            if (hipMemcpyAsync(const_cast<float*>(din), host_in + r0 * row_in, rows * row_in * sizeof(float),
                               hipMemcpyHostToDevice, up) != hipSuccess) {
                rc = BFLY_ERR_H2D;
                break;
            }
            const hipError_t e =
                kind == BFLY_KIND_C2C
                    ? bfly_kern::c2c(din, dout, n, rows, sign, variant, table, work, up, nullptr, 1.0f)
                    : bfly_kern::real(kind, din, dout, n, rows, n / 2 + 1, n + 2, variant, false, table, up);
            if (e != hipSuccess) {
                rc = BFLY_ERR_LAUNCH;
                break;
            }
            if (hipMemcpyAsync(host_out + r0 * row_out, dout, rows * row_out * sizeof(float), hipMemcpyDeviceToHost,
                               up) != hipSuccess)
                rc = BFLY_ERR_D2H;
        }
        const bool synced = hipStreamSynchronize(up) == hipSuccess;
        if (rc != BFLY_OK) (void)hipGetLastError();
        return rc == BFLY_OK && !synced ? BFLY_ERR_SYNC : rc;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kPipeDevices) return BFLY_ERR_NOT_LOADED;
    std::lock_guard<std::mutex> lock(g_pipe_lock[dev]);
    PipeStreams& P = g_pipe[dev];
    if (int e = pipe_ready(P)) return e;
    int rc = BFLY_OK;
    uint64_t k = 0;
    for (uint64_t r0 = 0; r0 < batch; r0 += cr, k++) {
        const uint32_t rows = (uint32_t)(batch - r0 < cr ? batch - r0 : cr);
        const float* hin = host_in + r0 * row_in;
        float* hout = host_out + r0 * row_out;
        const float* din = static_cast<const float*>(dev_in) + r0 * row_in;
        float* dout = static_cast<float*>(dev_out) + r0 * row_out;
        hipEvent_t loaded = P.ev[2 * (k % kPipeRing)], done = P.ev[2 * (k % kPipeRing) + 1];
        if (hipMemcpyAsync(const_cast<float*>(din), hin, rows * row_in * sizeof(float), hipMemcpyHostToDevice, up) !=
            hipSuccess) {
            rc = BFLY_ERR_H2D;
            break;
        }
        if (hipEventRecord(loaded, up) != hipSuccess || hipStreamWaitEvent(P.comp, loaded, 0) != hipSuccess) {
            rc = BFLY_ERR_LAUNCH;
            break;
        }
        const hipError_t e =
            kind == BFLY_KIND_C2C
                ? bfly_kern::c2c(din, dout, n, rows, sign, variant, table, work, P.comp, nullptr, 1.0f)
                : bfly_kern::real(kind, din, dout, n, rows, n / 2 + 1, n + 2, variant, false, table, P.comp);
        if (e != hipSuccess || hipEventRecord(done, P.comp) != hipSuccess ||
            hipStreamWaitEvent(P.down, done, 0) != hipSuccess) {
            rc = BFLY_ERR_LAUNCH;
            break;
        }
        if (hipMemcpyAsync(hout, dout, rows * row_out * sizeof(float), hipMemcpyDeviceToHost, P.down) != hipSuccess) {
            rc = BFLY_ERR_D2H;
            break;
        }
    }
    // every issued step ends before the call returns, also after an error
    const bool synced = hipStreamSynchronize(up) == hipSuccess && hipStreamSynchronize(P.comp) == hipSuccess &&
                        hipStreamSynchronize(P.down) == hipSuccess;
    if (rc != BFLY_OK) (void)hipGetLastError();
    if (rc == BFLY_OK && !synced) rc = BFLY_ERR_SYNC;
    return rc;
}

int bfly_rt_time_transform(int kind, int sign, const float* in, float* out, uint32_t n,
                           uint32_t batch, const float* table, void* work, int reps,
                           void* stream, float* ms) {
    if (!ms || reps <= 0) return BFLY_ERR_SHAPE;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return BFLY_ERR_ALLOC;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return BFLY_ERR_ALLOC;
    }
    int rc = hipEventRecord(e0, S(stream)) == hipSuccess ? BFLY_OK : BFLY_ERR_LAUNCH;
    for (int i = 0; i < reps && rc == BFLY_OK; i++) {
        if (kind == BFLY_KIND_C2C) rc = bfly_c2c(in, out, n, batch, sign, table, work, stream);
        else if (kind == BFLY_KIND_R2C) rc = bfly_r2c(in, out, n, batch, table, stream);
        else if (kind == BFLY_KIND_C2R) rc = bfly_c2r(in, out, n, batch, table, stream);
        else rc = BFLY_ERR_SHAPE;
    }
    if (hipEventRecord(e1, S(stream)) != hipSuccess && rc == BFLY_OK) rc = BFLY_ERR_LAUNCH;
    if (hipEventSynchronize(e1) != hipSuccess && rc == BFLY_OK) rc = BFLY_ERR_SYNC;
    if (rc == BFLY_OK && hipEventElapsedTime(ms, e0, e1) != hipSuccess) rc = BFLY_ERR_SYNC;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int bfly_mode_mix(int op, const float* a, const float* b, float* c, uint32_t batch, uint32_t c_in,
                  uint32_t c_out, uint32_t modes, void* stream) {
    if (int e = bfly_host::check_mode_mix(op, a, b, c, batch, c_in, c_out, modes)) return e;
    return launched(bfly_kern::mode_mix(op, a, b, c, batch, c_in, c_out, modes, S(stream)));
}

int bfly_probe_copy(const void* src, void* dst, size_t bytes, void* stream) {
    if (!src || !dst || bytes == 0 || (bytes & 15) || ((uintptr_t)src & 15) || ((uintptr_t)dst & 15))
        return BFLY_ERR_SHAPE;
    return launched(bfly_kern::copy(src, dst, bytes, S(stream)));
}

int bfly_probe_device(int device, char* json, size_t cap) {
    if (!json || cap == 0) return BFLY_ERR_SHAPE;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return BFLY_ERR_NOT_LOADED;
    // an attribute the runtime cannot report stays 0 in the JSON
    int lds_per_cu = 0, mem_clock_khz = 0, bus_width = 0, l2 = 0;
    if (hipDeviceGetAttribute(&lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess)
        lds_per_cu = 0;
    if (hipDeviceGetAttribute(&mem_clock_khz, hipDeviceAttributeMemoryClockRate, device) != hipSuccess)
        mem_clock_khz = 0;
    if (hipDeviceGetAttribute(&bus_width, hipDeviceAttributeMemoryBusWidth, device) != hipSuccess)
        bus_width = 0;
    if (hipDeviceGetAttribute(&l2, hipDeviceAttributeL2CacheSize, device) != hipSuccess) l2 = 0;
    int wall_khz = 0;
    if (hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess)
        wall_khz = 0;
    const int w = std::snprintf(
        json, cap,
        "{\"name\": \"%s\", \"gcn_arch\": \"%s\", \"compute_units\": %d, \"wave_size\": %d, "
        "\"lds_bytes_per_workgroup\": %zu, \"lds_bytes_per_cu\": %d, \"l2_bytes\": %d, "
        "\"global_mem_total_bytes\": %zu, \"max_threads_per_block\": %d, "
        "\"clock_khz\": %d, \"mem_clock_khz\": %d, \"mem_bus_width_bits\": %d, "
        "\"wall_clock_khz\": %d}",
        p.name, p.gcnArchName, p.multiProcessorCount, p.warpSize, (size_t)p.sharedMemPerBlock,
        lds_per_cu, l2, (size_t)p.totalGlobalMem, p.maxThreadsPerBlock, p.clockRate,
        mem_clock_khz, bus_width, wall_khz);
    return (w > 0 && (size_t)w < cap) ? BFLY_OK : BFLY_ERR_SHAPE;
}

}  // extern "C"
