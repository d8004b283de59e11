// bfly_noalloc.hip -- bfly_c2c_noalloc: the C2C entry point for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  Same kernels and results as
// bfly_c2c; the four-step chain's intermediate (the only scratch any C2C
// kernel uses) lives in the file-scope array `scratch` below instead of a
// caller-provided workspace.
//
// `scratch` is one copy per GPU, held until the process ends.  In host code
// its name is only a stand-in: the device address of each GPU's copy comes
// from hipGetSymbolAddress, looked up once per device (with that device
// current) and kept in `g_scratch`; only that address is ever passed to a
// kernel.  Every call on a GPU shares that GPU's copy, so calls that overlap
// on one GPU (different streams, or different host threads) race on it.
// Nothing in it needs to start at zero: the column kernel writes every
// element of the intermediate before the row kernel reads it.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>

#include "bfly_hip.h"
#include "bfly_kernels.hpp"
#include "bfly_tables.hpp"

namespace {

constexpr size_t kScratchBytes = BFLY_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "whole 256-byte blocks");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::atomic<void*> g_scratch[kMaxDevices] = {};
std::mutex g_lookup;

// Device address of the current GPU's scratch, looked up on first use.
int scratch_of_current_device(void** ptr) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return BFLY_ERR_NOT_LOADED;
    void* p = g_scratch[dev].load(std::memory_order_acquire);
    if (!p) {
        std::lock_guard<std::mutex> lock(g_lookup);
        p = g_scratch[dev].load(std::memory_order_relaxed);
        if (!p) {
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
                (void)hipGetLastError();
                return BFLY_ERR_NOT_LOADED;
            }
            g_scratch[dev].store(p, std::memory_order_release);
        }
    }
    *ptr = p;
    return BFLY_OK;
}

}  // namespace

extern "C" {

int bfly_noalloc_prepare(void) {
    void* p = nullptr;
    return scratch_of_current_device(&p);
}

int bfly_c2c_noalloc(const float* in, float* out, uint32_t n, uint32_t batch, int sign, const float* table,
                     void* stream) {
    const bool known = bfly_supported(BFLY_KIND_C2C, n) && batch;
    const int variant = known ? bfly_host::auto_variant(n, batch) : BFLY_VARIANT_AUTO;
    const bool chain = known && bfly_host::needs_work(BFLY_KIND_C2C, n, batch, variant);
    // the argument guards of bfly_c2c first (a non-null stand-in for the
    // workspace they would otherwise ask the caller for) ...
    void* const stand_in = reinterpret_cast<void*>(uintptr_t{256});
    if (int g = bfly_host::check_execute(BFLY_KIND_C2C, sign, in, out, n, batch, table, chain ? stand_in : nullptr,
                                         variant))
        return g;
    // ... then the scratch: the chain's intermediate is n * batch complex
    void* work = nullptr;
    if (chain) {
        if ((uint64_t)n * batch * 8ull > kScratchBytes) return BFLY_ERR_SCRATCH;
        if (int rc = scratch_of_current_device(&work)) return rc;
    }
    if (!bfly_host::disjoint(BFLY_KIND_C2C, in, out, work, n, batch, 0, 0)) return BFLY_ERR_SHAPE;
    const hipError_t e = bfly_kern::c2c(in, out, n, batch, sign, variant, table, work,
                                        reinterpret_cast<hipStream_t>(stream), nullptr, 1.0f);
    return e == hipSuccess ? BFLY_OK : BFLY_ERR_LAUNCH;
}

}  // extern "C"
