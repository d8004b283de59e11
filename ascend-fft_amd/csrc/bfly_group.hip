// bfly_group.hip -- grouped launch: many uneven transform requests run by one
// persistent kernel driven by a device-resident request table.
//
// Every request is cut into work units of WG*16 points: B = WG/T transforms
// of one request (T = n'/16 threads each, n' = n for C2C and n/2 for the real
// kinds), so every unit moves the same bytes however long its transforms
// are.  The host lists the units once (request, group index) in the order the
// configuration asks for; a launch walks that list either statically
// (workgroup w takes units w, w+grid, ...) or through an atomic counter.
// A unit runs the same device function the one-shot kernels run
// (c2c_unit / r2c_unit / c2r_unit), so with radix schedule 0 a grouped
// transform computes the same bits as its per-request launch.  With LDS
// twiddles (LT) a unit first stages its plan's two-level tables in LDS and
// then runs the per-transform code of the STREAM_LT kernels instead.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <new>
#include <utility>
#include <vector>

#include "bfly_hip.h"
#include "bfly_plans.hpp"
#include "bfly_tables.hpp"
#include "bfly_units.hpp"

using namespace bfly_dev;

namespace {

// One request as the kernel reads it (48 bytes).
struct GroupDesc {
    const void* in;
    void* out;
    const float2* tw;  // inter-pass table of the request's plan and schedule
    const float2* wn;  // real kinds: W_n^k, k = 0..n/2; C2C with LDS twiddles:
                       // [W_n^j, j < lo | W_n^{lo j}, j < n/lo], lo = min(64, n)
    uint32_t code;     // kind << 8 | inverse << 7 | plan index
    uint32_t batch;
    uint32_t pad[2];
};
static_assert(sizeof(GroupDesc) == 48, "request table entry");

constexpr uint32_t code_of(int kind, int dir, int plan) {
    return ((uint32_t)kind << 8) | (dir > 0 ? 0x80u : 0u) | (uint32_t)plan;
}

// LDS of one workgroup: WG*16 points with the 1-in-16 padding of lds_len.
template <int WG>
constexpr int group_lds_len() { return WG * bfly_plan::kPoints + WG; }

// LDS twiddle entries of the largest plan (C2cLdsTw / real_lds_tw_len).
constexpr int group_tw_len() {
    int len = 1;
    for (int i = 0; i < bfly_plan::kNumPlans; i++) {
        const int n = bfly_plan::kPlans[i].n, lo = n < TW2_LO ? n : TW2_LO;
        const int r1 = bfly_plan::kPlans[i].rad[bfly_plan::kPlans[i].nrad > 1 ? 1 : 0];
        const int c2c = lo + n / lo + bfly_plan::kPoints * r1;
        const int real = n >= 128 ? 64 + n / 64 + bfly_plan::kPoints * r1 + 64 + n / 64 + 1 : 0;
        len = c2c > len ? c2c : len;
        len = real > len ? real : len;
    }
    return len;
}

// Runs unit g of request d if d is of plan I; returns whether it matched.
// LT: twiddles from tables staged in `ltw` (group_tw_len entries).
template <int WG, int SC, bool LT, int I>
__device__ __forceinline__ bool try_plan(const GroupDesc& d, uint32_t g, float2* lds, float2* ltw, int tid) {
    constexpr int T = PlanOf<I>::T;
    if constexpr (LT && T <= WG) {
        static_assert(C2cLdsTw<I>::LEN <= group_tw_len(), "LDS twiddle area");
        if (d.code == code_of(BFLY_KIND_C2C, -1, I) || d.code == code_of(BFLY_KIND_C2C, 1, I)) {
            c2c_lds_tw<I, WG>(ltw, d.tw, d.wn, tid);
            __syncthreads();
            if (d.code == code_of(BFLY_KIND_C2C, -1, I))
                c2c_unit_lt<I, -1, WG>(static_cast<const float2*>(d.in), static_cast<float2*>(d.out), ltw, d.batch,
                                       g, lds, tid);
            else
                c2c_unit_lt<I, 1, WG>(static_cast<const float2*>(d.in), static_cast<float2*>(d.out), ltw, d.batch,
                                      g, lds, tid);
            return true;
        }
        // real kinds with n/2 >= 128 (the real STREAM_LT kernels' range)
        if constexpr (PlanOf<I>::N >= 128 && PlanOf<I>::N <= bfly_plan::kMaxLocalN) {
            constexpr uint32_t M = PlanOf<I>::N;
            static_assert(real_lds_tw_len<I>() <= group_tw_len(), "LDS twiddle area");
            const bool r2c = d.code == code_of(BFLY_KIND_R2C, -1, I);
            if (r2c || d.code == code_of(BFLY_KIND_C2R, 1, I)) {
                const RealTwLds<I> tws = real_lds_tw<I, WG>(ltw, d.tw, d.wn, tid);
                __syncthreads();
                if (r2c) {
                    float2 v[bfly_plan::kPoints];
                    r2c_load<I, WG>(v, static_cast<const float*>(d.in), d.batch, g, tid);
                    r2c_finish<I, WG, 0>(v, static_cast<float*>(d.out), tws, d.batch, M + 1, 2 * M + 2, g, lds,
                                         tid);
                } else {
                    float2 q[bfly_plan::kPoints + 1];
                    c2r_load<I, WG>(q, static_cast<const float*>(d.in), d.batch, M + 1, 2 * M + 2, g, tid);
                    c2r_finish<I, WG, 0>(q, static_cast<float*>(d.out), tws, d.batch, 1.0f / (float)(2 * M), g,
                                         lds, tid);
                }
                return true;
            }
        }
    }
    if constexpr (T <= WG) {
        if (d.code == code_of(BFLY_KIND_C2C, -1, I)) {
            c2c_unit<I, -1, LX_PAD, WG, SC>(static_cast<const float2*>(d.in), static_cast<float2*>(d.out),
                                            d.tw, d.batch, g, lds, tid);
            return true;
        }
        if (d.code == code_of(BFLY_KIND_C2C, 1, I)) {
            c2c_unit<I, 1, LX_PAD, WG, SC>(static_cast<const float2*>(d.in), static_cast<float2*>(d.out),
                                           d.tw, d.batch, g, lds, tid);
            return true;
        }
        // real kinds run the n/2-point plan; n/2 <= 4096
        if constexpr (PlanOf<I>::N >= 16 && PlanOf<I>::N <= bfly_plan::kMaxLocalN) {
            constexpr uint32_t M = PlanOf<I>::N;
            if (d.code == code_of(BFLY_KIND_R2C, -1, I)) {
                r2c_unit<I, WG, SC>(static_cast<const float*>(d.in), static_cast<float*>(d.out), d.tw, d.wn,
                                    d.batch, M + 1, 2 * M + 2, g, lds, tid);
                return true;
            }
            if (d.code == code_of(BFLY_KIND_C2R, 1, I)) {
                c2r_unit<I, WG, SC>(static_cast<const float*>(d.in), static_cast<float*>(d.out), d.tw, d.wn,
                                    d.batch, 1.0f / (float)(2 * M), M + 1, 2 * M + 2, g, lds, tid);
                return true;
            }
        }
    }
    return false;
}

template <int WG, int SC, bool LT, size_t... I>
__device__ __forceinline__ void run_unit(const GroupDesc& d, uint32_t g, float2* lds, float2* ltw, int tid,
                                         std::index_sequence<I...>) {
    (void)(try_plan<WG, SC, LT, (int)I>(d, g, lds, ltw, tid) || ...);
}

// ctr (dynamic schedule): ctr[0] = next unit, ctr[1] = workgroups finished.
// The last workgroup to finish puts both back to zero, so every launch (and
// every replay of a captured launch) starts from a cleared counter without a
// memset node.
template <int WG, int SC, bool LT>
__device__ __forceinline__ void group_body(const GroupDesc* __restrict__ descs, const uint2* __restrict__ units,
                                           uint32_t nunits, uint32_t* __restrict__ ctr,
                                           unsigned long long* __restrict__ stamps) {
    __shared__ float2 lds[group_lds_len<WG>()];
    __shared__ float2 ltw[LT ? group_tw_len() : 1];
    __shared__ uint32_t s_next;
    const bool dynamic = ctr != nullptr;
    if (stamps && threadIdx.x == 0) atomicMin(&stamps[0], (unsigned long long)wall_clock64());
    uint32_t u = blockIdx.x;
    if (dynamic) {
        if (threadIdx.x == 0) s_next = atomicAdd(&ctr[0], 1u);
        __syncthreads();
        u = __builtin_amdgcn_readfirstlane(s_next);
    }
    while (u < nunits) {
        const uint2 w = units[u];
        const GroupDesc d = descs[w.x];
        // Every address a unit derives from the thread index is loop
        // invariant; an opaque copy of the index per unit keeps the compiler
        // from hoisting those of all the unit bodies out of the loop at once
        // (their live ranges would overlap and spill).
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        run_unit<WG, SC, LT>(d, w.y, lds, ltw, tid, std::make_index_sequence<bfly_plan::kNumPlans>{});
        if (stamps) {
            // every store of this unit acknowledged, then one stamp per unit
            // (a wait on this thread's memory counters, not a device-wide
            // fence: that would write the L2 back and distort the timing)
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (threadIdx.x == 0) atomicMax(&stamps[1 + w.x], (unsigned long long)wall_clock64());
        }
        // the next unit reuses the LDS regions (and the twiddle tables)
        __syncthreads();
        if (dynamic) {
            if (threadIdx.x == 0) s_next = atomicAdd(&ctr[0], 1u);
            __syncthreads();
            u = __builtin_amdgcn_readfirstlane(s_next);
        } else {
            u += gridDim.x;
        }
    }
    // A workgroup counts itself out only after its last fetch from ctr[0]
    // returned, so the last one out knows every fetch is done.
    if (dynamic && threadIdx.x == 0) {
        if (atomicAdd(&ctr[1], 1u) == gridDim.x - 1) {
            atomicExch(&ctr[0], 0u);
            atomicExch(&ctr[1], 0u);
        }
    }
}

template <int WG, int SC>
__global__ __launch_bounds__(WG) void k_group(const GroupDesc* __restrict__ descs, const uint2* __restrict__ units,
                                             uint32_t nunits, uint32_t* __restrict__ ctr,
                                             unsigned long long* __restrict__ stamps) {
    group_body<WG, SC, false>(descs, units, nunits, ctr, stamps);
}

// LDS twiddles: at most 128 VGPRs, so two 512-thread workgroups fit a CU
// (their LDS, 2 x 75 KiB, does too)
template <int WG>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(4))) void k_group_lt(
    const GroupDesc* __restrict__ descs, const uint2* __restrict__ units, uint32_t nunits,
    uint32_t* __restrict__ ctr, unsigned long long* __restrict__ stamps) {
    group_body<WG, 0, true>(descs, units, nunits, ctr, stamps);
}

__global__ void k_stamp_init(unsigned long long* __restrict__ stamps, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) stamps[i] = i == 0 ? ~0ull : 0ull;
}

using GroupKernel = void (*)(const GroupDesc*, const uint2*, uint32_t, uint32_t*, unsigned long long*);

// LDS twiddles run schedule 0 only (bfly_host::group_plan enforces it).
GroupKernel group_kernel(int wg, int sched, bool lt) {
    if (wg == 512) return lt ? &k_group_lt<512> : sched ? &k_group<512, 1> : &k_group<512, 0>;
    if (wg == 1024) return lt ? &k_group_lt<1024> : sched ? &k_group<1024, 1> : &k_group<1024, 0>;
    return nullptr;
}

using bfly_host::group_local_len;

struct Group {
    void* dev = nullptr;
    size_t bytes = 0;
    const GroupDesc* descs = nullptr;
    const uint2* units = nullptr;
    uint32_t* ctr = nullptr;
    uint32_t nreq = 0, nunits = 0, grid = 0;
    int wg = 512, sched = 0, dynamic = 0;
    bool lt = false;
};

// Entries of [W_n^j, j < lo | W_n^{lo j}, j < n/lo], lo = min(64, n).
uint32_t two_level_len(uint32_t n) {
    const uint32_t lo = n < (uint32_t)TW2_LO ? n : (uint32_t)TW2_LO;
    return lo + n / lo;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

int bfly_group_supported(int kind, uint32_t n, int workgroup) {
    return bfly_host::group_ok(kind, n, workgroup) ? 1 : 0;
}

int bfly_group_create(const bfly_group_request* reqs, int nreq, const bfly_group_config* cfg,
                      void** group) {
    if (!group) return BFLY_ERR_SHAPE;
    *group = nullptr;
    bfly_host::GroupPlan P;
    if (int e = bfly_host::group_plan(reqs, nreq, cfg, &P)) return e;
    const bfly_group_config& c = P.cfg;
    const int wg = c.workgroup;
    const bool lt = c.twiddles == 1;
    const uint64_t total = P.total;

    std::map<std::pair<int, uint32_t>, size_t> tab_off;  // (real?, n) -> complex offset
    size_t tab_len = 0;
    for (int r = 0; r < nreq; r++) {
        const bool real = reqs[r].kind != BFLY_KIND_C2C;
        const auto key = std::make_pair((int)real, reqs[r].n);
        if (tab_off.count(key)) continue;
        tab_off[key] = tab_len;
        const int plan = bfly_plan::plan_index(group_local_len(reqs[r].kind, reqs[r].n));
        size_t len = bfly_plan::stage_twiddle_len(plan, c.schedule);
        if (real) len += reqs[r].n / 2 + 1;
        else if (lt) len += two_level_len(reqs[r].n);
        tab_len += (len + 1) & ~(size_t)1;  // keep every table 16-byte aligned
    }

    const size_t off_desc = 0;
    const size_t off_units = align256(off_desc + sizeof(GroupDesc) * (size_t)nreq);
    const size_t off_ctr = align256(off_units + sizeof(uint2) * (size_t)total);
    const size_t off_tab = align256(off_ctr + 2 * sizeof(uint32_t));
    const size_t bytes = align256(off_tab + tab_len * sizeof(float2));

    Group* G = new (std::nothrow) Group();
    if (!G) return BFLY_ERR_ALLOC;
    if (hipMalloc(&G->dev, bytes) != hipSuccess) {
        (void)hipGetLastError();  // a refused allocation is reported here, not by a later launch
        delete G;
        return BFLY_ERR_ALLOC;
    }
    std::vector<unsigned char> host(bytes, 0);
    unsigned char* base = static_cast<unsigned char*>(G->dev);
    float* tabs = reinterpret_cast<float*>(host.data() + off_tab);
    for (const auto& kv : tab_off) {
        const bool real = kv.first.first != 0;
        const uint32_t n = kv.first.second;
        const int plan = bfly_plan::plan_index(real ? n / 2 : n);
        float* dst = tabs + 2 * kv.second;
        const size_t st = bfly_host::fill_stage(plan, dst, c.schedule);
        if (real) {
            for (uint32_t k = 0; k <= n / 2; k++) bfly_host::twiddle(k, n, dst + 2 * (st + k));
        } else if (lt) {
            // the two-level factors, each rounded once from FP64 (the values of
            // the per-request tables)
            const uint32_t lo = n < (uint32_t)TW2_LO ? n : (uint32_t)TW2_LO;
            for (uint32_t j = 0; j < lo; j++) bfly_host::twiddle(j, n, dst + 2 * (st + j));
            for (uint32_t j = 0; j < n / lo; j++) bfly_host::twiddle((uint64_t)lo * j, n, dst + 2 * (st + lo + j));
        }
    }
    GroupDesc* d = reinterpret_cast<GroupDesc*>(host.data() + off_desc);
    for (int r = 0; r < nreq; r++) {
        const bfly_group_request& q = reqs[r];
        const bool real = q.kind != BFLY_KIND_C2C;
        const int plan = bfly_plan::plan_index(group_local_len(q.kind, q.n));
        const size_t to = tab_off[std::make_pair((int)real, q.n)];
        const float2* tw = reinterpret_cast<const float2*>(base + off_tab) + to;
        d[r].in = q.in;
        d[r].out = q.out;
        d[r].tw = tw;
        d[r].wn = real || lt ? tw + bfly_plan::stage_twiddle_len(plan, c.schedule) : nullptr;
        const int dir = q.kind == BFLY_KIND_C2C ? q.sign : (q.kind == BFLY_KIND_R2C ? -1 : 1);
        d[r].code = code_of(q.kind, dir, plan);
        d[r].batch = q.batch;
    }
    uint2* ul = reinterpret_cast<uint2*>(host.data() + off_units);
    for (uint64_t u = 0; u < total; u++) ul[u] = make_uint2(P.unit_req[u], P.unit_grp[u]);
    if (hipMemcpy(G->dev, host.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(G->dev);
        delete G;
        return BFLY_ERR_H2D;
    }

    int dev = 0, cus = 0, occ = 0;
    GroupKernel k = group_kernel(wg, c.schedule, lt);
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(k), wg, 0) !=
            hipSuccess ||
        occ <= 0) {
        (void)hipFree(G->dev);
        delete G;
        return BFLY_ERR_LAUNCH;
    }
    uint64_t grid = c.grid > 0 ? (uint64_t)c.grid : (uint64_t)cus * (uint64_t)occ;
    if (grid > total) grid = total;

    G->bytes = bytes;
    G->descs = reinterpret_cast<const GroupDesc*>(base + off_desc);
    G->units = reinterpret_cast<const uint2*>(base + off_units);
    G->ctr = reinterpret_cast<uint32_t*>(base + off_ctr);
    G->nreq = (uint32_t)nreq;
    G->nunits = (uint32_t)total;
    G->grid = (uint32_t)grid;
    G->wg = wg;
    G->sched = c.schedule;
    G->dynamic = c.dynamic;
    G->lt = lt;
    *group = G;
    return BFLY_OK;
}

int bfly_group_launch(void* group, unsigned long long* stamps, void* stream) {
    Group* G = static_cast<Group*>(group);
    if (!G || !G->dev) return BFLY_ERR_SHAPE;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (stamps) {
        hipLaunchKernelGGL(k_stamp_init, dim3(1), dim3(256), 0, s, stamps, G->nreq + 1);
        if (hipGetLastError() != hipSuccess) return BFLY_ERR_LAUNCH;
    }
    GroupKernel k = group_kernel(G->wg, G->sched, G->lt);
    hipLaunchKernelGGL(k, dim3(G->grid), dim3(G->wg), 0, s, G->descs, G->units, G->nunits,
                       G->dynamic ? G->ctr : nullptr, stamps);
    return hipGetLastError() == hipSuccess ? BFLY_OK : BFLY_ERR_LAUNCH;
}

int bfly_group_query(void* group, uint32_t* units, uint32_t* grid, size_t* bytes) {
    const Group* G = static_cast<const Group*>(group);
    if (!G) return BFLY_ERR_SHAPE;
    if (units) *units = G->nunits;
    if (grid) *grid = G->grid;
    if (bytes) *bytes = G->bytes;
    return BFLY_OK;
}

int bfly_group_destroy(void* group) {
    Group* G = static_cast<Group*>(group);
    if (!G) return BFLY_OK;
    const hipError_t e = G->dev ? hipFree(G->dev) : hipSuccess;
    delete G;
    return e == hipSuccess ? BFLY_OK : BFLY_ERR_SYNC;
}

}  // extern "C"
