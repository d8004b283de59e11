// ref_shim.cpp -- flat extern "C" view of the reference's own host code, built
// from the reference sources where they lie (see oracle/Makefile) into
// oracle/_ref/libbfly_ref.so.  Test infrastructure only: tests/ load it to
// check the oracle restatement (oracle/bfly_oracle.c) and the product's
// descriptor functions against the code they restate.  It is never linked
// into the shipped library.
//
// Compiled with plain g++ together with the reference's
// src/framework/reference.cpp and its header-only rules (fft_k.hpp,
// long_fft_ub.h, descriptors.hpp), against the reference's include/ tree
// only: none of these needs the AscendCL runtime, so no stand-in header is
// involved.  src/framework/butterfly.cpp (Generator tables, cost model) does
// need that runtime and is not built; its tables are pinned by closed forms
// and by the transforms they drive (tests/test_oracle.py).
#include "butterfly/descriptors.hpp"
#include "butterfly/fft_k.hpp"
#include "butterfly/long_fft_ub.h"
#include "butterfly/reference.hpp"

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#define REF_API extern "C" __attribute__((visibility("default")))

REF_API void ref_fft_f32(const float* in, float* out, uint32_t n) { bfly::refFftF32(in, out, n); }
REF_API void ref_r2c_f32(const float* in, float* out, uint32_t n) { bfly::refR2CF32(in, out, n); }
REF_API void ref_c2r_f32(const float* in, float* out, uint32_t n) { bfly::refC2RF32(in, out, n); }
REF_API double ref_max_rel_scaled(const float* got, const float* ref, size_t len) {
    return bfly::maxRelScaled(got, ref, len);
}
REF_API float ref_pattern_at(size_t i) { return bfly::patternAt(i); }
REF_API uint32_t ref_sample_batches(uint32_t batch, uint32_t want, uint32_t* out, uint32_t maxOut) {
    return bfly::sampleBatches(batch, want, out, maxOut);
}
REF_API uint32_t ref_plane_k_for(uint32_t n) { return bfly::planeKFor(n); }
REF_API uint32_t ref_fold_d_for(uint32_t n, uint32_t batch, uint32_t nblk) {
    return bfly::foldDFor(n, batch, nblk);
}
REF_API uint64_t ref_row_fft_ub_bytes(uint64_t n, uint64_t d, uint64_t k) {
    return AB_ROW_FFT_UB_BYTES(n, d, k);
}
REF_API int ref_resolve_row_fft(uint32_t len, uint32_t rows, uint32_t d, uint32_t k,
                                uint32_t* foldD, uint32_t* planeK) {
    const auto p = butterfly::resolve_row_fft(len, rows, d, k);
    *foldD = p.fold_d;
    *planeK = p.plane_k;
    return p.legal ? 1 : 0;
}

// query_lowering (include/butterfly/descriptors.hpp:409-667) over a flat
// mapping.  The AB_LT_* environment knobs the reference reads are passed as
// strings (nullptr = unset) and applied around the call.
struct RefMapping {
    uint32_t n, batch;
    int precision;          // 0 fp32, 1 fp16
    int direction;          // 0 c2c, 1 r2c, 2 c2r
    uint32_t n1, n2;        // stage lengths (stage_partition = 2 when n2 != 0)
    int stage_partition;    // overrides the count when > 0
    int stage_space, stage_time, data_space, data_time;
    int role0, role1;
    int residence;          // 0 none, 1 cell, 2 block, 3 worker
    int boundary_layout;    // 0 interleaved, 1 twiddled, 2 natural, 3 half-spectrum
    int boundary_home;      // 0 host, 1 device GM, 2 on-chip
    int boundary_impl;      // 0 separate, 1 fused
    int pipeline_buffers;
    uint32_t fold_d, plane_k;
    uint64_t unit_ub_bytes; // 0 = the unit's default
};

struct RefLowering {
    int supported, abstract_feasible, launches, gm_boundaries, host_assisted, on_chip;
    uint64_t plan_ub_bytes, modeled_payload_gm;
    char kinds[256];
    char reason[512];
};

static void setOrUnset(const char* key, const char* value) {
    if (value) setenv(key, value, 1); else unsetenv(key);
}

REF_API void ref_default_mapping(uint32_t n1, uint32_t n2, RefMapping* m) {
    const auto& hw = butterfly::builtin_hardware_profile();
    const auto d = butterfly::default_long_mapping(n1, n2, hw);
    std::memset(m, 0, sizeof(*m));
    m->n1 = n1; m->n2 = n2;
    m->stage_partition = d.stage_partition;
    m->stage_space = d.stage_space; m->stage_time = d.stage_time;
    m->data_space = d.data_space; m->data_time = d.data_time;
    m->role0 = d.role_stages[0]; m->role1 = d.role_stages[1];
    m->residence = (int)d.residence;
    m->boundary_layout = (int)d.boundary_layout;
    m->boundary_home = (int)d.boundary_home;
    m->boundary_impl = (int)d.boundary_impl;
    m->pipeline_buffers = d.pipeline_buffers;
}

REF_API void ref_query_lowering(const RefMapping* in, const char* ltTile, const char* ltStripeK,
                                const char* ltIdx, RefLowering* out) {
    using namespace butterfly;
    const auto& hw = builtin_hardware_profile();
    ArchitectureMapping m;
    m.stage_lengths = {in->n1, in->n2};
    m.stage_partition = in->stage_partition;
    m.stage_space = in->stage_space; m.stage_time = in->stage_time;
    m.data_space = in->data_space; m.data_time = in->data_time;
    m.role_stages = {in->role0, in->role1};
    m.residence = (Residence)in->residence;
    m.boundary_layout = (Layout)in->boundary_layout;
    m.boundary_home = (BoundaryHome)in->boundary_home;
    m.boundary_impl = (BoundaryImpl)in->boundary_impl;
    m.pipeline_buffers = in->pipeline_buffers;
    m.row_fft_fold_d = in->fold_d;
    m.row_fft_plane_k = in->plane_k;
    TransformSpec spec{in->n, in->batch, in->precision ? Precision::FP16 : Precision::FP32,
                       (Direction)in->direction};
    auto unit = local_fft_capability();
    if (in->unit_ub_bytes) unit.ub_bytes = (size_t)in->unit_ub_bytes;
    setOrUnset("AB_LT_TILE", ltTile);
    setOrUnset("AB_LT_STRIPE_K", ltStripeK);
    setOrUnset("AB_LT_IDX", ltIdx);
    const LoweringResult r = query_lowering(spec, m, unit, hw);
    unsetenv("AB_LT_TILE"); unsetenv("AB_LT_STRIPE_K"); unsetenv("AB_LT_IDX");
    std::memset(out, 0, sizeof(*out));
    out->supported = r.supported;
    out->abstract_feasible = r.abstract_feasible;
    out->launches = r.visible_launches;
    out->gm_boundaries = r.materialized_gm_boundaries;
    out->host_assisted = r.host_assisted;
    out->on_chip = r.whole_transform_on_chip;
    out->plan_ub_bytes = r.plan_ub_bytes;
    out->modeled_payload_gm = r.modeled_payload_gm_rw_bytes;
    std::string kinds;
    for (size_t i = 0; i < r.launch_manifest.size(); i++) {
        if (i) kinds += "|";
        kinds += launch_kind_name(r.launch_manifest[i].kind);
    }
    std::strncpy(out->kinds, kinds.c_str(), sizeof(out->kinds) - 1);
    std::strncpy(out->reason, r.reason.c_str(), sizeof(out->reason) - 1);
}
