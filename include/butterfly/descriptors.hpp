// Long-FFT lowering descriptors: the reference's architecture-descriptor
// vocabulary (include/butterfly/descriptors.hpp of the reference: transform
// spec G, mapping A, processing unit P, layouts L, hardware H and the
// query_lowering contract Q) describing the lowerings this library executes
// on gfx950.
//
// query_lowering is a pure function: it allocates and launches nothing and
// never substitutes a unit or mapping.  Its first level (is the tuple legal
// and resource-feasible?) applies the reference's rules with the reference's
// reasons (descriptors.hpp:409-486).  Its second level (does the runtime hold
// an executable lowering?) answers for the gfx950 runtime, whose lowerings
// differ from the Ascend chains:
//   * one stage (stage_partition 1, block-resident, boundary on chip): the
//     whole transform in one workgroup's LDS, one launch, no boundary
//     (k_c2c_local for n <= 4096, k_c2c_single for 8192 / 16384);
//   * two stages (stage_partition 2, boundary in device memory, fused
//     boundary): two launches -- column FFTs of N1 with the twiddle and the
//     transposed store fused in, then row FFTs of N2 reading that boundary
//     with the natural-order store fused in (k_long_cols, k_long_rows).
// The reference's host-memory chain and its separate six-launch device chain
// are legal mappings with no gfx950 lowering ("abstract-feasible-but-not-
// lowered").  Every count is derived from the launch manifest; the modelled
// payload traffic is the reference's (two tensor passes per launch).  The
// resource figure keeps the reference's field name `plan_ub_bytes` but holds
// LDS bytes per workgroup of the gfx950 kernels.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "butterfly/fft_k.hpp"

namespace butterfly {

// Geometry of the gfx950 kernels (the plan table of the HIP sources): 16
// points per thread, 256-thread workgroups up to 4096 points, one workgroup
// per transform up to 16384, LDS padded by one slot per 16 complex values.
constexpr uint32_t kGfxPoints = 16;
constexpr uint32_t kGfxBlock = 256;
constexpr uint32_t kGfxMaxLocalN = 4096;
constexpr uint32_t kGfxMaxSingleN = 16384;
constexpr uint64_t gfx_lds_len(uint64_t n) { return n + n / 16; }

// ---- G: transform ----
enum class Precision { FP32, FP16 };
enum class Direction { C2C_FWD, R2C, C2R };

struct TransformSpec {
    uint32_t n = 0;
    uint32_t batch = 0;
    Precision precision = Precision::FP32;
    Direction direction = Direction::C2C_FWD;
};

// ---- L: layouts ----
enum class Layout { InterleavedComplex, TwiddledStage, NaturalOrder, HalfSpectrumReal };

inline const char* layout_name(Layout l) {
    switch (l) {
        case Layout::InterleavedComplex: return "interleaved_complex";
        case Layout::TwiddledStage: return "twiddled_stage";
        case Layout::NaturalOrder: return "natural_order";
        case Layout::HalfSpectrumReal: return "half_spectrum_real";
    }
    return "unknown";
}
inline const char* precision_name(Precision p) { return p == Precision::FP32 ? "fp32" : "fp16"; }

// ---- A: mapping ----
enum class Residence { None, CellResident, BlockResident, WorkerResident };
enum class BoundaryHome { HostMemory, DeviceGM, OnChip };
enum class BoundaryImpl { Separate, Fused };

inline const char* boundary_impl_name(BoundaryImpl i) {
    return i == BoundaryImpl::Fused ? "fused" : "separate";
}
inline const char* residence_name(Residence r) {
    switch (r) {
        case Residence::None: return "none";
        case Residence::CellResident: return "cell-resident";
        case Residence::BlockResident: return "block-resident";
        case Residence::WorkerResident: return "worker-resident";
    }
    return "unknown";
}

struct ArchitectureMapping {
    int stage_partition = 0;              // stage groups in the chain
    std::vector<uint32_t> stage_lengths;  // points per stage group
    int stage_space = 0;                  // Us
    int stage_time = 0;                   // Ts
    int data_space = 0;                   // Ud
    int data_time = 0;                    // Td
    std::vector<int> role_stages;         // role slot -> stage it serves
    Residence residence = Residence::None;
    Layout boundary_layout = Layout::TwiddledStage;
    BoundaryHome boundary_home = BoundaryHome::HostMemory;
    BoundaryImpl boundary_impl = BoundaryImpl::Separate;
    int pipeline_buffers = 1;
    uint32_t row_fft_fold_d = 0;   // Ascend row-kernel overrides (0 = derive),
    uint32_t row_fft_plane_k = 0;  // validated for descriptor parity only
};

// ---- P: processing unit ----
enum class SyncScope { None, AIVIntraCore, CoreIntraGroup, CoreInterGroup };

inline const char* sync_scope_name(SyncScope s) {
    switch (s) {
        case SyncScope::None: return "none";
        case SyncScope::AIVIntraCore: return "aiv_intra_core";
        case SyncScope::CoreIntraGroup: return "core_intra_group";
        case SyncScope::CoreInterGroup: return "core_inter_group";
    }
    return "unknown";
}

struct ProcessingUnitCapability {
    std::string core;
    Precision precision = Precision::FP32;
    int min_local_log_n = 0;
    int max_local_log_n = 0;
    size_t ub_bytes = 0;  // on-chip bytes the unit may claim (LDS on gfx950)
    SyncScope synchronization_scope = SyncScope::None;  // AIVIntraCore = workgroup barrier
    std::vector<Layout> supported_layouts;
    bool supports_multi_role = false;
};

// The 256-thread row kernel: 16 .. 4096 points LDS-resident per transform;
// the long chain's stage lengths must lie in its range.
inline ProcessingUnitCapability local_fft_capability() {
    ProcessingUnitCapability u;
    u.core = "bfly_c2c_local";
    u.precision = Precision::FP32;
    u.min_local_log_n = 4;
    u.max_local_log_n = 12;
    u.ub_bytes = 163840;
    u.synchronization_scope = SyncScope::AIVIntraCore;
    u.supported_layouts = {Layout::InterleavedComplex, Layout::TwiddledStage, Layout::NaturalOrder};
    u.supports_multi_role = false;
    return u;
}

// One workgroup per whole transform (up to 1024 threads, 16384 points): every
// stage runs in the same workgroup, the boundary stays in LDS.
inline ProcessingUnitCapability single_fft_capability() {
    ProcessingUnitCapability u = local_fft_capability();
    u.core = "bfly_c2c_single";
    u.max_local_log_n = 14;
    u.supports_multi_role = true;
    return u;
}

// ---- H: hardware ----
struct HardwareProfile {
    const char* profile_id = "";
    const char* soc = "";
    int aicore_num = 0;        // compute units
    int vector_core_num = 0;   // compute units (one vector core each)
    size_t l2_bytes = 0;
    size_t global_mem_total_bytes = 0;
    size_t ub_bytes_per_core = 0;  // LDS bytes per compute unit
    int vec_lane_fp32 = 0;         // wave size
};

// MI355X (ascend-fft_amd/config/gfx950_profile.json).
inline const HardwareProfile& builtin_hardware_profile() {
    static const HardwareProfile p = {
        "mi355x_gfx950", "gfx950",
        256,                 // compute units
        256,                 // vector cores = compute units
        4194304ull,          // L2 per XCD (the 256 MiB MALL sits behind it)
        309220868096ull,     // HBM3E the runtime reports
        163840ull,           // 160 KiB LDS per compute unit
        64                   // wave64
    };
    return p;
}

// ---- Q: lowering contract ----
enum class LaunchKind { RowFFT, TransposeIn, Twiddle, TransposeBoundary, TransposeOut };

inline const char* launch_kind_name(LaunchKind k) {
    switch (k) {
        case LaunchKind::RowFFT: return "row_fft";
        case LaunchKind::TransposeIn: return "transpose_in";
        case LaunchKind::Twiddle: return "twiddle";
        case LaunchKind::TransposeBoundary: return "transpose_boundary";
        case LaunchKind::TransposeOut: return "transpose_out";
    }
    return "unknown";
}

struct LaunchRecord {
    LaunchKind kind;
    int stage;          // stage group the launch computes
    int boundary_edge;  // 1 when the launch consumes a materialised boundary
};

enum class CoreKind { AIVVectorCore };

inline const char* core_kind_name(CoreKind) { return "compute_unit"; }

struct KernelResource {
    const char* kernel;
    CoreKind core;
    size_t ub_bytes;  // LDS bytes per workgroup
    SyncScope sync_scope;
    const char* constraint;
};

// The reference's row-kernel (D, K) resolution (descriptors.hpp:231-251),
// kept for descriptor parity: an override the Ascend kernel contract forbids
// is rejected with the same reason.  The gfx950 kernels do not use D or K.
struct RowFftPlan {
    uint32_t fold_d;
    uint32_t plane_k;
    bool legal = true;
    const char* reason = "";
};
inline RowFftPlan resolve_row_fft(uint32_t len, uint32_t launch_rows, uint32_t fold_d_override,
                                  uint32_t plane_k_override) {
    RowFftPlan p;
    p.fold_d = fold_d_override ? fold_d_override : bfly::foldDFor(len, launch_rows, 48u);
    if (p.fold_d < 1u) p.fold_d = 1u;
    p.plane_k = plane_k_override ? plane_k_override : bfly::planeKFor(len);
    const bool kSet = p.plane_k == 8u || p.plane_k == 16u || p.plane_k == 32u;
    if (!kSet || p.plane_k * 8u > len || len % p.plane_k != 0u) {
        p.legal = false;
        p.reason = "plane K violates the kernel contract {8,16,32} with K | len and len/K >= 8";
    }
    return p;
}

// LDS of the kernel that runs a stage of `len` points in a chain of
// `stages` stage groups.
inline KernelResource gfx_stage_resource(uint32_t len, int stages, int stage) {
    if (stages == 1) {
        if (len <= kGfxMaxLocalN) {
            const uint64_t per = kGfxBlock / (len / kGfxPoints);
            return KernelResource{"k_c2c_local", CoreKind::AIVVectorCore,
                                  (size_t)(per * gfx_lds_len(len) * 8), SyncScope::AIVIntraCore,
                                  "256 threads, 256/(n/16) transforms, padded LDS"};
        }
        return KernelResource{"k_c2c_single", CoreKind::AIVVectorCore, (size_t)(gfx_lds_len(len) * 8),
                              SyncScope::AIVIntraCore, "n/16 threads, one transform, padded LDS"};
    }
    // the column kernel runs 512-thread workgroups (256 when the batch leaves
    // fewer column groups than fit the GPU: the figure here is the larger),
    // the row kernel 256
    const uint64_t threads = stage == 0 ? 2ull * kGfxBlock : kGfxBlock;
    const uint64_t cols = threads / (len / kGfxPoints);
    return KernelResource{stage == 0 ? "k_long_cols" : "k_long_rows", CoreKind::AIVVectorCore,
                          (size_t)(cols * (gfx_lds_len(len) + 1) * 8), SyncScope::AIVIntraCore,
                          stage == 0 ? "512 threads, 512/(len/16) columns, padded LDS regions + 1"
                                     : "256 threads, 256/(len/16) rows, padded LDS regions + 1"};
}

// Two tensor passes (read + write) per launch, the reference's payload
// model (descriptors.hpp:347-356): no tables, no host transfers.
inline uint64_t modeled_payload_gm_rw(LaunchKind, uint64_t tensor_bytes) { return 2ull * tensor_bytes; }

// Serial launches: the plan's on-chip figure is the peak over the chain.
inline KernelResource plan_peak(const std::vector<KernelResource>& ks) {
    KernelResource best = ks.empty()
        ? KernelResource{"<empty>", CoreKind::AIVVectorCore, 0, SyncScope::None, "n/a"}
        : ks.front();
    for (const auto& k : ks)
        if (k.ub_bytes > best.ub_bytes) best = k;
    return best;
}

struct LoweringResult {
    bool abstract_feasible = false;
    bool supported = false;
    std::string reason;
    std::vector<LaunchRecord> launch_manifest;
    int visible_launches = 0;
    int materialized_gm_boundaries = 0;
    size_t plan_ub_bytes = 0;  // LDS bytes per workgroup, peak over the manifest
    uint64_t modeled_payload_gm_rw_bytes = 0;
    Residence resident_subgraph = Residence::None;
    bool whole_transform_on_chip = false;
    bool host_assisted = false;
};

inline bool selectable(const LoweringResult& r) { return r.supported; }

inline LoweringResult query_lowering(const TransformSpec& spec, const ArchitectureMapping& m,
                                     const ProcessingUnitCapability& unit, const HardwareProfile& hw) {
    LoweringResult r;
    auto reject = [&r](std::string why) {
        r.supported = false;
        r.reason = std::move(why);
        return r;
    };
    // ---- level 1: the reference's legality rules and reasons ----
    if (spec.precision != unit.precision)
        return reject("precision mismatch: transform is " + std::string(precision_name(spec.precision)) +
                      ", unit " + unit.core + " provides " + precision_name(unit.precision));
    if (spec.direction != Direction::C2C_FWD)
        return reject(std::string("direction not supported by mapping for unit ") + unit.core);
    if (spec.n < 64 || (spec.n & (spec.n - 1))) return reject("transform length must be a power of two >= 64");
    if (m.stage_partition <= 0 || m.stage_lengths.empty()) return reject("mapping declares no stage partition");
    if ((int)m.stage_lengths.size() != m.stage_partition)
        return reject("stage_partition inconsistent with stage lengths");
    uint64_t product = 1;
    for (uint32_t len : m.stage_lengths) product *= len;
    if (product != spec.n)
        return reject("stage lengths product " + std::to_string(product) +
                      " does not match transform length " + std::to_string(spec.n));
    for (uint32_t len : m.stage_lengths) {
        if (len < 64 || (len & (len - 1))) return reject("stage length must be a power of two >= 64");
        int lg = 0;
        while ((1u << lg) < len) lg++;
        if (lg < unit.min_local_log_n || lg > unit.max_local_log_n)
            return reject("stage length " + std::to_string(len) + " (log2 " + std::to_string(lg) +
                          ") outside unit " + unit.core + " range [" + std::to_string(unit.min_local_log_n) +
                          "," + std::to_string(unit.max_local_log_n) + "]");
    }
    if (m.stage_space < 1) return reject("Us (stage services) must be >= 1");
    if (m.stage_time < 1) return reject("Ts (stage reuse) must be >= 1");
    if (m.data_time < 1) return reject("Td (tile reuse) must be >= 1");
    if (m.data_space < 1) return reject("Ud (concurrent data tiles) must be >= 1");
    if (m.data_space > hw.vector_core_num)
        return reject("Ud " + std::to_string(m.data_space) + " exceeds AIV count " +
                      std::to_string(hw.vector_core_num));
    if (m.stage_space > hw.vector_core_num)
        return reject("Us " + std::to_string(m.stage_space) + " exceeds AIV count " +
                      std::to_string(hw.vector_core_num));
    if (m.pipeline_buffers < 1) return reject("pipeline_buffers must be >= 1");
    for (int stage : m.role_stages)
        if (stage < 0 || stage >= m.stage_partition)
            return reject("role allocation references stage " + std::to_string(stage) + " of " +
                          std::to_string(m.stage_partition));
    bool layout_ok = false;
    for (Layout l : unit.supported_layouts)
        if (l == m.boundary_layout) layout_ok = true;
    if (!layout_ok)
        return reject(std::string("boundary layout ") + layout_name(m.boundary_layout) +
                      " not supported by unit " + unit.core);
    if (m.residence == Residence::BlockResident && !unit.supports_multi_role)
        return reject("block residence requires a multi-role unit: " + unit.core);
    if (m.residence == Residence::WorkerResident && !unit.supports_multi_role)
        return reject("worker residence requires a multi-role unit: " + unit.core);
    if ((m.residence == Residence::BlockResident || m.residence == Residence::WorkerResident) &&
        unit.synchronization_scope == SyncScope::None)
        return reject(std::string("residence ") + residence_name(m.residence) +
                      " requires a synchronization scope on unit " + unit.core);
    if (m.boundary_home == BoundaryHome::OnChip && m.residence != Residence::BlockResident)
        return reject("on-chip boundary requires block residence under one synchronized owner");
    const uint64_t gm = 24ull * (uint64_t)spec.n * (uint64_t)spec.batch;
    if (gm > (40ull << 30)) return reject("gm footprint " + std::to_string(gm) + " bytes exceeds 40 GiB guard");
    if (gm > (uint64_t)hw.global_mem_total_bytes) return reject("gm footprint exceeds device global memory");
    // the reference's row-kernel override contract, checked for every stage
    for (uint32_t len : m.stage_lengths) {
        const uint32_t rows = (spec.n / len) * spec.batch;
        const RowFftPlan rf = resolve_row_fft(len, rows, m.row_fft_fold_d, m.row_fft_plane_k);
        if (!rf.legal) return reject(std::string("illegal row-FFT plan: ") + rf.reason);
    }
    r.abstract_feasible = true;

    // ---- level 2: the gfx950 runtime's lowerings ----
    std::string not_lowered;
    const std::vector<int> roles_two{0, 1}, roles_one{0};
    if (m.stage_partition == 1) {
        if (m.boundary_home != BoundaryHome::OnChip || m.residence != Residence::BlockResident)
            not_lowered = "a one-stage chain is lowered only block-resident with the boundary on chip";
        else if (spec.n > kGfxMaxSingleN)
            not_lowered = "n " + std::to_string(spec.n) + " exceeds one workgroup's LDS (16384 points)";
        else if (m.stage_space != 1 || m.stage_time != 1)
            not_lowered = "one-stage chains run one stage service once (Us = Ts = 1)";
        else if (m.data_time != 1)
            not_lowered = "Td " + std::to_string(m.data_time) + ": no tile-reuse lowering exists";
        else if (m.pipeline_buffers != 1)
            not_lowered = "pipeline_buffers " + std::to_string(m.pipeline_buffers) +
                          ": only single-buffer staging is lowered";
        else if (m.role_stages != roles_one)
            not_lowered = "role_stages must map the one stage";
    } else if (m.stage_partition == 2) {
        if (m.stage_space != 1)
            not_lowered = "Us " + std::to_string(m.stage_space) + ": only a single stage service is lowered";
        else if (m.stage_time != m.stage_partition)
            not_lowered = "Ts " + std::to_string(m.stage_time) +
                          ": only sequential reuse across the stage groups is lowered";
        else if (m.data_time != 1)
            not_lowered = "Td " + std::to_string(m.data_time) + ": no tile-reuse lowering exists";
        else if (m.pipeline_buffers != 1)
            not_lowered = "pipeline_buffers " + std::to_string(m.pipeline_buffers) +
                          ": only single-buffer staging is lowered";
        else if (m.role_stages != roles_two)
            not_lowered = "role_stages must map the two stages 1:1";
        else if (m.residence != Residence::None)
            not_lowered = std::string("residence ") + residence_name(m.residence) +
                          ": no resident-subgraph lowering exists yet";
        else if (m.boundary_layout != Layout::TwiddledStage)
            not_lowered = std::string("boundary layout ") + layout_name(m.boundary_layout) +
                          ": only the twiddled-stage boundary is lowered";
        else if (m.boundary_home != BoundaryHome::DeviceGM)
            not_lowered = "boundary home " +
                          std::string(m.boundary_home == BoundaryHome::HostMemory ? "host memory" : "on chip") +
                          ": the gfx950 chain keeps the stage boundary in device memory";
        else if (m.boundary_impl != BoundaryImpl::Fused)
            not_lowered = "boundary impl separate: the gfx950 chain fuses transposes and twiddle into "
                          "the two row-FFT launches";
    } else {
        not_lowered = "runtime implements one-stage and two-segment chains only (stage_partition 1 or 2)";
    }
    if (!not_lowered.empty()) {
        r.supported = false;
        r.reason = "abstract-feasible-but-not-lowered: " + not_lowered;
        return r;
    }
    if (m.stage_partition == 1) {
        r.launch_manifest = {{LaunchKind::RowFFT, 0, 0}};
    } else {
        r.launch_manifest = {{LaunchKind::RowFFT, 0, 0}, {LaunchKind::RowFFT, 1, 1}};
    }
    r.visible_launches = (int)r.launch_manifest.size();
    const uint64_t tensor_bytes = 8ull * (uint64_t)spec.n * (uint64_t)spec.batch;
    for (const auto& rec : r.launch_manifest) r.modeled_payload_gm_rw_bytes += modeled_payload_gm_rw(rec.kind, tensor_bytes);

    const size_t have = unit.ub_bytes < hw.ub_bytes_per_core ? unit.ub_bytes : hw.ub_bytes_per_core;
    std::vector<KernelResource> used;
    for (const auto& rec : r.launch_manifest) {
        const KernelResource kr = gfx_stage_resource(m.stage_lengths[rec.stage], m.stage_partition, rec.stage);
        if (kr.ub_bytes > have) {
            r.abstract_feasible = false;
            r.supported = false;
            r.reason = std::string("LDS overflow: kernel ") + kr.kernel + " (" + kr.constraint + ") needs " +
                       std::to_string(kr.ub_bytes) + " bytes, have " + std::to_string(have);
            return r;
        }
        used.push_back(kr);
    }
    r.supported = true;
    r.reason.clear();
    r.plan_ub_bytes = plan_peak(used).ub_bytes;
    int edges = 0;
    for (const auto& rec : r.launch_manifest) edges += rec.boundary_edge;
    r.materialized_gm_boundaries = edges;
    r.resident_subgraph = m.residence;
    r.whole_transform_on_chip = m.stage_partition == 1;
    r.host_assisted = false;
    return r;
}

// The two-segment chain the runtime runs above 16384 points: one stage
// service reused for both groups, a tile grid over the compute units, the
// twiddled boundary in device memory, fused into the row-FFT launches.
inline ArchitectureMapping default_long_mapping(uint32_t n1, uint32_t n2, const HardwareProfile& hw) {
    ArchitectureMapping m;
    m.stage_partition = 2;
    m.stage_lengths = {n1, n2};
    m.stage_space = 1;
    m.stage_time = 2;
    m.data_space = hw.vector_core_num;
    m.data_time = 1;
    m.role_stages = {0, 1};
    m.residence = Residence::None;
    m.boundary_layout = Layout::TwiddledStage;
    m.boundary_home = BoundaryHome::DeviceGM;
    m.boundary_impl = BoundaryImpl::Fused;
    m.pipeline_buffers = 1;
    return m;
}

// The one-stage lowering of n <= 16384: the whole transform resident in one
// workgroup, every stage under its barrier.
inline ArchitectureMapping single_pass_mapping(uint32_t n, const HardwareProfile& hw) {
    ArchitectureMapping m;
    m.stage_partition = 1;
    m.stage_lengths = {n};
    m.stage_space = 1;
    m.stage_time = 1;
    m.data_space = hw.vector_core_num;
    m.data_time = 1;
    m.role_stages = {0};
    m.residence = Residence::BlockResident;
    m.boundary_layout = Layout::NaturalOrder;
    m.boundary_home = BoundaryHome::OnChip;
    m.pipeline_buffers = 1;
    return m;
}

// What the runtime runs for a C2C of n points: the one-stage lowering up to
// 16384, the two-segment chain with the reference's split above.
inline LoweringResult runtime_lowering(uint32_t n, uint32_t batch, const HardwareProfile& hw = builtin_hardware_profile()) {
    TransformSpec spec{n, batch, Precision::FP32, Direction::C2C_FWD};
    if (n <= kGfxMaxSingleN)
        return query_lowering(spec, single_pass_mapping(n, hw), single_fft_capability(), hw);
    uint32_t lg = 0;
    while ((1u << lg) < n) lg++;
    const uint32_t n1 = 1u << (lg / 2), n2 = n / n1;
    return query_lowering(spec, default_long_mapping(n1, n2, hw), local_fft_capability(), hw);
}

}  // namespace butterfly
