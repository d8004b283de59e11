// Evaluates include/butterfly/descriptors.hpp on mapping cases read from
// stdin, one per line:
//   hw n batch precision direction n1 n2 stage_partition Us Ts Ud Td role0 role1
//   residence layout home impl pipeline fold_d plane_k unit_ub unit
// hw: 0 = the built-in gfx950 profile, 1 = the reference's Ascend profile
// numbers (48 vector cores, 192 KiB per core); unit: 0 = local, 1 = single,
// 2 = the reference's row unit envelope ([2^6, 2^12], 192 KiB).
// n2 = 0 gives a one-entry stage list {n1}.  Prints one line per case:
//   supported abstract_feasible launches edges host_assisted on_chip plan_ub payload kinds#reason
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>

#include "butterfly/descriptors.hpp"

using namespace butterfly;

int main() {
    HardwareProfile ascend = {"ascend910_93", "Ascend910_9382", 24, 48, 201326592ull,
                              65787658240ull, 196608ull, 128};
    ProcessingUnitCapability ref_unit = local_fft_capability();
    ref_unit.core = "kfft_fwd_local_fft";
    ref_unit.min_local_log_n = 6;
    ref_unit.max_local_log_n = 12;
    ref_unit.ub_bytes = 196608;
    std::string line;
    while (std::getline(std::cin, line)) {
        if (line.empty()) continue;
        std::istringstream in(line);
        long v[23];
        for (long& x : v) in >> x;
        const HardwareProfile& hw = v[0] ? ascend : builtin_hardware_profile();
        TransformSpec spec{(uint32_t)v[1], (uint32_t)v[2], v[3] ? Precision::FP16 : Precision::FP32,
                           (Direction)v[4]};
        ArchitectureMapping m;
        if (v[6]) m.stage_lengths = {(uint32_t)v[5], (uint32_t)v[6]};
        else m.stage_lengths = {(uint32_t)v[5]};
        m.stage_partition = (int)v[7];
        m.stage_space = (int)v[8];
        m.stage_time = (int)v[9];
        m.data_space = (int)v[10];
        m.data_time = (int)v[11];
        if (v[6]) m.role_stages = {(int)v[12], (int)v[13]};
        else m.role_stages = {(int)v[12]};
        m.residence = (Residence)v[14];
        m.boundary_layout = (Layout)v[15];
        m.boundary_home = (BoundaryHome)v[16];
        m.boundary_impl = (BoundaryImpl)v[17];
        m.pipeline_buffers = (int)v[18];
        m.row_fft_fold_d = (uint32_t)v[19];
        m.row_fft_plane_k = (uint32_t)v[20];
        ProcessingUnitCapability unit = v[22] == 2 ? ref_unit : (v[22] == 1 ? single_fft_capability()
                                                                           : local_fft_capability());
        if (v[21]) unit.ub_bytes = (size_t)v[21];
        const LoweringResult r = query_lowering(spec, m, unit, hw);
        std::string kinds;
        for (size_t i = 0; i < r.launch_manifest.size(); i++) {
            if (i) kinds += "|";
            kinds += launch_kind_name(r.launch_manifest[i].kind);
        }
        std::printf("%d %d %d %d %d %d %zu %llu %s#%s\n", (int)r.supported, (int)r.abstract_feasible,
                    r.visible_launches, r.materialized_gm_boundaries, (int)r.host_assisted,
                    (int)r.whole_transform_on_chip, r.plan_ub_bytes,
                    (unsigned long long)r.modeled_payload_gm_rw_bytes, kinds.c_str(), r.reason.c_str());
    }
    return 0;
}
