// Candidate enumeration over the gfx950 design space (four states, as in
// the reference's include/butterfly/enumerate.hpp:12-58).
#pragma once
#include <string>
#include <vector>

#include "butterfly/objects.hpp"

namespace bfly {

enum class State : uint8_t {
    Infeasible = 0,  // illegal for the hardware or without a lowering; never measured
    Unverified = 1,  // legal, ranked by the model only
    Feasible = 2,    // passed the correctness gate
    Measured = 3,    // measured time recorded
};

const char* toString(State s);

struct Candidate {
    std::string id;  // e.g. "r16_ud256_ts1_shared_f1_R16"
    Mapping a;
    StagePlan p;
    Layout l;
    Fusion f;
    Metric q;
    State state = State::Unverified;
    std::string reason;  // why Infeasible
};

struct DesignSpace {
    std::vector<int> udCores;
    std::vector<int> radices;
    std::vector<std::string> localExchanges;
    std::vector<int> fusionStages;
    std::vector<std::string> layoutsIn;
    std::vector<std::string> layoutsOut;
    std::vector<int> tsValues;
    std::vector<int> coeffResidency;
    std::vector<int> dftMatrix;  // F.dft_matrix axis (0 plain, 1 MFMA)
    std::string raw;

    // Strict loading: missing/malformed files or axes throw std::exception.
    static DesignSpace load(const std::string& path);
    void applyDefaults();
    // R = radix^fusionLevel
    static int pointSize(int radix, int fusionLevel);
};

std::vector<Candidate> enumerate(const DesignSpace& sp, const Hardware& hw, uint32_t n,
                                 uint32_t batch);

// Non-Infeasible candidates, Measured > Feasible > Unverified, then by eta.
std::vector<const Candidate*> rank(const std::vector<Candidate>& cs);

}  // namespace bfly
