// Candidate enumeration over the gfx950 design space (four states, as in
// the reference's include/butterfly/enumerate.hpp:12-58).
#pragma once
#include <string>
#include <utility>
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
    std::string id;  // e.g. "r16_ud256_ts1_shared_f1_R16_stream"
    Mapping a;
    StagePlan p;
    Layout l;
    Fusion f;
    Metric q;
    State state = State::Unverified;
    std::string reason;  // why Infeasible
    // gfx950: the compiled C2C kernel (a BFLY_VARIANT_* of include/bfly_hip.h)
    // the candidate lowers to at the length it was enumerated for; 0 when
    // Infeasible.  A Plan re-lowers its candidate for every length it runs.
    int kernel = 0;
};

// Fitted cost model of one kernel: eta = c[0] + c[1] * MB moved + c[2] *
// Gpoint-log2 (n * batch * log2 n / 1e9), in microseconds, where MB moved is
// the kernel chain's HBM traffic 16 n batch (x2 for the four-step chain).
// Fitted by least squares from a measured sweep (bfly.calibrate); `loo` is
// the mean leave-one-out relative error of that fit.
struct KernelModel {
    double c[3] = {0.0, 0.0, 0.0};
    double loo = 0.0;
    int points = 0;
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
    std::vector<int> tbValues;   // A.tb: 1 = one tile per workgroup (one-shot grid), 0 = workgroups walk the batch
    std::vector<int> ubValues;   // A.ub: tiles in flight per workgroup (1 or 2)
    std::vector<std::pair<std::string, KernelModel>> models;  // Q.cost_model, by modelKey
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

// The C2C kernel (BFLY_VARIANT_*) a candidate runs at length n, or 0 with the
// reason in *why.  Pure function of the candidate's axes and n.
int lowerC2C(const Candidate& c, const Hardware& hw, uint32_t n, std::string* why);

// Name of a kernel variant as the logs spell it.
const char* kernelName(int variant);

// Cost-model key of a kernel at length n: its name, with "_single" for the
// one-workgroup-per-transform forms of 8192 / 16384 (plain, stream_lt).
std::string modelKey(int variant, uint32_t n);

// eta of `variant` on (n, batch) from the space's fitted model; the roofline
// estimate() when the space carries no fit for that kernel.
double kernelEta(const DesignSpace& sp, const Hardware& hw, int variant, uint32_t n, uint32_t batch);

}  // namespace bfly
