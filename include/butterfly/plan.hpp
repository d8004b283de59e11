// Context / Plan / Transform -- the drop-in C++ API (same classes, method
// signatures, layouts and return codes as the reference's
// include/butterfly/plan.hpp:19-91), executed on gfx950 through the C ABI of
// libbfly_hip.so (include/bfly_hip.h).  No HIP or torch type appears here.
//
// Execute return codes (reference docs/reference/api.md:38-48):
//   0 ok, -1 kernel not available, -2 table preparation failed,
//   -3 device allocation failed, -4 H2D failed, -5 launch failed,
//   -6 sync failed, -7 D2H failed, -8 shape / pointer / offset guard.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "butterfly/enumerate.hpp"
#include "butterfly/objects.hpp"

namespace bfly {

class Context;
class Plan;

// One materialising action of a plan (table generation, upload, launch).
class Transform {
public:
    enum Kind : uint8_t { GenTwiddles, GenBitReverse, LaunchKernel, Measure } kind;
    std::string label;
};

class Plan {
public:
    ~Plan();
    const Candidate& candidate() const { return cand_; }
    const std::vector<Transform>& transforms() const { return tf_; }

    // Upload the input-independent twiddle table of C2C length n (idempotent
    // per n; a failed rebuild never leaves the old key valid).  batch only
    // takes part in the shape guard.
    int prepare(uint32_t n, uint32_t batch = 0xFFFFFFFFu);
    // Forward C2C, host pointers [batch][2n] interleaved -> [batch][2n].
    // n a power of two in [64, 65536]; lengths above 4096 run the four-step
    // long-FFT kernels.  batch*2*n <= 2^32-1 as in the reference.
    int run(const float* in, float* out, uint32_t n, uint32_t batch);
    // Unnormalised inverse C2C (exponent sign +1), same layout and guards.
    int runInverse(const float* in, float* out, uint32_t n, uint32_t batch);
    // R2C: host [batch][n] reals -> [batch][n+2] half spectrum; n in [128, 8192].
    int runR2C(const float* in, float* out, uint32_t n, uint32_t batch);
    // C2R: host [batch][n+2] half spectrum -> [batch][n] reals with 1/n;
    // n in [64, 4096].  Imaginary parts of the DC and Nyquist bins are ignored.
    int runC2R(const float* in, float* out, uint32_t n, uint32_t batch);
    // Device-pointer C2C on the context stream, asynchronous (no copies).
    int runDevice(const float* devIn, float* devOut, uint32_t n, uint32_t batch, int sign = -1);
    // Measured us (HIP events, best of reps) + FP64 check of sampled rows;
    // on success records the Metric and moves the candidate to Measured.
    int measure(uint32_t n, uint32_t batch, Metric* out);

private:
    friend class Context;
    explicit Plan(Candidate c);
    int prepareKind(int kind, uint32_t n, uint32_t batch);
    bool validShape(int kind, uint32_t n, uint32_t batch, bool checkOffset = true) const;
    int execute(int kind, int sign, const float* in, float* out, uint32_t n, uint32_t batch);
    Candidate cand_;
    std::vector<Transform> tf_;
    struct Impl;
    std::unique_ptr<Impl> impl_;
    void* ctxStream_ = nullptr;  // injected by Context
};

class Context {
public:
    Context();
    ~Context();

    // 0 ok; -4 configuration error; -9 repeated initialisation;
    // -1 runtime init, -2 device, -3 stream creation failed.
    // kernelPath names the kernel library (informational: the kernels are
    // linked into libbfly_hip.so).
    int init(const std::string& profilePath, const std::string& spacePath,
             const std::string& kernelPath);
    const Hardware& hardware() const { return hw_; }
    const DesignSpace& space() const { return sp_; }
    const std::string& kernelPath() const { return kernelPath_; }

    std::vector<Candidate> enumerate(uint32_t n, uint32_t batch) const;

    // Model pre-rank -> build -> measure until topK candidates passed -> fastest.
    std::unique_ptr<Plan> select(uint32_t n, uint32_t batch, int topK = 3,
                                 std::string* log = nullptr);

    std::unique_ptr<Plan> makePlan(const Candidate& c);

    // The device stream the context's plans run on (a hipStream_t).
    void* stream() const { return stream_; }

private:
    Hardware hw_;
    DesignSpace sp_;
    std::string kernelPath_;
    bool initialized_ = false;
    void* stream_ = nullptr;
};

}  // namespace bfly
