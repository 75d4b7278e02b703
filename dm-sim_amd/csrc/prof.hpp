// prof.hpp -- optional per-launch timing with HIP events on the run's stream.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace dmx {

enum KernelClass : int {
    KC_FUSED4 = 0,    // k_fused4: one dense 4-qubit block, both sides of rho
    KC_PAIR_DENSE,    // k_sweep, pair kinds that rewrite both entries
    KC_PAIR_POS1,     // k_sweep, diagonal kinds (only the bit-set entry changes)
    KC_CX,            // k_sweep, controlled swap
    KC_QUAD,          // k_quad, 4x4 primitive
    KC_PERSISTENT,    // k_plain_persistent, a segment of primitives
    KC_FUSED5,        // k_fused5: one dense 5-qubit block, both sides of rho
    KC_FUSED4P,       // k_fused4_perm: a 4-qubit block with a trailing permutation, out of place
    KC_FUSED4H,       // k_fused4_herm: a 4-qubit block on the upper tiles of Hermitian rho, out of place
    KC_RELABEL,       // k_relabel: rho's index bits relabelled, out of place
    KC_COUNT
};

extern const char* const kKernelClassNames[KC_COUNT];

struct KernelStat {
    uint32_t calls = 0;
    double ms = 0.0;
    double bytes = 0.0;  // algorithmic bytes moved, summed over calls
};

class LaunchProfiler {
public:
    ~LaunchProfiler();
    void enable(bool on) { on_ = on; }
    bool enabled() const { return on_; }
    void start(hipStream_t s);  // new run
    void begin(int cls, double bytes);
    void end();
    void cancel();  // drops the record of the last begin() (its launch did not run)
    hipError_t collect();  // after the stream is synchronised
    const KernelStat& stat(int cls) const { return stats_[cls]; }

private:
    struct Rec {
        int cls;
        double bytes;
        hipEvent_t a, b;
    };
    hipEvent_t take();
    bool on_ = false;
    hipStream_t stream_ = nullptr;
    std::vector<hipEvent_t> pool_;
    size_t used_ = 0;
    std::vector<Rec> recs_;
    KernelStat stats_[KC_COUNT];
};

}  // namespace dmx
