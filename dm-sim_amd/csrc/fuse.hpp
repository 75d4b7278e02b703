// fuse.hpp -- fused execution: gates merged into dense k-qubit blocks, each
// applied to rho's row and column index in one sweep.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "lower.hpp"
#include "prim.hpp"
#include "prof.hpp"

namespace dmx {

struct FusedPlan {
    virtual ~FusedPlan() = default;
    virtual int upload(hipStream_t stream) = 0;
    // Applies the whole circuit once on `stream`, timing launches into prof.
    virtual int execute(double* re, double* im, int n, hipStream_t stream, LaunchProfiler& prof,
                        uint32_t* launches, uint32_t* sweeps) = 0;
};

int build_fused_plan(const std::vector<Gate>& circuit, int n, int max_k, FusedPlan** out);

}  // namespace dmx
