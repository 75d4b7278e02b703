// fuse.hpp -- fused execution: gates merged into dense k-qubit blocks, each
// applied to rho's row and column index in one sweep.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "lower.hpp"
#include "prim.hpp"
#include "fused.hpp"
#include "prof.hpp"

namespace dmx {

struct FusedPlan {
    virtual ~FusedPlan() = default;
    virtual int upload(hipStream_t stream) = 0;
    // Applies the whole circuit once on `stream`, timing launches into prof.
    virtual int execute(double* re, double* im, int n, hipStream_t stream, LaunchProfiler& prof,
                        uint32_t* launches, uint32_t* sweeps) = 0;
};

// Blocks of a circuit, in execution order interleaved (order) with the gates
// that are too wide for a block: order[i] = (0, block index) or (1, gate index).
struct BlockPlan {
    std::vector<FusedBlock> blocks;
    std::vector<uint32_t> gates_per_block;
    std::vector<std::pair<int, int>> order;
};

// Greedy planning into <= max_k-qubit blocks (padded to 4).  fusable = false
// (and no blocks) when the circuit must run the plain method (SRN, n < 4).
int plan_blocks(const std::vector<Gate>& circuit, int n, int max_k, BlockPlan& out, bool& fusable);

int build_fused_plan(const std::vector<Gate>& circuit, int n, int max_k, FusedPlan** out);

}  // namespace dmx
