// fused.hpp -- device record and launcher of the dense 4-qubit block sweep.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dmx {

// A dense 4-qubit block U (16 x 16, row-major, local index bit t <-> the t-th
// lowest qubit of qmask).
struct FusedBlock {
    uint32_t qmask;  // exactly four qubits
    uint32_t pad[3];
    double ur[256];
    double ui[256];
};

// One circuit's rho and the block to apply to it (batched launches).
struct BatchEntry {
    double* re;
    double* im;
    const FusedBlock* blk;
};

// rho <- U rho U^dagger for one block (n >= 4).
hipError_t launch_fused4(double* re, double* im, const FusedBlock* d_blk, int n, int max_blocks,
                         hipStream_t stream);

// The same for `count` circuits of n qubits each (one launch).
hipError_t launch_fused4_batch(const BatchEntry* d_entries, int count, int n, int max_blocks,
                               hipStream_t stream);

}  // namespace dmx
