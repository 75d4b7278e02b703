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

class CostModel;  // mlp.hpp

// What the planner chose.  predicted_us is the cost model's estimate of the
// run's launches (0 without a model); greedy_us its estimate for every block
// fused at the default tile width.
struct PlanInfo {
    uint32_t fused_blocks = 0;   // four-qubit blocks run fused
    uint32_t plain_blocks = 0;   // four-qubit blocks run as primitive sweeps on both indices
    uint32_t fused5_blocks = 0;  // five-qubit blocks run fused (a split one counts as its
                                 // four-qubit sub-blocks instead)
    uint32_t channels = 0;       // noise channels (one superoperator sweep each)
    uint32_t perm_blocks = 0;    // fused four-qubit blocks with a trailing permutation
    uint32_t herm_blocks = 0;    // four-qubit blocks run as Hermitian sweeps
    uint32_t relabels = 0;       // relabelling sweeps
    uint32_t grouping = 0;       // 0 least growth, 1 frontier, 2..5 frontier, other scores
    uint32_t tile_hist[9] = {};  // fused blocks per tile width m (4..8)
    double predicted_us = 0.0;
    double greedy_us = 0.0;
};

// One step of a fused plan as the executor runs it (host-side checks).
struct PlanStepView {
    enum Kind { FUSED4, FUSED5, HERM, RELABEL, CHANNEL, OTHER } kind;
    const FusedBlock* block;    // FUSED4, HERM, RELABEL: its record as launched
    const FusedBlock5* block5;  // FUSED5
    const PrimOp* prim;         // CHANNEL: its record for the stored orientation (the
                                // conj-transposed one follows it)
};

// Algorithmic bytes of one Hermitian block sweep (k_fused4_herm) on n >= 8
// qubits: the super-tiles on and above the diagonal read (16 * H(H+1)/2 of
// them, H = 2^(n-8), 64 KiB each); written: every entry, or (half, when the
// next step is another Hermitian sweep) the tiles on and above that sweep's
// diagonal, T(T+1)/2 of the T^2 tiles (T = 2^(n-4)).
inline double herm_bytes(int n, bool half)
{
    const double H = (double)(1ull << (n - 8));
    const double T = (double)(1ull << (n - 4));
    const double rho = 16.0 * (double)(1ull << (2 * n));
    return 16.0 * H * (H + 1.0) / 2.0 * 65536.0 + (half ? rho * (T + 1.0) / (2.0 * T) : rho);
}

struct FusedPlan {
    virtual ~FusedPlan() = default;
    virtual void describe(std::vector<PlanStepView>& out) const = 0;
    virtual int upload(hipStream_t stream) = 0;
    // Whether execute needs a second pair of planes (blocks with a trailing
    // permutation run out of place).
    virtual bool needs_alt() const = 0;
    // Applies the whole circuit once on `stream`, timing launches into prof.
    // transposed: rho is held conj-transposed (after a plain run), which
    // selects the channel superoperators' records.  rho is in (re, im); an
    // out-of-place sweep writes (re2, im2) and swaps the pairs, so on return
    // (re, im) holds rho again.
    virtual int execute(double*& re, double*& im, double*& re2, double*& im2, int n,
                        bool transposed, hipStream_t stream, LaunchProfiler& prof,
                        uint32_t* launches, uint32_t* sweeps) = 0;
    // After upload: times up to eight of the plan's four-qubit block sweeps (in
    // place, on the scratch planes re, im) at every tile width from 5 to
    // min(8, n) -- and 4 when every sampled block holds qubit 0 -- and gives
    // every block the fastest width where it is admissible.  *chosen = that
    // width (0: no block to time).  Updates info's width histogram.
    virtual int autotune_tiles(int n, double* re, double* im, hipStream_t stream, int* chosen,
                               PlanInfo& info) = 0;
    // Algorithmic bytes of one execute() on n qubits (rho not transposed):
    // the sum of what its launches record in the profiler.
    virtual double bytes(int n) const = 0;
};

// Blocks of a circuit in execution order: order[i] = (0, index into blocks)
// for a block of at most four qubits (padded to four), (2, index into
// blocks5) for a five-qubit block, (3, circuit index) for a channel, which no
// gate crosses.
struct BlockPlan {
    std::vector<FusedBlock> blocks;
    std::vector<std::vector<int>> members;  // circuit indices of each block's gates, in order
    // members[b][0, dense[b]) form the block matrix; the rest are trailing
    // permutation gates (X, CX, SWAP) applied as the block's index map
    std::vector<int> dense;
    std::vector<FusedBlock5> blocks5;
    std::vector<std::vector<int>> members5;
    std::vector<std::pair<int, int>> order;
};

// Greedy planning into blocks of at most max(max_k, widest gate) qubits
// (max_k <= 5; blocks below four qubits are padded to four).  fusable = false
// (and no blocks) when the circuit must run the plain method (SRN, n < 4).
// allow_perm (n <= 16): a four-qubit block may also take the X, CX and SWAP
// gates that become ready after it, run as a permutation of its output
// (launch_fused4_perm, out of place).
// herm: the plan runs Hermitian sweeps (k_fused4_herm), whose output map must
// send the block's index bits onto bits 0..3: a trailing CX then joins only if
// its control is not a block bit (after the block's earlier trailing gates)
// while its target is not.
// model: choose among the least-growth grouping, the frontier grouping and the
// frontier grouping scored by the gates' predicted unfused times (all gates,
// multi-qubit gates only, square roots) and by structural scores with heavy
// one-qubit gates, the one whose blocks the network prices cheapest (each the
// cheaper of its fused sweep and its gates unfused); without a model the one
// with fewer sweeps of the first two.  *grouping (may be null) = the choice:
// 0..5 in that order.
int plan_blocks(const std::vector<Gate>& circuit, int n, int max_k, BlockPlan& out, bool& fusable,
                bool allow_perm = false, bool herm = false, CostModel* model = nullptr,
                uint32_t* grouping = nullptr);

// Relative extra cost of a sweep with a trailing permutation on n qubits:
// about 2.5 us of index-map set-up against the sweep's own time (32 * 4^n bytes
// at ~4.6 TB/s plus ~8 us of launch latency), at least 3 % (measured at n = 8
// and n = 14).
inline double permute_overhead(int n)
{
    const double sweep_us = 32.0 * (double)(1ull << (2 * n)) / 4.6e6 + 8.0;
    const double r = 2.5 / sweep_us;
    return r > 0.03 ? r : 0.03;
}

// From this many qubits a permuting plan is kept without comparison; below, it
// is kept only if its sweeps, each permuting one charged permute_overhead(n)
// extra, are fewer than the in-place plan's.
constexpr int kPermuteSureQubits = 12;
inline bool permuting_pays(int n, double sweeps, double permuting, double sweeps_inplace)
{
    return n >= kPermuteSureQubits || sweeps + permute_overhead(n) * permuting < sweeps_inplace;
}



// Plan of a circuit for the fused executor.  With a cost model, each block is
// priced as a fused launch at every tile width and as its primitive sweeps
// (a five-qubit block: as its member gates re-planned into four-qubit
// blocks), and the cheapest is taken; a plan with five-qubit blocks is kept
// only if predicted faster than the four-qubit plan of the same circuit.
// Without a model, every block is fused at the default tile width (the
// greedy plan).
// tile_bits > 0 fixes every fused block's tile width where that width is
// admissible for the block (tile_bits_ok), and takes precedence over the model.
// herm (n = 8..16, with allow_perm): four-qubit blocks run as Hermitian
// sweeps (k_fused4_herm) with rho's index bits relabelled so that each
// block's output lands in aligned 16-column runs; rho returns to the
// circuit's bit order (k_relabel) before any other step and at the end.
int build_fused_plan(const std::vector<Gate>& circuit, int n, int max_k, CostModel* model,
                     int tile_bits, bool allow_perm, FusedPlan** out, PlanInfo* info,
                     bool herm = false);

// Smallest n for Hermitian plans, and the smallest where the engine uses
// them by default.
constexpr int kHermMinQubits = 8;
constexpr int kHermAutoQubits = 12;

}  // namespace dmx
