// engine.hpp -- the host-side simulation object behind the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "fuse.hpp"
#include "lower.hpp"
#include "prim.hpp"
#include "prof.hpp"

namespace dmx {

class Engine {
public:
    Engine() = default;
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    int init(uint32_t n_qubits, uint32_t n_gpus);
    int append(const Gate& g);
    int upload();
    int run();
    int clear_circuit();
    int reset_dm();
    int measure(uint32_t reps, uint32_t seed, uint32_t* out, double* total);
    int get_dm(double* re, double* im);
    int get_entries(uint64_t count, const uint64_t* idx, double* re, double* im);
    // The same with its device staging in a fixed per-device array (no
    // allocation): count <= kNoAllocMaxEntries; overlapping calls on one GPU race.
    int get_entries_noalloc(uint64_t count, const uint64_t* idx, double* re, double* im);
    int get_diag(double* diag);
    int expect_z(double* out);
    int set_method(int m);
    int set_fusion_max_k(uint32_t k);
    // Whether fused blocks may take trailing X/CX/SWAP gates as an output
    // permutation (out-of-place sweeps); takes effect at the next upload().
    int set_fusion_permute(bool on)
    {
        fusion_permute_ = on;
        return 0;
    }
    // Hermitian sweeps for four-qubit blocks (with permuting blocks on):
    // 0 off, 1 from kHermAutoQubits qubits (default), 2 from 8 qubits.  Takes
    // effect at the next upload().
    int set_fusion_hermitian(int mode)
    {
        fusion_herm_ = mode;
        return 0;
    }
    // Without a cost model, below the Hermitian-sweep size: time the plan's
    // block sweeps at every tile width at upload and keep the fastest (on by
    // default).  Takes effect at the next upload().
    int set_fusion_autotune(bool on)
    {
        fusion_autotune_ = on;
        return 0;
    }
    int tuned_tile_bits() const { return tuned_tile_bits_; }
    // Cost model for the fused planner (not owned; nullptr = greedy plan).
    // Takes effect at the next upload().
    void set_cost_model(CostModel* model) { model_ = model; }
    // Fixed super-tile width for every fused block that admits it (0 = chosen
    // by the model or the default).  Takes effect at the next upload().
    int set_tile_bits(int m);
    const PlanInfo& plan_info() const { return plan_info_; }
    // Timing-data generator: average time (microseconds) of `reps` back-to-back
    // launches of one kernel configuration on this simulation's rho, after one
    // untimed launch.  kind is a CostKind; fused: m, qmask (four qubits);
    // sweeps: target, control (cx control or 4x4 low qubit, else -1) and slow
    // (index side).  rho is left transformed by unitaries on one or both
    // sides: reset it before simulating.
    int time_launch(int kind, int m, uint32_t qmask, int target, int control, int slow, int reps,
                    double* us);
    std::string dump() const;

    uint32_t n_qubits() const { return (uint32_t)n_; }
    uint64_t dim() const { return dim_; }
    uint32_t n_gates() const { return (uint32_t)circuit_.size(); }
    double last_ms() const { return last_ms_; }
    uint32_t last_launches() const { return last_launches_; }
    uint32_t last_sweeps() const { return last_sweeps_; }
    // Persistent-loop segments of the last run that ran as one launch per
    // primitive because the cooperative launch was refused.
    uint32_t last_fallbacks() const { return last_fallbacks_; }
    // Prints the reference's per-run report after every run (off by default).
    void set_report(bool on) { report_ = on; }
    // Device memory this simulation holds (rho, second planes, tables).
    uint64_t device_bytes() const;
    LaunchProfiler& profiler() { return prof_; }

private:
    int run_plain(bool per_launch);
    int run_fused();
    void drop_graph();
    int ensure_ops_capacity(size_t count);
    int to_reference_orientation();

    int n_ = 0;
    uint64_t dim_ = 0;
    uint64_t num_ = 0;  // entries per plane, 4^n
    double* d_re_ = nullptr;  // rho (the planes swap with the second pair after an
    double* d_im_ = nullptr;  // out-of-place sweep)
    double* d_re2_ = nullptr;  // second pair, allocated for plans with permuting blocks
    double* d_im2_ = nullptr;
    double* d_diag_ = nullptr;
    PrimOp* d_ops_ = nullptr;
    size_t d_ops_cap_ = 0;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;

    bool transposed_ = false;  // device holds conj(res^T) of the reference's result
    bool uploaded_ = false;
    int method_ = 0;
    int fusion_max_k_ = 4;
    bool fusion_permute_ = true;
    int fusion_herm_ = 1;
    bool fusion_autotune_ = true;
    int tuned_tile_bits_ = 0;  // width the last upload's timing chose (0: none)
    bool herm_wanted() const;
    bool permuting_pays(const PlanInfo& perm, const PlanInfo& inplace) const;
    int persistent_blocks_ = 0;
    int max_blocks_ = 8192;  // grid cap for one-primitive launches

    std::vector<Gate> circuit_;
    std::vector<PrimOp> prims_;
    FusedPlan* plan_ = nullptr;
    // the fused run captured and replayed (profiling off): valid for the
    // orientation and the planes it starts from; it records the planes it ends
    // in (permuting sweeps swap the pairs, so a plan with an odd number of them
    // alternates between two graphs)
    struct RunGraph {
        hipGraphExec_t exec = nullptr;
        bool transposed = false;
        double *re = nullptr, *im = nullptr;  // rho at the start
        double *re_end = nullptr, *im_end = nullptr, *re2_end = nullptr, *im2_end = nullptr;
        uint32_t launches = 0;
    };
    RunGraph graphs_[2];
    int graph_next_ = 0;
    CostModel* model_ = nullptr;
    int tile_bits_ = 0;
    PlanInfo plan_info_;
    FusedBlock* d_probe_blk_ = nullptr;  // time_launch's operator records
    FusedBlock5* d_probe_blk5_ = nullptr;
    PrimOp* d_probe_op_ = nullptr;
    LaunchProfiler prof_;

    double last_ms_ = 0.0;
    uint32_t last_launches_ = 0;
    uint32_t last_sweeps_ = 0;
    uint32_t last_fallbacks_ = 0;
    bool report_ = false;
};

// thread-local error slot for the C ABI
void set_error(const std::string& msg);
const char* get_error();

}  // namespace dmx
