// engine.hpp -- the host-side simulation object behind the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "lower.hpp"
#include "prim.hpp"
#include "prof.hpp"

namespace dmx {

struct FusedPlan;  // fuse.hpp

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
    int get_diag(double* diag);
    int expect_z(double* out);
    int set_method(int m);
    int set_fusion_max_k(uint32_t k);
    std::string dump() const;

    uint32_t n_qubits() const { return (uint32_t)n_; }
    uint64_t dim() const { return dim_; }
    uint32_t n_gates() const { return (uint32_t)circuit_.size(); }
    double last_ms() const { return last_ms_; }
    uint32_t last_launches() const { return last_launches_; }
    uint32_t last_sweeps() const { return last_sweeps_; }
    LaunchProfiler& profiler() { return prof_; }

private:
    int run_plain(bool per_launch);
    int run_fused();
    int ensure_ops_capacity(size_t count);
    int to_reference_orientation();

    int n_ = 0;
    uint64_t dim_ = 0;
    uint64_t num_ = 0;  // entries per plane, 4^n
    double* d_re_ = nullptr;
    double* d_im_ = nullptr;
    double* d_diag_ = nullptr;
    PrimOp* d_ops_ = nullptr;
    size_t d_ops_cap_ = 0;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;

    bool transposed_ = false;  // device holds conj(res^T) of the reference's result
    bool uploaded_ = false;
    int method_ = 0;
    int fusion_max_k_ = 4;
    int persistent_blocks_ = 0;
    int max_blocks_ = 8192;  // grid cap for one-primitive launches

    std::vector<Gate> circuit_;
    std::vector<PrimOp> prims_;
    FusedPlan* plan_ = nullptr;
    LaunchProfiler prof_;

    double last_ms_ = 0.0;
    uint32_t last_launches_ = 0;
    uint32_t last_sweeps_ = 0;
};

// thread-local error slot for the C ABI
void set_error(const std::string& msg);
const char* get_error();

}  // namespace dmx
