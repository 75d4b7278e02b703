// batch.hpp -- batched executor: many circuits of uneven width and depth, all
// rho resident in one pool, one graph launch per run.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "fuse.hpp"
#include "fused.hpp"
#include "lower.hpp"
#include "prof.hpp"

namespace dmx {

struct PauliTerm {
    uint64_t xmask;   // qubits carrying X or Y
    uint64_t zymask;  // qubits carrying Z or Y
    int32_t ny;       // number of Y factors
    int32_t pad;
    double weight;
};

struct EnergyJob {
    const double* re;
    const double* im;
    int32_t n;
    int32_t first_term;
    int32_t n_terms;
    int32_t pad;
};

hipError_t launch_pool_reset(double* pool, uint64_t n_doubles, const uint64_t* d_re_off,
                             int count, hipStream_t stream);
hipError_t launch_pauli_energy(const EnergyJob* d_jobs, const PauliTerm* d_terms, int count,
                               double* d_out, hipStream_t stream);
// out[diag_off[c] + i] = Re rho_c[i][i] for the `count` circuits of jobs.
hipError_t launch_batch_diag(const EnergyJob* d_jobs, const uint64_t* d_diag_off, int count,
                             double* d_out, hipStream_t stream);

class Batch {
public:
    Batch() = default;
    ~Batch();
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    int init();
    int add_circuit(uint32_t n, std::vector<Gate> gates, uint32_t* id);
    int add_pauli(uint32_t id, const char* pauli, double weight);
    int upload();
    int run();
    int energies(double* out) const;
    // rho of a circuit: only for circuits of the last wave when the batch
    // runs in more than one (the earlier ones' pool space is reused)
    int get_dm(uint32_t id, double* re, double* im);
    // Re rho_ii of any circuit (kept for every wave)
    int get_diag(uint32_t id, double* out);
    // Device bytes one pool may take (0: free memory at upload less 2 GiB,
    // halved when permuting blocks need a second pool).  Circuits are run in
    // waves that fit, one after another inside a run.
    int set_pool_limit(uint64_t bytes);
    uint32_t waves() const { return (uint32_t)wave_first_.size(); }
    void set_graph(bool on) { use_graph_ = on; }
    // Tile-width autotuning at upload (default on): per qubit count, the
    // super-tile width whose batched block launch times fastest on this GPU;
    // off = the default width.
    int set_autotune(bool on);
    // Super-tile width used for circuits of n qubits (0 if none in the batch).
    int tile_bits(int n) const;
    // Schedule chosen at upload: rho bytes per group of batched circuits that
    // run all their block steps before the next group (0: each wave is one
    // group), and the number of groups.  Timed among several budgets when
    // autotuning is on; otherwise 0.
    uint64_t group_bytes() const { return group_bytes_; }
    uint32_t groups() const { return groups_; }
    // Fixes the group budget (bytes, 0 = whole waves) instead of timing it.
    int set_group_bytes(uint64_t bytes);
    // Planner network for the circuits that run on their own (used at upload;
    // null: the greedy planner).
    void set_cost_model(CostModel* m) { model_ = m; }
    // From this many qubits a circuit's rho alone fills the GPU: it runs on
    // its own inside its wave, planned as the single-circuit engine plans it
    // (Hermitian block sweeps, five-qubit gates kept whole), instead of in
    // the batched block launches.
    static constexpr int kSoloQubits = kHermAutoQubits;
    // Circuits of kBatchHermQubits..kSoloQubits-1 qubits run Hermitian block
    // sweeps (batched k_fused4_herm / k_relabel launches; planned as the
    // single-circuit engine plans with dmx_set_fusion_hermitian(s, 2)) when on
    // (the default); off: in-order blocks.  Measured on 43 hardware-efficient
    // circuits of one width, Hermitian against in-order: 0.75x the time at 11
    // qubits, 0.98x at 10, 1.26x at 9 and 1.58x at 8, whose sweeps have too
    // few items to fill the GPU.
    int set_hermitian(bool on);
    static constexpr int kBatchHermQubits = 10;
    // Workgroups of one batched Hermitian launch in all (shared among its
    // circuits), so that each takes several upper-tile items in turn and its
    // next item's loads overlap the transform.  256 hardware-efficient
    // circuits of 6..11 qubits, one count for both widths: 61.2 ms per run at
    // 2048, 62.0 at 1536, 63.0 at 3072, 64.8 at 4096 (the in-order and relabel
    // launches are indifferent); then 10-qubit launches alone: 59.5 ms at
    // 2048, 58.9 at 1536 and at 1024, 60.9 at 3072.
    static constexpr int herm_batch_blocks(int n) { return n >= 11 ? 2048 : 1536; }

    double last_ms() const { return last_ms_; }
    uint32_t launches() const
    {
        uint32_t solo = 0;
        for (const Circ& c : circs_) solo += c.solo_launches;
        return (uint32_t)launches_.size() + solo +
               (uint32_t)wave_first_.size() * (waves() > 1 ? 4 : 3);
    }
    uint32_t circuits() const { return (uint32_t)circs_.size(); }
    uint64_t gates() const;
    // Algorithmic bytes of one run's block sweeps: read + write of both planes
    // (32 * 4^n bytes) per batched block, and what the steps of the circuits
    // that run on their own move.
    double sweep_bytes() const;

private:
    struct Circ {
        int n;
        std::vector<Gate> orig;   // as added
        std::vector<Gate> gates;  // C4X expanded (batched blocks have at most four qubits)
        bool solo = false;        // runs on its own (kSoloQubits and up)
        bool herm = false;        // its batched steps are Hermitian sweeps and relabellings
        double bytes = 0;         // algorithmic bytes of its steps
        std::unique_ptr<FusedPlan> plan;  // its plan then
        uint32_t solo_launches = 0;
        std::vector<PauliTerm> paulis;
        uint64_t re_off = 0, im_off = 0;
        size_t blk_off = 0;
        size_t n_blocks = 0;
        bool final_alt = false;  // rho ends in the second pool (odd number of permuting blocks)
        int wave = 0;
        uint64_t diag_off = 0;  // its diagonal in the kept diagonals (several waves)
    };
    enum StepKind : uint8_t { SK_INORDER = 0, SK_HERM = 1, SK_RELABEL = 2 };
    struct Launch {
        int n;
        int m;
        int count;
        size_t entry_off;
        int wave;
        int kind;  // StepKind
    };
    hipError_t launch(const Launch& l, const BatchEntry* d_entries);
    using LaunchBuilder =
        std::function<size_t(uint64_t, std::vector<BatchEntry>&, std::vector<Launch>&)>;
    int enqueue();  // the whole run on stream_ (captured or eager)
    int autotune_groups(const LaunchBuilder& build);
    int autotune(std::vector<FusedBlock>& blocks);
    void free_device();

    std::vector<Circ> circs_;
    std::vector<Launch> launches_;
    CostModel* model_ = nullptr;
    bool hermitian_ = true;
    std::vector<uint8_t> kinds_;  // StepKind of every batched step (parallel to the blocks)
    LaunchProfiler quiet_;  // disabled: the solo plans' executor times nothing
    bool uploaded_ = false;
    bool use_graph_ = true;
    bool autotune_ = true;
    bool ran_ = false;
    std::map<int, int> tile_bits_;  // qubit count -> super-tile width
    uint64_t pool_limit_ = 0;
    uint64_t group_bytes_ = 0;
    uint32_t groups_ = 0;
    bool group_fixed_ = false;
    std::vector<int> wave_first_, wave_count_;  // circuits of each wave (contiguous ids)
    std::vector<uint64_t> wave_doubles_;        // pool doubles each wave uses
    double* d_diag_ = nullptr;                  // kept diagonals (several waves)
    uint64_t* d_diag_off_ = nullptr;

    double* pool_ = nullptr;
    double* pool_alt_ = nullptr;  // second pool, for circuits with permuting blocks
    uint64_t pool_doubles_ = 0;
    FusedBlock* d_blocks_ = nullptr;
    BatchEntry* d_entries_ = nullptr;
    uint64_t* d_re_off_ = nullptr;
    PauliTerm* d_terms_ = nullptr;
    EnergyJob* d_jobs_ = nullptr;
    double* d_energy_ = nullptr;
    std::vector<double> energy_;

    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    double last_ms_ = 0.0;
};

}  // namespace dmx
