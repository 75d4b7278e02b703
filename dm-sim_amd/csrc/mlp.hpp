// mlp.hpp -- the fusion cost model: network kernels (mlp.hip), the host model
// object, and the feature encoding of a candidate launch.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <vector>

namespace dmx {

constexpr int kMlpFeatures = 14;
constexpr int kMlpHidden = 64;
constexpr int kMlpParams = kMlpHidden * kMlpFeatures + kMlpHidden + kMlpHidden * kMlpHidden +
                           kMlpHidden + kMlpHidden + 1;

hipError_t mlp_forward(const float* d_params, const float* d_X, int count, float* d_Y,
                       hipStream_t stream);
hipError_t mlp_backward(const float* d_params, const float* d_X, const float* d_T, int count,
                        float norm, float* d_G, hipStream_t stream);
hipError_t mlp_adam(float* d_P, float* d_G, float* d_M, float* d_V, float lr, int step,
                    hipStream_t stream);
hipError_t mlp_zero(float* d_G, hipStream_t stream);

// Launch kinds the model prices.
enum CostKind : int {
    CK_FUSED = 0,  // k_fused4, one sweep for both sides
    CK_DENSE,      // k_sweep of a dense pair kind, one side
    CK_DIAG,       // k_sweep of a diagonal kind, one side
    CK_CX,         // k_sweep controlled swap, one side
    CK_QUAD,       // k_quad, one side
    CK_FUSED5,     // k_fused5, one sweep for both sides
    CK_HERM,       // k_fused4_herm, one sweep for both sides (upper tiles, half writes)
};

// Feature vector of one launch (values scaled to about [0, 1]).  A sweep
// launch on the slow index (slow = 1) acts on flattened index bit q + n.
//  0 n/16 | 1..7 kind one-hot | 8 tile bits m/8 (fused kinds) |
//  9 log2(workgroups launched)/16 |
//  10 fused: block qubits below 4, /4; sweep: narrow-path flag (bit < 7) |
//  11 fused: highest block qubit/16; sweep: target bit/32 |
//  12 fused: lowest block qubit/16; sweep: low bit of a 4x4, else target, /32 |
//  13 fused: contiguous low super-tile column bits/8; cx: control bit+1 /32
// Hermitian sweeps: m = 8, qmask = the block's qubits in the circuit's order.
// fused: qmask = the block's four (five) qubits; sweeps: target (4x4: high
// qubit), control (cx: control qubit; 4x4: low qubit; else -1).
void cost_features(int n, int kind, int m, uint32_t qmask, int target, int control, int slow,
                   float* out);

// Grid sizes the launchers use (fused: k_fused4 with this cap).
constexpr int kFusedGridCap = 4096;
constexpr int kSweepGridCap = 8192;

class CostModel {
public:
    ~CostModel();
    int init(uint32_t seed);
    // Trains on targets y (any scale): stores their mean/std and fits the
    // network to the standardised targets.
    int train(int count, const float* X, const float* y, int epochs, int batch, float lr,
              uint32_t seed, float* final_loss);
    // Predictions in the targets' original scale.
    int predict(int count, const float* X, float* out);
    // Gradient of mean((f(x) - y)^2) with respect to the parameters, for the
    // network as it is (no target standardisation).
    int gradient(int count, const float* X, const float* y, float* out);
    int get_params(float* out) const;
    int set_params(const float* in);
    void target_scale(float* mu, float* sigma) const
    {
        *mu = mu_;
        *sigma = sigma_;
    }
    void set_target_scale(float mu, float sigma)
    {
        mu_ = mu;
        sigma_ = sigma;
    }
    // Planner helper: predicted launch time (microseconds) with a host-side
    // memo of feature vectors already priced.
    int price(const std::vector<std::vector<float>>& feats, std::vector<float>& us);

private:
    int ensure(int count);
    float* d_P_ = nullptr;
    float* d_G_ = nullptr;
    float* d_M_ = nullptr;
    float* d_V_ = nullptr;
    float* d_X_ = nullptr;
    float* d_T_ = nullptr;
    float* d_Y_ = nullptr;
    int cap_ = 0;
    int step_ = 0;
    float mu_ = 0.f, sigma_ = 1.f;
    hipStream_t stream_ = nullptr;
    std::vector<float> host_P_;
    std::map<std::vector<float>, float> memo_;
};

}  // namespace dmx
