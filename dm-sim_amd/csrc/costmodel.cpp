// costmodel.cpp -- host side of the fusion cost model (network kernels in mlp.hip).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <string>

#include "dmsim_amd.h"
#include "engine.hpp"
#include "mlp.hpp"

namespace dmx {

#define CTRY(call)                                                                   \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            set_error(std::string(#call) + ": " + hipGetErrorString(e_));            \
            return e_ == hipErrorOutOfMemory ? DMX_ENOMEM : DMX_EHIP;                \
        }                                                                            \
    } while (0)

void cost_features(int n, int kind, int m, uint32_t qmask, int target, int control, int slow,
                   float* out)
{
    for (int i = 0; i < kMlpFeatures; i++) out[i] = 0.f;
    out[0] = n / 16.f;
    out[1 + kind] = 1.f;
    const uint64_t N = 1ull << (2 * n);
    uint64_t groups;
    if (kind == CK_FUSED || kind == CK_FUSED5 || kind == CK_HERM) {
        const int k = kind == CK_FUSED5 ? 5 : 4;
        out[8] = m / 8.f;
        if (kind == CK_HERM) {  // super-tiles on and above the diagonal, a multiple of 128
            const uint64_t H = 1ull << (n - 8);
            const uint64_t items = 16 * (H * (H + 1) / 2);
            groups = (std::min<uint64_t>(items, kFusedGridCap) + 127) / 128 * 128;
        } else {
            groups = std::min<uint64_t>(1ull << ((n - k) + (n - m)), kFusedGridCap);
        }
        out[10] = __builtin_popcount(qmask & 0xFu) / 4.f;
        out[11] = (31 - __builtin_clz(qmask)) / 16.f;
        out[12] = __builtin_ctz(qmask) / 16.f;
        uint32_t cmask = qmask;  // super-tile column bits, as the fused kernels form them
        for (int q = 0; q < n && __builtin_popcount(cmask) < m; q++) cmask |= 1u << q;
        out[13] = __builtin_ctz(~cmask) / 8.f;
    } else {
        const int shift = slow ? n : 0;
        const int b = target + shift;
        if (kind == CK_QUAD)
            groups = std::min<uint64_t>((N / 4 + 255) / 256, kSweepGridCap);
        else
            groups = N < 256 ? 1 : std::min<uint64_t>((N / 256 + 3) / 4, kSweepGridCap);
        out[10] = b < 7 ? 1.f : 0.f;
        out[11] = b / 32.f;
        out[12] = (kind == CK_QUAD ? control + shift : b) / 32.f;
        if (kind == CK_CX) out[13] = (control + shift + 1) / 32.f;
    }
    out[9] = (float)std::log2((double)groups) / 16.f;
}

CostModel::~CostModel()
{
    for (float* p : {d_P_, d_G_, d_M_, d_V_, d_X_, d_T_, d_Y_})
        if (p) (void)hipFree(p);
    if (stream_) (void)hipStreamDestroy(stream_);
}

int CostModel::init(uint32_t seed)
{
    CTRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    CTRY(hipMalloc(&d_P_, kMlpParams * sizeof(float)));
    CTRY(hipMalloc(&d_G_, kMlpParams * sizeof(float)));
    CTRY(hipMalloc(&d_M_, kMlpParams * sizeof(float)));
    CTRY(hipMalloc(&d_V_, kMlpParams * sizeof(float)));
    // He-uniform initial weights, zero biases
    std::mt19937 rng(seed);
    host_P_.assign(kMlpParams, 0.f);
    const int H = kMlpHidden, F = kMlpFeatures;
    auto fill = [&](int off, int cnt, int fan_in) {
        std::uniform_real_distribution<float> u(-std::sqrt(6.f / fan_in), std::sqrt(6.f / fan_in));
        for (int i = 0; i < cnt; i++) host_P_[off + i] = u(rng);
    };
    fill(0, H * F, F);
    fill(H * F + H, H * H, H);
    fill(H * F + H + H * H + H, H, H);
    CTRY(hipMemcpy(d_P_, host_P_.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    CTRY(mlp_zero(d_G_, stream_));
    CTRY(mlp_zero(d_M_, stream_));
    CTRY(mlp_zero(d_V_, stream_));
    CTRY(hipStreamSynchronize(stream_));
    return DMX_OK;
}

int CostModel::ensure(int count)
{
    if (count <= cap_) return DMX_OK;
    for (float** p : {&d_X_, &d_T_, &d_Y_})
        if (*p) {
            CTRY(hipFree(*p));
            *p = nullptr;
        }
    const int cap = std::max(count, 1024);
    CTRY(hipMalloc(&d_X_, (size_t)cap * kMlpFeatures * sizeof(float)));
    CTRY(hipMalloc(&d_T_, (size_t)cap * sizeof(float)));
    CTRY(hipMalloc(&d_Y_, (size_t)cap * sizeof(float)));
    cap_ = cap;
    return DMX_OK;
}

int CostModel::train(int count, const float* X, const float* y, int epochs, int batch, float lr,
                     uint32_t seed, float* final_loss)
{
    if (count < 2 || epochs < 1 || batch < 1) {
        set_error("training needs at least 2 samples, 1 epoch and batch >= 1");
        return DMX_EINVAL;
    }
    int rc = ensure(count);
    if (rc) return rc;
    double mean = 0, var = 0;
    for (int i = 0; i < count; i++) mean += y[i];
    mean /= count;
    for (int i = 0; i < count; i++) var += (y[i] - mean) * (y[i] - mean);
    var /= count;
    mu_ = (float)mean;
    sigma_ = (float)std::sqrt(var > 1e-12 ? var : 1.0);
    std::vector<int> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::mt19937 rng(seed);
    std::vector<float> xs((size_t)count * kMlpFeatures), ts(count);
    for (int ep = 0; ep < epochs; ep++) {
        CTRY(hipStreamSynchronize(stream_));  // the previous epoch's copies are done with xs, ts
        std::shuffle(order.begin(), order.end(), rng);
        for (int i = 0; i < count; i++) {
            std::copy(X + (size_t)order[i] * kMlpFeatures,
                      X + (size_t)(order[i] + 1) * kMlpFeatures, xs.begin() + (size_t)i * kMlpFeatures);
            ts[i] = (y[order[i]] - mu_) / sigma_;
        }
        CTRY(hipMemcpyAsync(d_X_, xs.data(), xs.size() * sizeof(float), hipMemcpyHostToDevice,
                            stream_));
        CTRY(hipMemcpyAsync(d_T_, ts.data(), ts.size() * sizeof(float), hipMemcpyHostToDevice,
                            stream_));
        for (int b0 = 0; b0 < count; b0 += batch) {
            const int nb = std::min(batch, count - b0);
            CTRY(mlp_backward(d_P_, d_X_ + (size_t)b0 * kMlpFeatures, d_T_ + b0, nb, (float)nb,
                              d_G_, stream_));
            CTRY(mlp_adam(d_P_, d_G_, d_M_, d_V_, lr, ++step_, stream_));
        }
    }
    // loss on the whole set (standardised targets)
    CTRY(hipStreamSynchronize(stream_));
    for (int i = 0; i < count; i++) ts[i] = (y[i] - mu_) / sigma_;
    CTRY(hipMemcpyAsync(d_X_, X, (size_t)count * kMlpFeatures * sizeof(float),
                        hipMemcpyHostToDevice, stream_));
    CTRY(mlp_forward(d_P_, d_X_, count, d_Y_, stream_));
    std::vector<float> out(count);
    CTRY(hipMemcpyAsync(out.data(), d_Y_, count * sizeof(float), hipMemcpyDeviceToHost, stream_));
    CTRY(hipStreamSynchronize(stream_));
    double loss = 0;
    for (int i = 0; i < count; i++) loss += (out[i] - ts[i]) * (out[i] - ts[i]);
    if (final_loss) *final_loss = (float)(loss / count);
    CTRY(hipMemcpy(host_P_.data(), d_P_, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    memo_.clear();
    return DMX_OK;
}

int CostModel::predict(int count, const float* X, float* out)
{
    if (count <= 0) return DMX_OK;
    int rc = ensure(count);
    if (rc) return rc;
    CTRY(hipMemcpyAsync(d_X_, X, (size_t)count * kMlpFeatures * sizeof(float),
                        hipMemcpyHostToDevice, stream_));
    CTRY(mlp_forward(d_P_, d_X_, count, d_Y_, stream_));
    CTRY(hipMemcpyAsync(out, d_Y_, count * sizeof(float), hipMemcpyDeviceToHost, stream_));
    CTRY(hipStreamSynchronize(stream_));
    for (int i = 0; i < count; i++) out[i] = out[i] * sigma_ + mu_;
    return DMX_OK;
}

int CostModel::gradient(int count, const float* X, const float* y, float* out)
{
    if (count <= 0) {
        set_error("gradient needs samples");
        return DMX_EINVAL;
    }
    int rc = ensure(count);
    if (rc) return rc;
    CTRY(hipMemcpyAsync(d_X_, X, (size_t)count * kMlpFeatures * sizeof(float),
                        hipMemcpyHostToDevice, stream_));
    CTRY(hipMemcpyAsync(d_T_, y, count * sizeof(float), hipMemcpyHostToDevice, stream_));
    CTRY(mlp_zero(d_G_, stream_));
    CTRY(mlp_backward(d_P_, d_X_, d_T_, count, (float)count, d_G_, stream_));
    CTRY(hipMemcpyAsync(out, d_G_, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, stream_));
    CTRY(mlp_zero(d_G_, stream_));
    CTRY(hipStreamSynchronize(stream_));
    return DMX_OK;
}

int CostModel::get_params(float* out) const
{
    std::copy(host_P_.begin(), host_P_.end(), out);
    return DMX_OK;
}

int CostModel::set_params(const float* in)
{
    host_P_.assign(in, in + kMlpParams);
    CTRY(hipMemcpy(d_P_, host_P_.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    memo_.clear();
    return DMX_OK;
}

int CostModel::price(const std::vector<std::vector<float>>& feats, std::vector<float>& us)
{
    us.assign(feats.size(), 0.f);
    std::vector<size_t> todo;
    std::vector<float> X;
    for (size_t i = 0; i < feats.size(); i++) {
        auto it = memo_.find(feats[i]);
        if (it != memo_.end()) {
            us[i] = it->second;
        } else {
            todo.push_back(i);
            X.insert(X.end(), feats[i].begin(), feats[i].end());
        }
    }
    if (todo.empty()) return DMX_OK;
    std::vector<float> y(todo.size());
    int rc = predict((int)todo.size(), X.data(), y.data());
    if (rc) return rc;
    for (size_t k = 0; k < todo.size(); k++) {
        const float t = std::exp(y[k]);  // the model is trained on ln(microseconds)
        us[todo[k]] = t;
        memo_[feats[todo[k]]] = t;
    }
    return DMX_OK;
}

}  // namespace dmx
