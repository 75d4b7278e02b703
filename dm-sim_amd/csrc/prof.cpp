// prof.cpp -- per-launch HIP-event timing (see prof.hpp).
#include "prof.hpp"

namespace dmx {

const char* const kKernelClassNames[KC_COUNT] = {
    "k_fused4", "k_sweep(dense pair)", "k_sweep(diagonal pair)", "k_sweep(cx)", "k_quad",
    "k_plain_persistent", "k_fused5", "k_fused4_perm", "k_fused4_herm", "k_relabel"};

LaunchProfiler::~LaunchProfiler()
{
    for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
}

hipEvent_t LaunchProfiler::take()
{
    if (used_ == pool_.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        pool_.push_back(e);
    }
    return pool_[used_++];
}

void LaunchProfiler::start(hipStream_t s)
{
    stream_ = s;
    used_ = 0;
    recs_.clear();
    for (auto& st : stats_) st = KernelStat{};
}

void LaunchProfiler::begin(int cls, double bytes)
{
    if (!on_) return;
    Rec r{cls, bytes, take(), nullptr};
    if (r.a) (void)hipEventRecord(r.a, stream_);
    recs_.push_back(r);
}

void LaunchProfiler::end()
{
    if (!on_ || recs_.empty()) return;
    recs_.back().b = take();
    if (recs_.back().b) (void)hipEventRecord(recs_.back().b, stream_);
}

void LaunchProfiler::cancel()
{
    if (on_ && !recs_.empty()) recs_.pop_back();
}

hipError_t LaunchProfiler::collect()
{
    for (const Rec& r : recs_) {
        if (!r.a || !r.b) continue;
        float ms = 0.f;
        hipError_t e = hipEventElapsedTime(&ms, r.a, r.b);
        if (e != hipSuccess) return e;
        KernelStat& s = stats_[r.cls];
        s.calls++;
        s.ms += ms;
        s.bytes += r.bytes;
    }
    return hipSuccess;
}

}  // namespace dmx
