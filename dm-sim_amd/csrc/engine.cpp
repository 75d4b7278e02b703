// engine.cpp -- Simulation host object: rho resident in HBM, a packed device
// gate table, runs, measurement.  Reference: src/dmsim_amdgpu_omp.hpp:208-829.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>

#include "dmsim_amd.h"
#include "fuse.hpp"
#include "noalloc.hpp"
#include "fused.hpp"
#include "mlp.hpp"
#include "sweep.hpp"

namespace dmx {

namespace {
thread_local std::string t_error;
}

void set_error(const std::string& msg) { t_error = msg; }
const char* get_error() { return t_error.c_str(); }

#define HIPTRY(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            set_error(std::string(#call) + ": " + hipGetErrorString(e_));             \
            return e_ == hipErrorOutOfMemory ? DMX_ENOMEM : DMX_EHIP;                 \
        }                                                                             \
    } while (0)

Engine::~Engine()
{
    drop_graph();
    delete plan_;
    if (d_re_) (void)hipFree(d_re_);
    if (d_im_) (void)hipFree(d_im_);
    if (d_re2_) (void)hipFree(d_re2_);
    if (d_im2_) (void)hipFree(d_im2_);
    if (d_diag_) (void)hipFree(d_diag_);
    if (d_ops_) (void)hipFree(d_ops_);
    if (d_probe_blk_) (void)hipFree(d_probe_blk_);
    if (d_probe_blk5_) (void)hipFree(d_probe_blk5_);
    if (d_probe_op_) (void)hipFree(d_probe_op_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

int Engine::init(uint32_t n_qubits, uint32_t n_gpus)
{
    if (n_gpus != 1) {
        set_error("n_gpus must be 1: this engine runs one GPU per process");
        return DMX_EINVAL;
    }
    if (n_qubits < 1 || n_qubits > 16) {
        set_error("n_qubits must be in 1..16 (rho needs 16*4^n bytes of HBM)");
        return DMX_EINVAL;
    }
    n_ = (int)n_qubits;
    dim_ = 1ull << n_;
    num_ = dim_ * dim_;
    HIPTRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    HIPTRY(hipEventCreate(&ev0_));
    HIPTRY(hipEventCreate(&ev1_));
    HIPTRY(hipMalloc(&d_re_, num_ * sizeof(double)));
    HIPTRY(hipMalloc(&d_im_, num_ * sizeof(double)));
    HIPTRY(hipMalloc(&d_diag_, dim_ * sizeof(double)));
    HIPTRY(plain_persistent_grid(n_, &persistent_blocks_));
    void* scratch = nullptr;  // the allocation-free gather's staging on this device
    HIPTRY(noalloc_scratch(&scratch));
    return reset_dm();
}

int Engine::append(const Gate& g)
{
    if (g.op < 0 || g.op >= OP_COUNT) {
        set_error("unknown op code " + std::to_string(g.op));
        return DMX_EINVAL;
    }
    const int k = g.op == OP_MAT ? g.k : op_arity(g.op);
    if (g.op == OP_KRAUS && (g.mre.size() != 32 || g.mim.size() != 32)) {
        set_error("channel record needs two 4x4 superoperators");
        return DMX_EINVAL;
    }
    if (g.op == OP_MAT) {
        if (k < 1 || k > 5 || (int)g.mre.size() != (1 << (2 * k)) ||
            (int)g.mim.size() != (1 << (2 * k))) {
            set_error("matrix gate needs 1 <= k <= 5 and 4^k entries per plane");
            return DMX_EINVAL;
        }
        if (k > n_) {
            set_error("matrix gate wider than the register");
            return DMX_EINVAL;
        }
    }
    for (int i = 0; i < k; i++) {
        if ((int)g.qb[i] >= n_) {
            set_error("qubit index " + std::to_string(g.qb[i]) + " out of range for " +
                      std::to_string(n_) + " qubits");
            return DMX_EINVAL;
        }
        for (int j = 0; j < i; j++)
            if (g.qb[i] == g.qb[j]) {
                set_error(std::string(g.op == OP_MAT ? "MAT" : kOpNames[g.op]) +
                          ": operands must be distinct qubits");
                return DMX_EINVAL;
            }
    }
    Gate copy = g;
    for (int i = k; i < 5; i++) copy.qb[i] = 0;
    circuit_.push_back(std::move(copy));
    return DMX_OK;
}

int Engine::ensure_ops_capacity(size_t count)
{
    if (count <= d_ops_cap_) return DMX_OK;
    if (d_ops_) HIPTRY(hipFree(d_ops_));
    d_ops_ = nullptr;
    d_ops_cap_ = 0;
    const size_t cap = std::max<size_t>(count, 64);
    HIPTRY(hipMalloc(&d_ops_, cap * sizeof(PrimOp)));
    d_ops_cap_ = cap;
    return DMX_OK;
}

int Engine::upload()
{
    if (uploaded_) {
        set_error("upload() called again without clear_circuit()");
        return DMX_ESTATE;
    }
    prims_.clear();
    bool wide = false;
    for (const Gate& g : circuit_)
        if (!lower_gate(g, prims_)) wide = true;
    drop_graph();
    delete plan_;
    plan_ = nullptr;
    plan_info_ = PlanInfo{};
    if (wide && method_ != DMX_METHOD_FUSED) {
        // a k >= 3 matrix has no primitive sweep; it needs the fused executor
        set_error("matrix gates on 3..5 qubits need DMX_METHOD_FUSED");
        return DMX_ESTATE;
    }
    if (!prims_.empty()) {
        int rc = ensure_ops_capacity(prims_.size());
        if (rc) return rc;
        HIPTRY(hipMemcpyAsync(d_ops_, prims_.data(), prims_.size() * sizeof(PrimOp),
                              hipMemcpyHostToDevice, stream_));
    }
    if (method_ == DMX_METHOD_FUSED) {
        // no plan (SRN in the circuit, or n < 4): the run uses the plain sweeps
        const bool herm = herm_wanted();
        int rc = build_fused_plan(circuit_, n_, fusion_max_k_, model_, tile_bits_,
                                  fusion_permute_, &plan_, &plan_info_, herm);
        if (rc) return rc;
        if (plan_ && plan_info_.perm_blocks > 0 && n_ < kPermuteSureQubits && fusion_herm_ != 2) {
            // a short sweep is latency-bound and a permuting one pays its index
            // map's set-up: keep the permuting plan only if it still wins
            FusedPlan* inplace = nullptr;
            PlanInfo ii;
            rc = build_fused_plan(circuit_, n_, fusion_max_k_, model_, tile_bits_, false,
                                  &inplace, &ii);
            if (rc) return rc;
            if (inplace && !permuting_pays(plan_info_, ii)) {
                drop_graph();
                delete plan_;
                plan_ = inplace;
                plan_info_ = ii;
            } else {
                delete inplace;
            }
        }
        if (plan_ && plan_->needs_alt() && !d_re2_) {
            // permuting blocks run out of place: a second pair of planes, or
            // (no room for it) a plan without them
            double *re2 = nullptr, *im2 = nullptr;
            hipError_t e = hipMalloc(&re2, num_ * sizeof(double));
            const bool got_re = e == hipSuccess;
            if (got_re) e = hipMalloc(&im2, num_ * sizeof(double));
            if (e == hipSuccess) {
                d_re2_ = re2;
                d_im2_ = im2;
            } else {
                (void)hipGetLastError();
                if (got_re) (void)hipFree(re2);
                drop_graph();
                delete plan_;
                plan_ = nullptr;
                rc = build_fused_plan(circuit_, n_, fusion_max_k_, model_, tile_bits_, false,
                                      &plan_, &plan_info_);
                if (rc) return rc;
            }
        }
        if (plan_) {
            rc = plan_->upload(stream_);
            if (rc) return rc;
            tuned_tile_bits_ = 0;
            if (!model_ && tile_bits_ == 0 && fusion_autotune_ && plan_info_.herm_blocks == 0) {
                // the greedy plan's tile width, measured: on the second pair of
                // planes if there is one, else on scratch planes (rho untouched)
                double *sr = d_re2_, *si = d_im2_;
                const bool temp = !sr;
                if (temp) {
                    HIPTRY(hipMalloc(&sr, num_ * sizeof(double)));
                    const hipError_t e = hipMalloc(&si, num_ * sizeof(double));
                    if (e != hipSuccess) {
                        (void)hipFree(sr);
                        HIPTRY(e);
                    }
                }
                rc = plan_->autotune_tiles(n_, sr, si, stream_, &tuned_tile_bits_, plan_info_);
                if (temp) {
                    (void)hipFree(sr);
                    (void)hipFree(si);
                }
                if (rc) return rc;
            }
        } else if (wide) {
            set_error("matrix gates on 3..5 qubits need a fusable circuit (n >= 4, no SRN)");
            return DMX_ESTATE;
        }
    }
    HIPTRY(hipStreamSynchronize(stream_));
    uploaded_ = true;
    return DMX_OK;
}

bool Engine::herm_wanted() const
{
    if (!fusion_permute_ || fusion_herm_ == 0) return false;
    return n_ >= (fusion_herm_ == 2 ? kHermMinQubits : kHermAutoQubits);
}

bool Engine::permuting_pays(const PlanInfo& perm, const PlanInfo& inplace) const
{
    if (perm.predicted_us > 0 && inplace.predicted_us > 0) {  // the cost model's estimates
        const double share = perm.fused_blocks ? (double)perm.perm_blocks / perm.fused_blocks : 0.0;
        return perm.predicted_us * (1.0 + permute_overhead(n_) * share) < inplace.predicted_us;
    }
    return dmx::permuting_pays(n_, perm.fused_blocks + perm.fused5_blocks + perm.plain_blocks,
                               perm.perm_blocks,
                               inplace.fused_blocks + inplace.fused5_blocks + inplace.plain_blocks);
}

int Engine::run()
{
    if (!uploaded_) {
        set_error("run() before upload()");
        return DMX_ESTATE;
    }
    const int rc = method_ == DMX_METHOD_FUSED && plan_ ? run_fused()
                                                        : run_plain(method_ != DMX_METHOD_PLAIN);
    if (rc == DMX_OK && report_) {
        // the reference's per-run report (src/dmsim_amdgpu_omp.hpp:498-503,
        // on by default through PRINT_MEA_PER_CIRCUIT); one GPU: no
        // communication time
        const double mb = (double)device_bytes() / 1024.0 / 1024.0;
        printf("\n============== DM-Sim ===============\n");
        printf("nqubits:%d, ngates:%u, ngpus:%d, comp:%.3lf ms, comm:%.3lf ms, sim:%.3lf ms, "
               "mem:%.3lf MB, mem_per_gpu:%.3lf MB\n",
               n_, n_gates(), 1, last_ms_, 0.0, last_ms_, mb, mb);
        printf("=====================================\n");
        fflush(stdout);
    }
    return rc;
}

uint64_t Engine::device_bytes() const
{
    uint64_t b = 2 * num_ * sizeof(double) + dim_ * sizeof(double) + d_ops_cap_ * sizeof(PrimOp);
    if (d_re2_) b += 2 * num_ * sizeof(double);
    return b;
}

int Engine::run_plain(bool per_launch)
{
    const int n_ops = (int)prims_.size();
    last_launches_ = 0;
    last_sweeps_ = 0;
    last_fallbacks_ = 0;
    bool coop_ok = true;  // until a cooperative launch of this run is refused
    prof_.start(stream_);
    HIPTRY(hipEventRecord(ev0_, stream_));
    // Segments of gate primitives between channels: each segment runs its two
    // index passes (the reference's forward and backward pass); a channel then
    // acts once on both indices, in the record for the current orientation.
    int s0 = 0;
    while (s0 < n_ops) {
        int s1 = s0;
        while (s1 < n_ops && prims_[s1].kind != PK_SUP) s1++;
        if (s1 > s0) {
            // reference orientation: fast index first; transposed: slow first
            const bool first_slow = transposed_;
            for (int pass = 0; pass < 2; pass++) {
                const bool slow = pass == 0 ? first_slow : !first_slow;
                int t = s0;
                while (t < s1) {
                    if (per_launch || prims_[t].kind == PK_C2) {
                        prof_.begin(prim_class(prims_[t]), prim_bytes(prims_[t], n_));
                        HIPTRY(launch_plain_one(d_re_, d_im_, d_ops_ + t, prims_[t], n_, slow,
                                                max_blocks_, stream_));
                        prof_.end();
                        last_launches_++;
                        t++;
                        continue;
                    }
                    int end = t;
                    double bytes = 0;
                    while (end < s1 && prims_[end].kind != PK_C2)
                        bytes += prim_bytes(prims_[end++], n_);
                    hipError_t e = hipErrorCooperativeLaunchTooLarge;
                    if (coop_ok) {
                        prof_.begin(KC_PERSISTENT, bytes);
                        e = launch_plain_persistent(d_re_, d_im_, d_ops_, t, end, n_, slow,
                                                    persistent_blocks_, stream_);
                        prof_.end();
                    }
                    if (e == hipSuccess) {
                        last_launches_++;
                        t = end;
                        continue;
                    }
                    // The cooperative launch was refused (ROCm grants the
                    // grid barrier's hardware to one process on a device at
                    // a time) and did not run: the same primitives, one
                    // launch each, give the same bits.
                    (void)hipGetLastError();
                    if (coop_ok) prof_.cancel();
                    coop_ok = false;
                    last_fallbacks_++;
                    for (; t < end; t++) {
                        prof_.begin(prim_class(prims_[t]), prim_bytes(prims_[t], n_));
                        HIPTRY(launch_plain_one(d_re_, d_im_, d_ops_ + t, prims_[t], n_, slow,
                                                max_blocks_, stream_));
                        prof_.end();
                        last_launches_++;
                    }
                }
            }
            last_sweeps_ += 2u * (uint32_t)(s1 - s0);
            transposed_ = !transposed_;
        }
        if (s1 < n_ops) {  // channel: records s1 (stored orientation), s1 + 1 (transposed)
            const int pick = s1 + (transposed_ ? 1 : 0);
            prof_.begin(KC_QUAD, prim_bytes(prims_[pick], n_));
            HIPTRY(launch_plain_one(d_re_, d_im_, d_ops_ + pick, prims_[pick], n_, false,
                                    max_blocks_, stream_));
            prof_.end();
            last_launches_++;
            last_sweeps_++;
        }
        s0 = s1 + 2;
    }
    HIPTRY(hipEventRecord(ev1_, stream_));
    HIPTRY(hipEventSynchronize(ev1_));
    float ms = 0.f;
    HIPTRY(hipEventElapsedTime(&ms, ev0_, ev1_));
    last_ms_ = ms;
    HIPTRY(prof_.collect());
    return DMX_OK;
}

void Engine::drop_graph()
{
    for (RunGraph& g : graphs_) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = RunGraph{};
    }
    graph_next_ = 0;
}

int Engine::run_fused()
{
    last_launches_ = 0;
    last_sweeps_ = 0;
    last_fallbacks_ = 0;
    if (!prof_.enabled()) {
        // launch-bound at small n: one graph launch instead of one per step
        RunGraph* rg = nullptr;
        for (RunGraph& g : graphs_)
            if (g.exec && g.transposed == transposed_ && g.re == d_re_ && g.im == d_im_) rg = &g;
        if (!rg) {
            RunGraph& g = graphs_[graph_next_];
            graph_next_ ^= 1;
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g = RunGraph{};
            double *re = d_re_, *im = d_im_, *re2 = d_re2_, *im2 = d_im2_;
            uint32_t launches = 0, sweeps = 0;
            HIPTRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            const int rc = plan_->execute(re, im, re2, im2, n_, transposed_, stream_, prof_,
                                          &launches, &sweeps);
            hipGraph_t cg = nullptr;
            const hipError_t ec = hipStreamEndCapture(stream_, &cg);
            if (rc) {
                if (cg) (void)hipGraphDestroy(cg);
                return rc;
            }
            HIPTRY(ec);
            hipGraphExec_t exec = nullptr;
            const hipError_t ei = hipGraphInstantiate(&exec, cg, nullptr, nullptr, 0);
            (void)hipGraphDestroy(cg);
            HIPTRY(ei);
            g = RunGraph{exec, transposed_, d_re_, d_im_, re, im, re2, im2, launches};
            rg = &g;
        }
        HIPTRY(hipEventRecord(ev0_, stream_));
        HIPTRY(hipGraphLaunch(rg->exec, stream_));
        HIPTRY(hipEventRecord(ev1_, stream_));
        HIPTRY(hipEventSynchronize(ev1_));
        d_re_ = rg->re_end;
        d_im_ = rg->im_end;
        d_re2_ = rg->re2_end;
        d_im2_ = rg->im2_end;
        float ms = 0.f;
        HIPTRY(hipEventElapsedTime(&ms, ev0_, ev1_));
        last_ms_ = ms;
        last_launches_ = rg->launches;
        last_sweeps_ = rg->launches;
        return DMX_OK;
    }
    prof_.start(stream_);
    HIPTRY(hipEventRecord(ev0_, stream_));
    int rc = plan_->execute(d_re_, d_im_, d_re2_, d_im2_, n_, transposed_, stream_, prof_,
                            &last_launches_, &last_sweeps_);
    if (rc) return rc;
    HIPTRY(hipEventRecord(ev1_, stream_));
    HIPTRY(hipEventSynchronize(ev1_));
    float ms = 0.f;
    HIPTRY(hipEventElapsedTime(&ms, ev0_, ev1_));
    last_ms_ = ms;
    HIPTRY(prof_.collect());
    return DMX_OK;
}

int Engine::clear_circuit()
{
    circuit_.clear();
    prims_.clear();
    drop_graph();
    delete plan_;
    plan_ = nullptr;
    uploaded_ = false;
    return DMX_OK;
}

int Engine::reset_dm()
{
    HIPTRY(hipMemsetAsync(d_re_, 0, num_ * sizeof(double), stream_));
    HIPTRY(hipMemsetAsync(d_im_, 0, num_ * sizeof(double), stream_));
    const double one = 1.0;
    HIPTRY(hipMemcpyAsync(d_re_, &one, sizeof(double), hipMemcpyHostToDevice, stream_));
    HIPTRY(hipStreamSynchronize(stream_));
    transposed_ = false;
    return DMX_OK;
}

int Engine::get_diag(double* diag)
{
    HIPTRY(launch_diag(d_re_, n_, d_diag_, stream_));
    HIPTRY(hipMemcpyAsync(diag, d_diag_, dim_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPTRY(hipStreamSynchronize(stream_));
    return DMX_OK;
}

// measure(): src/dmsim_amdgpu_omp.hpp:536-564.  Sequential prefix sum of
// |Re rho_ii| (same order, so the same doubles), glibc srand/rand; the shot is
// the j with scan[j] <= r < scan[j+1], 0 if there is none.  scan is
// non-decreasing, so that j is unique and a binary search finds it.
int Engine::measure(uint32_t reps, uint32_t seed, uint32_t* out, double* total)
{
    std::vector<double> diag(dim_);
    int rc = get_diag(diag.data());
    if (rc) return rc;
    std::vector<double> scan(dim_ + 1);
    scan[0] = 0;
    for (uint64_t i = 1; i < dim_ + 1; i++) scan[i] = scan[i - 1] + std::fabs(diag[i - 1]);
    srand(seed);
    for (uint32_t s = 0; s < reps; s++) {
        const double r = (double)rand() / (double)RAND_MAX;
        // first index with scan[idx] > r
        const uint64_t idx =
            (uint64_t)(std::upper_bound(scan.begin(), scan.end(), r) - scan.begin());
        uint32_t res = 0;
        if (idx >= 1 && idx <= dim_) res = (uint32_t)(idx - 1);
        out[s] = res;
    }
    if (std::fabs(scan[dim_] - 1.0) > 1e-3)  // ERROR_BAR, src/config.hpp:53
        printf("Sum of probability along diag is far from 1.0 with %lf\n", scan[dim_]);
    if (total) *total = scan[dim_];
    return DMX_OK;
}

int Engine::to_reference_orientation()
{
    if (!transposed_) return DMX_OK;
    HIPTRY(launch_conj_transpose(d_re_, d_im_, n_, stream_));
    transposed_ = false;
    return DMX_OK;
}

int Engine::get_dm(double* re, double* im)
{
    int rc = to_reference_orientation();
    if (rc) return rc;
    if (re)
        HIPTRY(hipMemcpyAsync(re, d_re_, num_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (im)
        HIPTRY(hipMemcpyAsync(im, d_im_, num_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPTRY(hipStreamSynchronize(stream_));
    return DMX_OK;
}

int Engine::get_entries(uint64_t count, const uint64_t* idx, double* re, double* im)
{
    for (uint64_t t = 0; t < count; t++)
        if (idx[t] >= num_) {
            set_error("entry index out of range (must be < 4^n)");
            return DMX_EINVAL;
        }
    if (count == 0) return DMX_OK;
    void* buf = nullptr;  // indices, then the two output planes
    HIPTRY(hipMalloc(&buf, count * (sizeof(uint64_t) + 2 * sizeof(double))));
    uint64_t* d_idx = static_cast<uint64_t*>(buf);
    double* d_re = reinterpret_cast<double*>(d_idx + count);
    double* d_im = d_re + count;
    hipError_t e = hipMemcpyAsync(d_idx, idx, count * sizeof(uint64_t), hipMemcpyHostToDevice,
                                  stream_);
    if (e == hipSuccess)
        e = launch_gather(d_re_, d_im_, n_, d_idx, count, transposed_, d_re, d_im, stream_);
    if (e == hipSuccess && re)
        e = hipMemcpyAsync(re, d_re, count * sizeof(double), hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess && im)
        e = hipMemcpyAsync(im, d_im, count * sizeof(double), hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    (void)hipFree(buf);
    HIPTRY(e);
    return DMX_OK;
}

int Engine::get_entries_noalloc(uint64_t count, const uint64_t* idx, double* re, double* im)
{
    if (count > kNoAllocMaxEntries) {
        set_error("at most " + std::to_string(kNoAllocMaxEntries) +
                  " entries per allocation-free gather");
        return DMX_EINVAL;
    }
    for (uint64_t t = 0; t < count; t++)
        if (idx[t] >= num_) {
            set_error("entry index out of range (must be < 4^n)");
            return DMX_EINVAL;
        }
    HIPTRY(gather_noalloc(d_re_, d_im_, n_, idx, count, transposed_, re, im, stream_));
    return DMX_OK;
}

int Engine::expect_z(double* out)
{
    std::vector<double> diag(dim_);
    int rc = get_diag(diag.data());
    if (rc) return rc;
    for (int q = 0; q < n_; q++) {
        double s = 0;
        for (uint64_t i = 0; i < dim_; i++) s += ((i >> q) & 1) ? -diag[i] : diag[i];
        out[q] = s;
    }
    return DMX_OK;
}

int Engine::set_method(int m)
{
    if (m != DMX_METHOD_PLAIN && m != DMX_METHOD_PLAIN_LAUNCH && m != DMX_METHOD_FUSED) {
        set_error("unknown method " + std::to_string(m));
        return DMX_EINVAL;
    }
    if (uploaded_ && ((m == DMX_METHOD_FUSED) != (method_ == DMX_METHOD_FUSED))) {
        // the uploaded table was built for the other executor
        method_ = m;
        uploaded_ = false;
        return upload();
    }
    method_ = m;
    return DMX_OK;
}

int Engine::set_tile_bits(int m)
{
    if (m != 0 && (m < 4 || m > 8)) {
        set_error("tile width must be 0 (automatic) or 4..8");
        return DMX_EINVAL;
    }
    tile_bits_ = m;
    return DMX_OK;
}

int Engine::set_fusion_max_k(uint32_t k)
{
    if (k < 1 || k > 5) {
        set_error("fusion width must be in 1..5");
        return DMX_EINVAL;
    }
    fusion_max_k_ = (int)k;
    return DMX_OK;
}

int Engine::time_launch(int kind, int m, uint32_t qmask, int target, int control, int slow,
                        int reps, double* us)
{
    if (reps < 1 || reps > 100000 || !us) {
        set_error("time_launch needs 1 <= reps <= 100000 and an output");
        return DMX_EINVAL;
    }
    const double pi = 3.14159265358979323846;
    PrimOp op;
    std::memset(&op, 0, sizeof(op));
    op.q1 = -1;
    if (kind == CK_FUSED) {
        if (n_ < 4 || __builtin_popcount(qmask) != 4 || (qmask >> n_) != 0 ||
            !tile_bits_ok(n_, m, qmask)) {
            set_error("fused timing needs n >= 4, four qubits < n and 4 <= m <= min(8, n) "
                      "(m = 4 only with qubit 0 in the block)");
            return DMX_EINVAL;
        }
    } else if (kind == CK_FUSED5) {
        if (n_ < 5 || __builtin_popcount(qmask) != 5 || (qmask >> n_) != 0 ||
            !tile_bits5_ok(n_, m, qmask)) {
            set_error("five-qubit timing needs n >= 5, five qubits < n and 5 <= m <= min(8, n) "
                      "(m = 5 only with qubit 0 in the block)");
            return DMX_EINVAL;
        }
    } else if (kind == CK_HERM) {
        if (n_ < 8 || __builtin_popcount(qmask) != 4 || (qmask >> n_) != 0 || m != 8) {
            set_error("Hermitian-sweep timing needs n >= 8, four qubits < n and m = 8");
            return DMX_EINVAL;
        }
        if (!d_re2_) {  // the sweep writes the second pair of planes
            double *re2 = nullptr, *im2 = nullptr;
            HIPTRY(hipMalloc(&re2, num_ * sizeof(double)));
            hipError_t e = hipMalloc(&im2, num_ * sizeof(double));
            if (e != hipSuccess) {
                (void)hipFree(re2);
                HIPTRY(e);
            }
            d_re2_ = re2;
            d_im2_ = im2;
        }
    } else if (kind >= CK_DENSE && kind <= CK_QUAD) {
        const bool two = kind == CK_CX || kind == CK_QUAD;
        if (target < 0 || target >= n_ || (two && (control < 0 || control >= n_ || control == target))) {
            set_error("sweep timing needs a target (and a distinct control) below n");
            return DMX_EINVAL;
        }
        op.q0 = target;
        if (kind == CK_DENSE) {  // Hadamard-like 2x2
            op.kind = PK_C1;
            const double h = 1.0 / std::sqrt(2.0);
            op.c[0] = h, op.c[2] = h, op.c[4] = h, op.c[6] = -h;
        } else if (kind == CK_DIAG) {
            op.kind = PK_U1;
            op.c[0] = std::cos(0.3);
            op.c[1] = std::sin(0.3);
        } else if (kind == CK_CX) {
            op.kind = PK_X;
            op.q1 = control;
        } else {  // 4x4 DFT / 2
            op.kind = PK_C2;
            op.q1 = control;
            for (int r = 0; r < 4; r++)
                for (int c = 0; c < 4; c++) {
                    op.c[2 * (4 * r + c)] = 0.5 * std::cos(2 * pi * r * c / 4);
                    op.c[2 * (4 * r + c) + 1] = 0.5 * std::sin(2 * pi * r * c / 4);
                }
        }
    } else {
        set_error("unknown launch kind " + std::to_string(kind));
        return DMX_EINVAL;
    }
    if (!d_probe_blk_) HIPTRY(hipMalloc(&d_probe_blk_, sizeof(FusedBlock)));
    if (!d_probe_blk5_) HIPTRY(hipMalloc(&d_probe_blk5_, sizeof(FusedBlock5)));
    if (!d_probe_op_) HIPTRY(hipMalloc(&d_probe_op_, sizeof(PrimOp)));
    if (kind == CK_FUSED5) {  // 32 x 32 DFT / sqrt(32)
        std::vector<FusedBlock5> fb(1);
        std::memset(fb.data(), 0, sizeof(FusedBlock5));
        fb[0].qmask = qmask;
        fb[0].m = (uint32_t)m;
        const double s = 1.0 / std::sqrt(32.0);
        for (int r = 0; r < 32; r++)
            for (int c = 0; c < 32; c++) {
                fb[0].ur[32 * r + c] = s * std::cos(2 * pi * r * c / 32);
                fb[0].ui[32 * r + c] = s * std::sin(2 * pi * r * c / 32);
            }
        HIPTRY(hipMemcpy(d_probe_blk5_, fb.data(), sizeof(FusedBlock5), hipMemcpyHostToDevice));
    } else if (kind == CK_FUSED || kind == CK_HERM) {  // 16 x 16 DFT / 4
        FusedBlock fb;
        std::memset(&fb, 0, sizeof(fb));
        fb.qmask = qmask;
        fb.m = (uint32_t)m;
        if (kind == CK_HERM) {
            // the output order a plan gives it: the block's qubits at bits 0..3,
            // the super-tile's four low free bits at 4..7, the rest in order;
            // the next sweep's bits at the top (so half the entries are written)
            uint32_t cm = qmask;
            for (int q = 0; q < n_ && __builtin_popcount(cm) < 8; q++) cm |= 1u << q;
            int next = 0;
            for (uint32_t b = qmask; b; b &= b - 1) fb.perm_L[__builtin_ctz(b)] = 1u << next++;
            for (uint32_t b = cm & ~qmask; b; b &= b - 1) fb.perm_L[__builtin_ctz(b)] = 1u << next++;
            for (int q = 0; q < n_; q++)
                if (!((cm >> q) & 1u)) fb.perm_L[q] = 1u << next++;
            fb.perm_on = 1;
            fb.next_q = 0xFu << (n_ - 4);
        }
        for (int r = 0; r < 16; r++)
            for (int c = 0; c < 16; c++) {
                fb.ur[16 * r + c] = 0.25 * std::cos(2 * pi * r * c / 16);
                fb.ui[16 * r + c] = 0.25 * std::sin(2 * pi * r * c / 16);
            }
        HIPTRY(hipMemcpy(d_probe_blk_, &fb, sizeof(fb), hipMemcpyHostToDevice));
    } else {
        HIPTRY(hipMemcpy(d_probe_op_, &op, sizeof(op), hipMemcpyHostToDevice));
    }
    auto launch = [&]() {
        if (kind == CK_FUSED5)
            return launch_fused5(d_re_, d_im_, d_probe_blk5_, n_, m, kFusedGridCap, stream_);
        if (kind == CK_HERM)
            return launch_fused4_herm(d_re_, d_im_, d_re2_, d_im2_, d_probe_blk_, n_,
                                      kFusedGridCap, stream_);
        return kind == CK_FUSED
                   ? launch_fused4(d_re_, d_im_, d_probe_blk_, n_, m, kFusedGridCap, stream_)
                   : launch_plain_one(d_re_, d_im_, d_probe_op_, op, n_, slow != 0, kSweepGridCap,
                                      stream_);
    };
    HIPTRY(launch());
    HIPTRY(hipEventRecord(ev0_, stream_));
    for (int r = 0; r < reps; r++) HIPTRY(launch());
    HIPTRY(hipEventRecord(ev1_, stream_));
    HIPTRY(hipEventSynchronize(ev1_));
    float ms = 0.f;
    HIPTRY(hipEventElapsedTime(&ms, ev0_, ev1_));
    *us = 1000.0 * ms / reps;
    return DMX_OK;
}

std::string Engine::dump() const
{
    std::ostringstream ss;
    for (const Gate& g : circuit_) ss << g.dump() << "\n";
    return ss.str();
}

}  // namespace dmx
