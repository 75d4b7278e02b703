// batch.cpp -- batched executor (host side).
//
// Each circuit is planned into <= 4-qubit blocks by the fused planner.  Run
// order, per wave: reset every rho of the wave; for block step s = 0, 1, ...,
// one launch per (qubit count, tile width) covering every circuit of that
// width that still has a block s (grid.y = circuit); then each circuit of
// kSoloQubits or more, whose rho alone fills the GPU, through its own fused
// plan (the single-circuit engine's: Hermitian block sweeps); then one
// workgroup per circuit sums its Pauli terms.  The sequence is captured once
// into a hipGraph and replayed by each run.
#include "batch.hpp"

#include <algorithm>
#include <map>
#include <tuple>

#include "dmsim_amd.h"
#include "engine.hpp"

namespace dmx {

#define BTRY(call)                                                                   \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            set_error(std::string(#call) + ": " + hipGetErrorString(e_));            \
            return e_ == hipErrorOutOfMemory ? DMX_ENOMEM : DMX_EHIP;                \
        }                                                                            \
    } while (0)

Batch::~Batch()
{
    free_device();
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Batch::free_device()
{
    if (stream_) (void)hipStreamSynchronize(stream_);  // nothing in flight reads them
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
    for (void* p : {(void*)pool_, (void*)pool_alt_, (void*)d_blocks_, (void*)d_entries_,
                    (void*)d_re_off_, (void*)d_terms_, (void*)d_jobs_, (void*)d_energy_,
                    (void*)d_diag_, (void*)d_diag_off_})
        if (p) (void)hipFree(p);
    d_diag_ = nullptr;
    d_diag_off_ = nullptr;
    pool_ = nullptr;
    pool_alt_ = nullptr;
    d_blocks_ = nullptr;
    d_entries_ = nullptr;
    d_re_off_ = nullptr;
    d_terms_ = nullptr;
    d_jobs_ = nullptr;
    d_energy_ = nullptr;
}

int Batch::init()
{
    BTRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    BTRY(hipEventCreate(&ev0_));
    BTRY(hipEventCreate(&ev1_));
    return DMX_OK;
}

uint64_t Batch::gates() const
{
    uint64_t g = 0;
    for (const Circ& c : circs_) g += c.orig.size();
    return g;
}

double Batch::sweep_bytes() const
{
    double b = 0;
    for (const Circ& c : circs_) b += c.bytes;
    return b;
}

int Batch::add_circuit(uint32_t n, std::vector<Gate> gates, uint32_t* id)
{
    if (uploaded_) {
        set_error("add_circuit after upload");
        return DMX_ESTATE;
    }
    if (n < 4 || n > 16) {
        set_error("batched circuits need 4 <= n_qubits <= 16");
        return DMX_EINVAL;
    }
    for (const Gate& g : gates) {
        if (g.op < 0 || g.op >= OP_MAT) {
            set_error("unknown op code " + std::to_string(g.op));
            return DMX_EINVAL;
        }
        const int k = op_arity(g.op);
        for (int i = 0; i < k; i++) {
            if ((int)g.qb[i] >= (int)n) {
                set_error("qubit index out of range in a batched circuit");
                return DMX_EINVAL;
            }
            for (int j2 = 0; j2 < i; j2++)
                if (g.qb[i] == g.qb[j2]) {
                    set_error("operands must be distinct qubits");
                    return DMX_EINVAL;
                }
        }
    }
    Circ c;
    c.n = (int)n;
    c.orig = gates;
    // batched blocks have at most four qubits: a C4X runs as the gates the
    // reference composes it from (qelib1 c4x as src/dmsim_cpu_omp.hpp:2701-2714
    // spells it, whose primitive sweeps are the C4X's own)
    for (const Gate& g : gates) {
        if (g.op != OP_C4X) {
            c.gates.push_back(g);
            continue;
        }
        const uint32_t a = g.qb[0], b = g.qb[1], cc = g.qb[2], d = g.qb[3], e = g.qb[4];
        auto one = [&](int op, uint32_t q) {
            Gate x;
            x.op = op;
            x.qb[0] = q;
            c.gates.push_back(x);
        };
        auto cu1 = [&](double lambda, uint32_t ctl, uint32_t tgt) {
            Gate x;
            x.op = OP_CU1;
            x.qb[0] = ctl;
            x.qb[1] = tgt;
            x.lambda = lambda;
            c.gates.push_back(x);
        };
        auto c3 = [&](int op, uint32_t t) {
            Gate x;
            x.op = op;
            x.qb[0] = a, x.qb[1] = b, x.qb[2] = cc, x.qb[3] = t;
            c.gates.push_back(x);
        };
        const double pi = 3.14159265358979323846;
        one(OP_H, e);
        cu1(-pi / 2, d, e);
        one(OP_H, e);
        c3(OP_C3X, d);
        one(OP_H, d);
        cu1(pi / 4, d, e);
        one(OP_H, d);
        c3(OP_C3X, d);
        c3(OP_C3SQRTX, e);
    }
    circs_.push_back(std::move(c));
    if (id) *id = (uint32_t)(circs_.size() - 1);
    return DMX_OK;
}

int Batch::set_pool_limit(uint64_t bytes)
{
    if (uploaded_) {
        set_error("set_pool_limit after upload");
        return DMX_ESTATE;
    }
    pool_limit_ = bytes;
    return DMX_OK;
}

int Batch::add_pauli(uint32_t id, const char* pauli, double weight)
{
    if (uploaded_) {
        set_error("add_pauli after upload");
        return DMX_ESTATE;
    }
    if (id >= circs_.size() || !pauli) {
        set_error("bad circuit id or null Pauli string");
        return DMX_EINVAL;
    }
    Circ& c = circs_[id];
    const std::string s(pauli);
    if ((int)s.size() != c.n) {
        set_error("Pauli string length must equal the circuit's qubit count");
        return DMX_EINVAL;
    }
    PauliTerm t{0, 0, 0, 0, weight};
    for (int q = 0; q < c.n; q++) {
        switch (s[q]) {
        case 'I': break;
        case 'X': t.xmask |= 1ull << q; break;
        case 'Y': t.xmask |= 1ull << q; t.zymask |= 1ull << q; t.ny++; break;
        case 'Z': t.zymask |= 1ull << q; break;
        default:
            set_error("Pauli string characters must be I, X, Y or Z");
            return DMX_EINVAL;
        }
    }
    c.paulis.push_back(t);
    return DMX_OK;
}

int Batch::upload()
{
    if (uploaded_) {
        set_error("batch already uploaded");
        return DMX_ESTATE;
    }
    if (circs_.empty() || circs_.size() > 65535) {
        set_error("a batch needs 1..65535 circuits");
        return DMX_EINVAL;
    }
    // a failed earlier attempt may have left device buffers and launches
    free_device();
    launches_.clear();
    kinds_.clear();
    for (Circ& c : circs_) {
        c.plan.reset();
        c.solo = c.herm = false;
        c.solo_launches = 0;
    }
    std::vector<FusedBlock> blocks;
    std::vector<PauliTerm> terms;
    std::vector<uint64_t> re_off;
    uint64_t off = 0;
    bool any_perm = false;
    std::vector<uint8_t> kinds;
    for (Circ& c : circs_) {
        c.blk_off = blocks.size();
        c.bytes = 0;
        c.herm = false;
        c.n_blocks = 0;
        if (c.n >= kSoloQubits) {
            // planned as the single-circuit engine plans it alone (fused
            // method, blocks of up to four qubits, Hermitian sweeps)
            FusedPlan* p = nullptr;
            PlanInfo pi;
            int rc = build_fused_plan(c.orig, c.n, 4, model_, 0, true, &p, &pi, true);
            if (rc) return rc;
            if (!p) {
                set_error("batched circuits must be fusable (no SRN)");
                return DMX_EINVAL;
            }
            c.plan.reset(p);
            rc = p->upload(stream_);
            if (rc) return rc;
            c.solo = true;
            any_perm = any_perm || p->needs_alt();
            // where rho ends: each out-of-place step moves it to the other pool
            std::vector<PlanStepView> steps;
            p->describe(steps);
            bool alt = false;
            for (const PlanStepView& v : steps)
                if (v.kind == PlanStepView::HERM || v.kind == PlanStepView::RELABEL ||
                    (v.kind == PlanStepView::FUSED4 && v.block->perm_on))
                    alt = !alt;
            c.final_alt = alt;
            c.bytes = p->bytes(c.n);
            continue;
        }
        const double sweep = 32.0 * (double)(1ull << (2 * c.n));
        if (hermitian_ && c.n >= kBatchHermQubits) {
            // Hermitian sweeps and relabellings, planned as the single-circuit
            // engine plans with Hermitian sweeps from 8 qubits
            FusedPlan* p = nullptr;
            PlanInfo pi;
            int rc = build_fused_plan(c.gates, c.n, 4, model_, 0, true, &p, &pi, true);
            if (rc) return rc;
            std::unique_ptr<FusedPlan> hold(p);
            std::vector<PlanStepView> steps;
            bool ok = p != nullptr;
            if (ok) {
                p->describe(steps);
                for (const PlanStepView& v : steps)
                    ok = ok && (v.kind == PlanStepView::HERM || v.kind == PlanStepView::RELABEL ||
                                v.kind == PlanStepView::FUSED4);
            }
            if (ok) {
                for (const PlanStepView& v : steps) {
                    blocks.push_back(*v.block);
                    const uint8_t k = v.kind == PlanStepView::HERM      ? SK_HERM
                                      : v.kind == PlanStepView::RELABEL ? SK_RELABEL
                                                                        : SK_INORDER;
                    kinds.push_back(k);
                    c.bytes += k == SK_HERM ? herm_bytes(c.n, v.block->next_q != 0) : sweep;
                }
                c.n_blocks = steps.size();
                c.herm = true;
                continue;
            }
            // a step that is no batched sweep (an unfused segment): in-order blocks
        }
        // the single-circuit engine's choice (same planner, same rule), so a
        // batched rho is bit-identical to the circuit's own fused run
        BlockPlan bp;
        bool fusable = false;
        int rc = plan_blocks(c.gates, c.n, 4, bp, fusable, true);
        if (rc) return rc;
        size_t permuting = 0;
        for (const FusedBlock& fb : bp.blocks) permuting += fb.perm_on ? 1 : 0;
        if (fusable && permuting > 0 && c.n < kPermuteSureQubits) {
            BlockPlan ip;
            bool f2 = false;
            rc = plan_blocks(c.gates, c.n, 4, ip, f2, false);
            if (rc) return rc;
            if (!permuting_pays(c.n, (double)(bp.blocks.size() + bp.blocks5.size()),
                                (double)permuting, (double)(ip.blocks.size() + ip.blocks5.size())))
                bp = std::move(ip);
        }
        if (!fusable) {
            set_error("batched circuits must be fusable (no SRN)");
            return DMX_EINVAL;
        }
        for (const auto& o : bp.order)
            if (o.first != 0) {
                set_error("batched circuits take gates on at most 4 qubits");
                return DMX_EINVAL;
            }
        c.n_blocks = bp.blocks.size();
        blocks.insert(blocks.end(), bp.blocks.begin(), bp.blocks.end());
        kinds.insert(kinds.end(), bp.blocks.size(), (uint8_t)SK_INORDER);
        c.bytes = (double)c.n_blocks * sweep;
    }
    kinds_ = kinds;
    for (const FusedBlock& fb : blocks) any_perm = any_perm || fb.perm_on;
    // waves: consecutive circuits whose rho fit one pool (and its twin for
    // permuting blocks); every wave reuses the pool from offset 0
    uint64_t limit = pool_limit_;
    if (limit == 0) {
        size_t fr = 0, tot = 0;
        BTRY(hipMemGetInfo(&fr, &tot));
        const uint64_t reserve = 2ull << 30;
        limit = (fr > reserve ? fr - reserve : 0) / (any_perm ? 2 : 1);
    }
    const uint64_t cap = limit / sizeof(double);
    wave_first_.clear();
    wave_count_.clear();
    wave_doubles_.clear();
    uint64_t diag_total = 0;
    for (int i = 0; i < (int)circs_.size(); i++) {
        Circ& c = circs_[i];
        const uint64_t need = 2ull << (2 * c.n);
        if (need > cap) {
            set_error("a batched circuit of " + std::to_string(c.n) +
                      " qubits does not fit the pool limit");
            return DMX_ENOMEM;
        }
        if (wave_first_.empty() || off + need > cap) {
            wave_first_.push_back(i);
            wave_count_.push_back(0);
            wave_doubles_.push_back(0);
            off = 0;
        }
        c.wave = (int)wave_first_.size() - 1;
        c.re_off = off;
        c.im_off = off + (1ull << (2 * c.n));
        off += need;
        wave_count_.back()++;
        wave_doubles_.back() = off;
        re_off.push_back(c.re_off);
        c.diag_off = diag_total;
        diag_total += 1ull << c.n;
    }
    pool_doubles_ = *std::max_element(wave_doubles_.begin(), wave_doubles_.end());
    BTRY(hipMalloc(&pool_, pool_doubles_ * sizeof(double)));
    if (any_perm) BTRY(hipMalloc(&pool_alt_, pool_doubles_ * sizeof(double)));
    if (waves() > 1) {
        std::vector<uint64_t> doff;
        for (const Circ& c : circs_) doff.push_back(c.diag_off);
        BTRY(hipMalloc(&d_diag_, diag_total * sizeof(double)));
        BTRY(hipMalloc(&d_diag_off_, doff.size() * sizeof(uint64_t)));
        BTRY(hipMemcpy(d_diag_off_, doff.data(), doff.size() * sizeof(uint64_t),
                       hipMemcpyHostToDevice));
    }
    // which pool holds each circuit's rho before each of its blocks (static,
    // so the captured graph replays the same pointers)
    auto planes = [&](const Circ& c, bool alt, double*& re, double*& im) {
        double* base = alt ? pool_alt_ : pool_;
        re = base + c.re_off;
        im = base + c.im_off;
    };
    std::vector<std::vector<char>> alt_before(circs_.size());
    for (size_t i = 0; i < circs_.size(); i++) {
        Circ& c = circs_[i];
        if (c.solo) continue;  // final_alt set with its plan
        bool alt = false;
        for (size_t b = 0; b < c.n_blocks; b++) {
            alt_before[i].push_back(alt ? 1 : 0);
            if (blocks[c.blk_off + b].perm_on) alt = !alt;
        }
        c.final_alt = alt;
    }
    auto entry = [&](int i, size_t s) {
        const Circ& c = circs_[i];
        const bool alt = alt_before[i][s] != 0;
        BatchEntry e;
        planes(c, alt, e.re, e.im);
        if (blocks[c.blk_off + s].perm_on)
            planes(c, !alt, e.re_out, e.im_out);
        else
            e.re_out = e.re, e.im_out = e.im;
        e.blk = d_blocks_ + c.blk_off + s;
        return e;
    };
    if (!blocks.empty()) BTRY(hipMalloc(&d_blocks_, blocks.size() * sizeof(FusedBlock)));
    for (const Circ& c : circs_) tile_bits_[c.n] = default_tile_bits(c.n);
    if (autotune_ && !blocks.empty()) {
        BTRY(hipMemcpy(d_blocks_, blocks.data(), blocks.size() * sizeof(FusedBlock),
                       hipMemcpyHostToDevice));
        const int rc = autotune(blocks);
        if (rc) return rc;
    }
    BTRY(hipMalloc(&d_re_off_, re_off.size() * sizeof(uint64_t)));
    BTRY(hipMalloc(&d_energy_, circs_.size() * sizeof(double)));

    // Launches, per wave: its batched circuits packed into groups whose rho
    // take at most `budget` bytes (largest circuits first; 0: the whole wave
    // is one group), and per group and block step one launch per (qubit
    // count, tile width).  A group runs all its steps before the next starts,
    // so with a budget below the Infinity Cache its rho stay in the cache
    // from one step to the next.  Returns the number of groups.
    auto build = [&](uint64_t budget, std::vector<BatchEntry>& ents, std::vector<Launch>& lst) {
        size_t n_groups = 0;
        for (int w = 0; w < (int)waves(); w++) {
            std::vector<int> ids;
            for (int i = wave_first_[w]; i < wave_first_[w] + wave_count_[w]; i++)
                if (circs_[i].n_blocks > 0) ids.push_back(i);
            if (budget > 0)
                std::stable_sort(ids.begin(), ids.end(),
                                 [&](int a, int b) { return circs_[a].n > circs_[b].n; });
            size_t k = 0;
            while (k < ids.size()) {
                std::vector<int> grp;
                uint64_t bytes = 0;
                while (k < ids.size()) {
                    const uint64_t need = 16ull << (2 * circs_[ids[k]].n);
                    if (budget > 0 && !grp.empty() && bytes + need > budget) break;
                    grp.push_back(ids[k++]);
                    bytes += need;
                }
                std::sort(grp.begin(), grp.end(), [&](int a, int b) {
                    return circs_[a].n != circs_[b].n ? circs_[a].n < circs_[b].n : a < b;
                });
                n_groups++;
                size_t steps = 0;
                for (int i : grp) steps = std::max(steps, circs_[i].n_blocks);
                for (size_t s = 0; s < steps; s++) {
                    // per (qubit count, step kind, tile width)
                    std::map<std::tuple<int, int, int>, std::vector<int>> by;
                    for (int i : grp)
                        if (s < circs_[i].n_blocks) {
                            const size_t b = circs_[i].blk_off + s;
                            const int kind = kinds_[b];
                            const int m = kind == SK_INORDER ? (int)blocks[b].m : 8;
                            by[{circs_[i].n, kind, m}].push_back(i);
                        }
                    for (const auto& kv : by) {
                        lst.push_back(Launch{std::get<0>(kv.first), std::get<2>(kv.first),
                                             (int)kv.second.size(), ents.size(), w,
                                             std::get<1>(kv.first)});
                        for (int i : kv.second) ents.push_back(entry(i, s));
                    }
                }
            }
        }
        return n_groups;
    };
    if (!group_fixed_) {
        group_bytes_ = 0;
        if (autotune_ && !blocks.empty()) {
            const int rc = autotune_groups(build);
            if (rc) return rc;
        }
    }
    std::vector<BatchEntry> entries;
    groups_ = (uint32_t)build(group_bytes_, entries, launches_);
    if (!entries.empty()) BTRY(hipMalloc(&d_entries_, entries.size() * sizeof(BatchEntry)));

    std::vector<EnergyJob> jobs;
    for (const Circ& c : circs_) {
        double *re = nullptr, *im = nullptr;
        planes(c, c.final_alt, re, im);
        jobs.push_back({re, im, c.n, (int32_t)terms.size(), (int32_t)c.paulis.size(), 0});
        terms.insert(terms.end(), c.paulis.begin(), c.paulis.end());
    }
    if (terms.empty()) terms.push_back(PauliTerm{0, 0, 0, 0, 0.0});  // keep a valid pointer
    BTRY(hipMalloc(&d_terms_, terms.size() * sizeof(PauliTerm)));
    BTRY(hipMalloc(&d_jobs_, jobs.size() * sizeof(EnergyJob)));

    if (!blocks.empty())
        BTRY(hipMemcpy(d_blocks_, blocks.data(), blocks.size() * sizeof(FusedBlock),
                       hipMemcpyHostToDevice));
    if (!entries.empty())
        BTRY(hipMemcpy(d_entries_, entries.data(), entries.size() * sizeof(BatchEntry),
                       hipMemcpyHostToDevice));
    BTRY(hipMemcpy(d_re_off_, re_off.data(), re_off.size() * sizeof(uint64_t),
                   hipMemcpyHostToDevice));
    BTRY(hipMemcpy(d_terms_, terms.data(), terms.size() * sizeof(PauliTerm),
                   hipMemcpyHostToDevice));
    BTRY(hipMemcpy(d_jobs_, jobs.data(), jobs.size() * sizeof(EnergyJob), hipMemcpyHostToDevice));

    // capture the run once; replays reset the pool themselves
    BTRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue();
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(stream_, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    BTRY(ec);
    graph_ = g;
    BTRY(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
    energy_.assign(circs_.size(), 0.0);
    uploaded_ = true;
    return DMX_OK;
}

int Batch::set_group_bytes(uint64_t bytes)
{
    if (uploaded_) {
        set_error("set_group_bytes after upload");
        return DMX_ESTATE;
    }
    group_bytes_ = bytes;
    group_fixed_ = true;
    return DMX_OK;
}

int Batch::set_autotune(bool on)
{
    if (uploaded_) {
        set_error("set_autotune after upload");
        return DMX_ESTATE;
    }
    autotune_ = on;
    return DMX_OK;
}

int Batch::tile_bits(int n) const
{
    auto it = tile_bits_.find(n);
    return it == tile_bits_.end() ? 0 : it->second;
}

// Times, for each qubit count, one batched launch of every circuit's first
// block at each admissible super-tile width (5..min(8, n): every block of the
// group must admit it, and m = 4 needs qubit 0 in the block), and gives all
// blocks of that width class the fastest.  Runs before the pool is first
// reset; the sweeps act on whatever the pool holds.
int Batch::autotune(std::vector<FusedBlock>& blocks)
{
    std::map<int, std::vector<int>> by_n;  // circuits of in-order blocks by qubit count
    for (int i = 0; i < (int)circs_.size(); i++)
        if (circs_[i].n_blocks > 0 && !circs_[i].herm) by_n[circs_[i].n].push_back(i);
    if (by_n.empty()) return DMX_OK;
    BTRY(hipMemsetAsync(pool_, 0, pool_doubles_ * sizeof(double), stream_));
    size_t most = 0;
    for (const auto& kv : by_n) most = std::max(most, kv.second.size());
    BatchEntry* d_tmp = nullptr;
    BTRY(hipMalloc(&d_tmp, most * sizeof(BatchEntry)));
    int rc = DMX_OK;
    for (const auto& kv : by_n) {
        const int n = kv.first;
        const int lo = n > 4 ? 5 : 4, hi = std::min(8, n);
        if (lo >= hi) continue;
        std::vector<BatchEntry> ents;
        for (int i : kv.second) {
            const Circ& c = circs_[i];
            BatchEntry e{pool_ + c.re_off, pool_ + c.im_off, pool_ + c.re_off, pool_ + c.im_off,
                         d_blocks_ + c.blk_off};
            if (blocks[c.blk_off].perm_on)  // its first block writes the second pool
                e.re_out = pool_alt_ + c.re_off, e.im_out = pool_alt_ + c.im_off;
            ents.push_back(e);
        }
        hipError_t e = hipMemcpy(d_tmp, ents.data(), ents.size() * sizeof(BatchEntry),
                                 hipMemcpyHostToDevice);
        float best_ms = 0.f;
        int best_m = default_tile_bits(n);
        for (int m = lo; m <= hi && e == hipSuccess; m++) {
            e = launch_fused4_batch(d_tmp, (int)ents.size(), n, m, 4096, stream_);
            if (e == hipSuccess) e = hipEventRecord(ev0_, stream_);
            for (int r = 0; r < 3 && e == hipSuccess; r++)
                e = launch_fused4_batch(d_tmp, (int)ents.size(), n, m, 4096, stream_);
            if (e == hipSuccess) e = hipEventRecord(ev1_, stream_);
            if (e == hipSuccess) e = hipEventSynchronize(ev1_);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0_, ev1_);
            if (e == hipSuccess && (m == lo || ms < best_ms)) {
                best_ms = ms;
                best_m = m;
            }
        }
        if (e != hipSuccess) {
            set_error(std::string("batch autotune: ") + hipGetErrorString(e));
            rc = DMX_EHIP;
            break;
        }
        tile_bits_[n] = best_m;
    }
    (void)hipFree(d_tmp);
    if (rc) return rc;
    for (const Circ& c : circs_)
        if (!c.herm)
            for (size_t b = 0; b < c.n_blocks; b++)
                blocks[c.blk_off + b].m = (uint32_t)tile_bits_[c.n];
    return DMX_OK;
}

// Times the batched block launches of one run (after a warm-up pass) under
// each group budget -- the whole wave as one group, and 16, 32, 64 and 128 MiB
// of rho per group (well inside the 256 MB Infinity Cache, which also holds
// the second pool of permuting blocks) -- and keeps the fastest.  The sweeps
// act on the zeroed pools; the schedule changes no circuit's arithmetic.
int Batch::autotune_groups(const LaunchBuilder& build)
{
    const uint64_t MiB = 1ull << 20;
    const uint64_t cands[] = {0, 16 * MiB, 32 * MiB, 64 * MiB, 128 * MiB};
    constexpr int kCands = 5;
    std::vector<BatchEntry> ents[kCands];
    std::vector<Launch> lsts[kCands];
    size_t most = 1;
    for (int c = 0; c < kCands; c++) {
        build(cands[c], ents[c], lsts[c]);
        most = std::max(most, ents[c].size());
    }
    BTRY(hipMemsetAsync(pool_, 0, pool_doubles_ * sizeof(double), stream_));
    if (pool_alt_) BTRY(hipMemsetAsync(pool_alt_, 0, pool_doubles_ * sizeof(double), stream_));
    BTRY(hipStreamSynchronize(stream_));
    BatchEntry* d_tmp = nullptr;
    BTRY(hipMalloc(&d_tmp, most * sizeof(BatchEntry)));
    float best = 0.f;
    int pick = 0;
    hipError_t e = hipSuccess;
    for (int c = 0; c < kCands && e == hipSuccess; c++) {
        // the previous candidate's launches have finished (event synchronised)
        e = hipMemcpy(d_tmp, ents[c].data(), ents[c].size() * sizeof(BatchEntry),
                      hipMemcpyHostToDevice);
        for (int rep = 0; rep < 2 && e == hipSuccess; rep++) {  // rep 0 warms up
            if (rep == 1) e = hipEventRecord(ev0_, stream_);
            for (const Launch& l : lsts[c])
                if (e == hipSuccess) e = launch(l, d_tmp);
        }
        if (e == hipSuccess) e = hipEventRecord(ev1_, stream_);
        if (e == hipSuccess) e = hipEventSynchronize(ev1_);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0_, ev1_);
        if (e == hipSuccess && (c == 0 || ms < best)) {
            best = ms;
            pick = c;
        }
    }
    (void)hipFree(d_tmp);
    if (e != hipSuccess) {
        set_error(std::string("batch schedule timing: ") + hipGetErrorString(e));
        return DMX_EHIP;
    }
    group_bytes_ = cands[pick];
    return DMX_OK;
}

hipError_t Batch::launch(const Launch& l, const BatchEntry* d_entries)
{
    const BatchEntry* e = d_entries + l.entry_off;
    switch (l.kind) {
    case SK_HERM: return launch_fused4_herm_batch(e, l.count, l.n, herm_batch_blocks(l.n), stream_);
    case SK_RELABEL: return launch_relabel_batch(e, l.count, l.n, 4096, stream_);
    default: return launch_fused4_batch(e, l.count, l.n, l.m, 4096, stream_);
    }
}

int Batch::set_hermitian(bool on)
{
    if (uploaded_) {
        set_error("set_hermitian after upload");
        return DMX_ESTATE;
    }
    hermitian_ = on;
    return DMX_OK;
}

int Batch::enqueue()
{
    for (int w = 0; w < (int)waves(); w++) {
        const int first = wave_first_[w], count = wave_count_[w];
        BTRY(launch_pool_reset(pool_, wave_doubles_[w], d_re_off_ + first, count, stream_));
        for (const Launch& l : launches_)
            if (l.wave == w) BTRY(launch(l, d_entries_));
        for (int i = first; i < first + count; i++) {
            Circ& c = circs_[i];
            if (!c.solo) continue;
            double* re = pool_ + c.re_off;
            double* im = pool_ + c.im_off;
            double* re2 = pool_alt_ ? pool_alt_ + c.re_off : nullptr;
            double* im2 = pool_alt_ ? pool_alt_ + c.im_off : nullptr;
            uint32_t la = 0, sw = 0;
            const int rc = c.plan->execute(re, im, re2, im2, c.n, false, stream_, quiet_, &la, &sw);
            if (rc) return rc;
            if (re != (c.final_alt ? pool_alt_ : pool_) + c.re_off) {
                set_error("internal: a batched circuit's rho ends in the other pool");
                return DMX_EINVAL;
            }
            c.solo_launches = la;
        }
        BTRY(launch_pauli_energy(d_jobs_ + first, d_terms_, count, d_energy_ + first, stream_));
        if (waves() > 1)  // the next wave reuses this pool space: keep the diagonals
            BTRY(launch_batch_diag(d_jobs_ + first, d_diag_off_ + first, count, d_diag_, stream_));
    }
    return DMX_OK;
}

int Batch::run()
{
    if (!uploaded_) {
        set_error("batch run before upload");
        return DMX_ESTATE;
    }
    BTRY(hipEventRecord(ev0_, stream_));
    if (use_graph_) {
        BTRY(hipGraphLaunch(exec_, stream_));
    } else {
        const int rc = enqueue();
        if (rc) return rc;
    }
    BTRY(hipEventRecord(ev1_, stream_));
    BTRY(hipEventSynchronize(ev1_));
    float ms = 0.f;
    BTRY(hipEventElapsedTime(&ms, ev0_, ev1_));
    last_ms_ = ms;
    BTRY(hipMemcpy(energy_.data(), d_energy_, energy_.size() * sizeof(double),
                   hipMemcpyDeviceToHost));
    ran_ = true;
    return DMX_OK;
}

int Batch::energies(double* out) const
{
    if (!ran_) {
        set_error("batch energies before run");
        return DMX_ESTATE;
    }
    std::copy(energy_.begin(), energy_.end(), out);
    return DMX_OK;
}

int Batch::get_dm(uint32_t id, double* re, double* im)
{
    if (!ran_ || id >= circs_.size()) {
        set_error(!ran_ ? "batch get_dm before run" : "bad circuit id");
        return !ran_ ? DMX_ESTATE : DMX_EINVAL;
    }
    const Circ& c = circs_[id];
    if (c.wave != (int)waves() - 1) {
        set_error("rho of a circuit of an earlier wave is not kept (its pool space is reused); "
                  "its diagonal is: dmx_batch_get_diag");
        return DMX_ESTATE;
    }
    const size_t bytes = (1ull << (2 * c.n)) * sizeof(double);
    const double* base = c.final_alt ? pool_alt_ : pool_;
    if (re) BTRY(hipMemcpy(re, base + c.re_off, bytes, hipMemcpyDeviceToHost));
    if (im) BTRY(hipMemcpy(im, base + c.im_off, bytes, hipMemcpyDeviceToHost));
    return DMX_OK;
}

int Batch::get_diag(uint32_t id, double* out)
{
    if (!ran_ || id >= circs_.size() || !out) {
        set_error(!ran_ ? "batch get_diag before run" : "bad circuit id or null output");
        return !ran_ ? DMX_ESTATE : DMX_EINVAL;
    }
    const Circ& c = circs_[id];
    const uint64_t dim = 1ull << c.n;
    if (waves() > 1) {
        BTRY(hipMemcpy(out, d_diag_ + c.diag_off, dim * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        const double* base = (c.final_alt ? pool_alt_ : pool_) + c.re_off;
        BTRY(hipMemcpy2D(out, sizeof(double), base, (dim + 1) * sizeof(double), sizeof(double),
                         dim, hipMemcpyDeviceToHost));
    }
    return DMX_OK;
}

}  // namespace dmx
