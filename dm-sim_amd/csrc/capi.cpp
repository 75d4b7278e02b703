// capi.cpp -- extern "C" entry points of libdmsim_amd.so (include/dmsim_amd.h).
#include <cstring>
#include <exception>
#include <new>
#include <set>
#include <string>
#include <vector>

#include "dmsim_amd.h"
#include "batch.hpp"
#include "engine.hpp"
#include "fused.hpp"
#include "mlp.hpp"
#include "qasm.hpp"

struct dmx_cost;

struct dmx_sim {
    dmx::Engine eng;
    dmx_cost* model = nullptr;  // the planner model set on this simulation
};

// A model remembers the simulations that plan with it, so destroying it
// first detaches it from them (their later uploads, including the one a
// method switch makes, plan greedily instead of reading freed memory).
struct dmx_batch;

struct dmx_cost {
    dmx::CostModel m;
    std::set<dmx_sim*> users;
    std::set<dmx_batch*> batches;
};

static_assert(DMX_COST_FEATURES == dmx::kMlpFeatures, "feature count");
static_assert(DMX_COST_PARAMS == dmx::kMlpParams, "parameter count");

struct dmx_batch {
    dmx::Batch b;
    dmx_cost* model = nullptr;  // the planner model set on this batch
};

namespace {

template <class F>
int guarded(F&& f)
{
    try {
        return f();
    } catch (const std::bad_alloc&) {
        dmx::set_error("host allocation failed");
        return DMX_ENOMEM;
    } catch (const std::exception& e) {
        dmx::set_error(e.what());
        return DMX_EINVAL;
    }
}

int null_handle()
{
    dmx::set_error("null simulation handle");
    return DMX_EINVAL;
}

}  // namespace

extern "C" {

int dmx_create(uint32_t n_qubits, uint32_t n_gpus, dmx_sim** out)
{
    if (!out) {
        dmx::set_error("null output pointer");
        return DMX_EINVAL;
    }
    *out = nullptr;
    return guarded([&] {
        dmx_sim* s = new dmx_sim();
        int rc = s->eng.init(n_qubits, n_gpus);
        if (rc) {
            delete s;
            return rc;
        }
        *out = s;
        return DMX_OK;
    });
}

int dmx_destroy(dmx_sim* sim)
{
    if (sim && sim->model) sim->model->users.erase(sim);
    delete sim;
    return DMX_OK;
}

int dmx_append(dmx_sim* sim, int32_t op, const uint32_t qb[5], double theta, double phi,
               double lambda)
{
    if (!sim) return null_handle();
    if (!qb) {
        dmx::set_error("null qubit array");
        return DMX_EINVAL;
    }
    if (op < 0 || op >= DMX_OP_COUNT) {
        dmx::set_error("unknown op code " + std::to_string(op));
        return DMX_EINVAL;
    }
    return guarded([&] {
        dmx::Gate g;
        g.op = op;
        std::memcpy(g.qb, qb, sizeof(g.qb));
        g.theta = theta;
        g.phi = phi;
        g.lambda = lambda;
        return sim->eng.append(g);
    });
}

int dmx_append_matrix(dmx_sim* sim, uint32_t k, const uint32_t* qubits, const double* mat_re,
                      const double* mat_im)
{
    if (!sim) return null_handle();
    if (k < 1 || k > 5 || !qubits || !mat_re || !mat_im) {
        dmx::set_error("matrix gate needs 1 <= k <= 5 and non-null arrays");
        return DMX_EINVAL;
    }
    return guarded([&] {
        dmx::Gate g;
        g.op = dmx::OP_MAT;
        g.k = (int)k;
        for (uint32_t i = 0; i < k; i++) g.qb[i] = qubits[i];
        const size_t cnt = (size_t)1 << (2 * k);
        g.mre.assign(mat_re, mat_re + cnt);
        g.mim.assign(mat_im, mat_im + cnt);
        return sim->eng.append(g);
    });
}

int dmx_append_kraus(dmx_sim* sim, uint32_t qubit, uint32_t n_ops, const double* k_re,
                     const double* k_im)
{
    if (!sim) return null_handle();
    if (n_ops < 1 || n_ops > 64 || !k_re || !k_im) {
        dmx::set_error("a channel needs 1..64 Kraus operators and non-null arrays");
        return DMX_EINVAL;
    }
    return guarded([&] {
        // stored A = conj(rho):  A' = sum_i conj(K_i) A K_i^T, so on the entry quad
        // (row bit a, column bit b) S[(a,b),(a',b')] = sum_i conj(K_i[a][a']) K_i[b][b'];
        // for the conj-transposed orientation S'[(a,b),(a',b')] = conj(S[(b,a),(b',a')]).
        double sr[16] = {}, si[16] = {};
        for (uint32_t i = 0; i < n_ops; i++) {
            const double* kr = k_re + 4 * i;
            const double* ki = k_im + 4 * i;
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++)
                    for (int a2 = 0; a2 < 2; a2++)
                        for (int b2 = 0; b2 < 2; b2++) {
                            const double xr = kr[2 * a + a2], xi = -ki[2 * a + a2];  // conj
                            const double yr = kr[2 * b + b2], yi = ki[2 * b + b2];
                            const int idx = (2 * a + b) * 4 + (2 * a2 + b2);
                            sr[idx] += xr * yr - xi * yi;
                            si[idx] += xr * yi + xi * yr;
                        }
        }
        dmx::Gate g;
        g.op = dmx::OP_KRAUS;
        g.k = 1;
        g.qb[0] = qubit;
        g.mre.assign(32, 0.0);
        g.mim.assign(32, 0.0);
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                for (int a2 = 0; a2 < 2; a2++)
                    for (int b2 = 0; b2 < 2; b2++) {
                        const int idx = (2 * a + b) * 4 + (2 * a2 + b2);
                        const int tdx = (2 * b + a) * 4 + (2 * b2 + a2);
                        g.mre[idx] = sr[idx];
                        g.mim[idx] = si[idx];
                        g.mre[16 + idx] = sr[tdx];
                        g.mim[16 + idx] = -si[tdx];
                    }
        return sim->eng.append(g);
    });
}

int dmx_upload(dmx_sim* sim)
{
    if (!sim) return null_handle();
    return guarded([&] { return sim->eng.upload(); });
}

int dmx_run(dmx_sim* sim)
{
    if (!sim) return null_handle();
    return guarded([&] { return sim->eng.run(); });
}

int dmx_clear_circuit(dmx_sim* sim)
{
    if (!sim) return null_handle();
    return guarded([&] { return sim->eng.clear_circuit(); });
}

int dmx_reset(dmx_sim* sim)
{
    if (!sim) return null_handle();
    return guarded([&] {
        int rc = sim->eng.clear_circuit();
        return rc ? rc : sim->eng.reset_dm();
    });
}

int dmx_reset_dm(dmx_sim* sim)
{
    if (!sim) return null_handle();
    return guarded([&] { return sim->eng.reset_dm(); });
}

int dmx_measure(dmx_sim* sim, uint32_t repetition, uint32_t seed, uint32_t* out,
                double* prob_sum)
{
    if (!sim) return null_handle();
    if (repetition > 0 && !out) {
        dmx::set_error("null output array");
        return DMX_EINVAL;
    }
    return guarded([&] { return sim->eng.measure(repetition, seed, out, prob_sum); });
}

int dmx_get_dm(dmx_sim* sim, double* re, double* im)
{
    if (!sim) return null_handle();
    return guarded([&] { return sim->eng.get_dm(re, im); });
}

int dmx_get_entries(dmx_sim* sim, uint64_t count, const uint64_t* idx, double* re, double* im)
{
    if (!sim) return null_handle();
    if (count && !idx) {
        dmx::set_error("null index array");
        return DMX_EINVAL;
    }
    return guarded([&] { return sim->eng.get_entries(count, idx, re, im); });
}

int dmx_get_entries_noalloc(dmx_sim* sim, uint64_t count, const uint64_t* idx, double* re,
                            double* im)
{
    if (!sim) return null_handle();
    if (count && !idx) {
        dmx::set_error("null index array");
        return DMX_EINVAL;
    }
    return guarded([&] { return sim->eng.get_entries_noalloc(count, idx, re, im); });
}

int dmx_get_diag(dmx_sim* sim, double* diag)
{
    if (!sim) return null_handle();
    if (!diag) {
        dmx::set_error("null output array");
        return DMX_EINVAL;
    }
    return guarded([&] { return sim->eng.get_diag(diag); });
}

int dmx_expect_z(dmx_sim* sim, double* out)
{
    if (!sim) return null_handle();
    if (!out) {
        dmx::set_error("null output array");
        return DMX_EINVAL;
    }
    return guarded([&] { return sim->eng.expect_z(out); });
}

int dmx_info(const dmx_sim* sim, uint32_t* n_qubits, uint64_t* dim, uint32_t* n_gates)
{
    if (!sim) return null_handle();
    if (n_qubits) *n_qubits = sim->eng.n_qubits();
    if (dim) *dim = sim->eng.dim();
    if (n_gates) *n_gates = sim->eng.n_gates();
    return DMX_OK;
}

int dmx_dump(const dmx_sim* sim, char* buf, size_t cap, size_t* needed)
{
    if (!sim) return null_handle();
    return guarded([&] {
        const std::string s = sim->eng.dump();
        if (needed) *needed = s.size() + 1;
        if (buf && cap > 0) {
            const size_t m = s.size() < cap - 1 ? s.size() : cap - 1;
            std::memcpy(buf, s.data(), m);
            buf[m] = '\0';
        }
        return DMX_OK;
    });
}

int dmx_set_method(dmx_sim* sim, int method)
{
    if (!sim) return null_handle();
    return guarded([&] { return sim->eng.set_method(method); });
}

int dmx_set_fusion_max_k(dmx_sim* sim, uint32_t k)
{
    if (!sim) return null_handle();
    return guarded([&] { return sim->eng.set_fusion_max_k(k); });
}

int dmx_set_fusion_permute(dmx_sim* sim, uint32_t on)
{
    if (!sim) return null_handle();
    if (on > 1) {
        dmx::set_error("fusion permute flag must be 0 or 1");
        return DMX_EINVAL;
    }
    return guarded([&] { return sim->eng.set_fusion_permute(on != 0); });
}

int dmx_set_fusion_autotune(dmx_sim* sim, uint32_t on)
{
    if (!sim) return null_handle();
    if (on > 1) {
        dmx::set_error("fusion autotune flag must be 0 or 1");
        return DMX_EINVAL;
    }
    return guarded([&] { return sim->eng.set_fusion_autotune(on != 0); });
}

int dmx_tuned_tile_bits(const dmx_sim* sim, uint32_t* m)
{
    if (!sim) return null_handle();
    if (m) *m = (uint32_t)sim->eng.tuned_tile_bits();
    return DMX_OK;
}

int dmx_set_fusion_hermitian(dmx_sim* sim, uint32_t mode)
{
    if (!sim) return null_handle();
    if (mode > 2) {
        dmx::set_error("fusion Hermitian mode must be 0, 1 or 2");
        return DMX_EINVAL;
    }
    return guarded([&] { return sim->eng.set_fusion_hermitian((int)mode); });
}

int dmx_last_run_stats(const dmx_sim* sim, double* ms, uint32_t* launches, uint32_t* sweeps)
{
    if (!sim) return null_handle();
    if (ms) *ms = sim->eng.last_ms();
    if (launches) *launches = sim->eng.last_launches();
    if (sweeps) *sweeps = sim->eng.last_sweeps();
    return DMX_OK;
}

int dmx_set_report(dmx_sim* sim, int on)
{
    if (!sim) return null_handle();
    sim->eng.set_report(on != 0);
    return DMX_OK;
}

int dmx_last_run_fallbacks(const dmx_sim* sim, uint32_t* segments)
{
    if (!sim) return null_handle();
    if (segments) *segments = sim->eng.last_fallbacks();
    return DMX_OK;
}

int dmx_set_profiling(dmx_sim* sim, int on)
{
    if (!sim) return null_handle();
    sim->eng.profiler().enable(on != 0);
    return DMX_OK;
}

int dmx_profile_get(dmx_sim* sim, uint32_t cls, char* name, size_t cap, uint32_t* calls,
                    double* ms, double* bytes, uint32_t* count)
{
    if (!sim) return null_handle();
    if (count) *count = dmx::KC_COUNT;
    if (cls >= (uint32_t)dmx::KC_COUNT) {
        dmx::set_error("kernel class index out of range");
        return DMX_EINVAL;
    }
    const dmx::KernelStat& st = sim->eng.profiler().stat((int)cls);
    if (name && cap > 0) {
        std::strncpy(name, dmx::kKernelClassNames[cls], cap - 1);
        name[cap - 1] = '\0';
    }
    if (calls) *calls = st.calls;
    if (ms) *ms = st.ms;
    if (bytes) *bytes = st.bytes;
    return DMX_OK;
}

int dmx_qasm_parse(const char* source, uint32_t cap, int32_t* ops, uint32_t* qb, double* params,
                   uint32_t* n_qubits, uint32_t* n_gates)
{
    if (!source || (cap && (!ops || !qb || !params))) {
        dmx::set_error("null source or output arrays");
        return DMX_EINVAL;
    }
    return guarded([&] {
        dmx::QasmProgram prog;
        int rc = dmx::parse_qasm(source, prog);
        if (rc) return rc;
        const size_t m = prog.gates.size() < cap ? prog.gates.size() : cap;
        for (size_t i = 0; i < m; i++) {
            const dmx::Gate& g = prog.gates[i];
            ops[i] = g.op;
            for (int j = 0; j < 5; j++) qb[5 * i + j] = g.qb[j];
            params[3 * i] = g.theta;
            params[3 * i + 1] = g.phi;
            params[3 * i + 2] = g.lambda;
        }
        if (n_qubits) *n_qubits = (uint32_t)prog.n_qubits;
        if (n_gates) *n_gates = (uint32_t)prog.gates.size();
        return DMX_OK;
    });
}

int dmx_append_qasm(dmx_sim* sim, const char* source, uint32_t* n_gates)
{
    if (!sim) return null_handle();
    if (!source) {
        dmx::set_error("null source");
        return DMX_EINVAL;
    }
    return guarded([&] {
        dmx::QasmProgram prog;
        int rc = dmx::parse_qasm(source, prog);
        if (rc) return rc;
        if ((uint32_t)prog.n_qubits > sim->eng.n_qubits()) {
            dmx::set_error("the program declares " + std::to_string(prog.n_qubits) +
                           " qubits; the simulation has " +
                           std::to_string(sim->eng.n_qubits()));
            return DMX_EINVAL;
        }
        for (const dmx::Gate& g : prog.gates) {
            rc = sim->eng.append(g);
            if (rc) return rc;
        }
        if (n_gates) *n_gates = (uint32_t)prog.gates.size();
        return DMX_OK;
    });
}

int dmx_cost_create(uint32_t seed, dmx_cost** out)
{
    if (!out) {
        dmx::set_error("null output pointer");
        return DMX_EINVAL;
    }
    *out = nullptr;
    return guarded([&] {
        dmx_cost* c = new dmx_cost();
        int rc = c->m.init(seed);
        if (rc) {
            delete c;
            return rc;
        }
        *out = c;
        return DMX_OK;
    });
}

int dmx_cost_destroy(dmx_cost* model)
{
    if (model) {
        for (dmx_sim* s : model->users) {
            s->eng.set_cost_model(nullptr);
            s->model = nullptr;
        }
        for (dmx_batch* b : model->batches) {
            b->b.set_cost_model(nullptr);
            b->model = nullptr;
        }
    }
    delete model;
    return DMX_OK;
}

int dmx_cost_features(uint32_t n_qubits, int32_t kind, int32_t m, uint32_t qmask, int32_t target,
                      int32_t control, int32_t slow, float* out)
{
    if (!out || n_qubits < 1 || n_qubits > 16 || kind < 0 || kind > DMX_COST_HERM) {
        dmx::set_error("cost features need an output, 1 <= n <= 16 and a known kind");
        return DMX_EINVAL;
    }
    const int n = (int)n_qubits;
    if (kind == DMX_COST_FUSED) {
        if (n < 4 || __builtin_popcount(qmask) != 4 || (qmask >> n) != 0 ||
            !dmx::tile_bits_ok(n, m, qmask)) {
            dmx::set_error("fused features need n >= 4, four qubits < n and 4 <= m <= min(8, n) "
                           "(m = 4 only with qubit 0 in the block)");
            return DMX_EINVAL;
        }
    } else if (kind == DMX_COST_FUSED5) {
        if (n < 5 || __builtin_popcount(qmask) != 5 || (qmask >> n) != 0 ||
            !dmx::tile_bits5_ok(n, m, qmask)) {
            dmx::set_error("five-qubit features need n >= 5, five qubits < n and "
                           "5 <= m <= min(8, n) (m = 5 only with qubit 0 in the block)");
            return DMX_EINVAL;
        }
    } else if (kind == DMX_COST_HERM) {
        if (n < 8 || __builtin_popcount(qmask) != 4 || (qmask >> n) != 0 || m != 8) {
            dmx::set_error("Hermitian-sweep features need n >= 8, four qubits < n and m = 8");
            return DMX_EINVAL;
        }
    } else {
        const bool two = kind == DMX_COST_CX || kind == DMX_COST_QUAD;
        if (target < 0 || target >= n ||
            (two && (control < 0 || control >= n || control == target))) {
            dmx::set_error("sweep features need a target (and a distinct control) below n");
            return DMX_EINVAL;
        }
    }
    dmx::cost_features(n, kind, m, qmask, target, control, slow, out);
    return DMX_OK;
}

int dmx_cost_train(dmx_cost* model, uint32_t count, const float* X, const float* y,
                   uint32_t epochs, uint32_t batch, float lr, uint32_t seed, float* final_loss)
{
    if (!model || !X || !y) {
        dmx::set_error("null model or data");
        return DMX_EINVAL;
    }
    if (count > (1u << 24) || epochs > 100000 || batch > (1u << 24)) {
        dmx::set_error("training sizes out of range");
        return DMX_EINVAL;
    }
    return guarded([&] {
        return model->m.train((int)count, X, y, (int)epochs, (int)batch, lr, seed, final_loss);
    });
}

int dmx_cost_predict(dmx_cost* model, uint32_t count, const float* X, float* out)
{
    if (!model || (count && (!X || !out)) || count > (1u << 24)) {
        dmx::set_error("null model or data, or too many samples");
        return DMX_EINVAL;
    }
    return guarded([&] { return model->m.predict((int)count, X, out); });
}

int dmx_cost_gradient(dmx_cost* model, uint32_t count, const float* X, const float* y,
                      float* out)
{
    if (!model || !X || !y || !out || count > (1u << 24)) {
        dmx::set_error("null model or data, or too many samples");
        return DMX_EINVAL;
    }
    return guarded([&] { return model->m.gradient((int)count, X, y, out); });
}

int dmx_cost_get_params(const dmx_cost* model, float* params, float* mean, float* dev)
{
    if (!model) {
        dmx::set_error("null model");
        return DMX_EINVAL;
    }
    if (params) model->m.get_params(params);
    float mu = 0.f, sigma = 1.f;
    model->m.target_scale(&mu, &sigma);
    if (mean) *mean = mu;
    if (dev) *dev = sigma;
    return DMX_OK;
}

int dmx_cost_set_params(dmx_cost* model, const float* params, float mean, float dev)
{
    if (!model || !params || !(dev > 0.f)) {
        dmx::set_error("null model or parameters, or deviation <= 0");
        return DMX_EINVAL;
    }
    model->m.set_target_scale(mean, dev);
    return guarded([&] { return model->m.set_params(params); });
}

int dmx_set_cost_model(dmx_sim* sim, dmx_cost* model)
{
    if (!sim) return null_handle();
    if (sim->model) sim->model->users.erase(sim);
    sim->model = model;
    if (model) model->users.insert(sim);
    sim->eng.set_cost_model(model ? &model->m : nullptr);
    return DMX_OK;
}

int dmx_set_tile_bits(dmx_sim* sim, uint32_t m)
{
    if (!sim) return null_handle();
    return sim->eng.set_tile_bits((int)m);
}

int dmx_plan_info(const dmx_sim* sim, uint32_t* fused_blocks, uint32_t* plain_blocks,
                  uint32_t* fused5_blocks, double* predicted_us, double* greedy_us, uint32_t* tiles)
{
    if (!sim) return null_handle();
    const dmx::PlanInfo& pi = sim->eng.plan_info();
    if (fused_blocks) *fused_blocks = pi.fused_blocks;
    if (plain_blocks) *plain_blocks = pi.plain_blocks;
    if (fused5_blocks) *fused5_blocks = pi.fused5_blocks;
    if (predicted_us) *predicted_us = pi.predicted_us;
    if (greedy_us) *greedy_us = pi.greedy_us;
    if (tiles)
        for (int m = 4; m <= 8; m++) tiles[m - 4] = pi.tile_hist[m];
    return DMX_OK;
}

int dmx_plan_sweeps(const dmx_sim* sim, uint32_t* perm_blocks, uint32_t* herm_blocks,
                    uint32_t* relabels, uint32_t* grouping)
{
    if (!sim) return null_handle();
    const dmx::PlanInfo& pi = sim->eng.plan_info();
    if (perm_blocks) *perm_blocks = pi.perm_blocks;
    if (herm_blocks) *herm_blocks = pi.herm_blocks;
    if (relabels) *relabels = pi.relabels;
    if (grouping) *grouping = pi.grouping;
    return DMX_OK;
}

int dmx_time_launch(dmx_sim* sim, int32_t kind, int32_t m, uint32_t qmask, int32_t target,
                    int32_t control, int32_t slow, uint32_t reps, double* us)
{
    if (!sim) return null_handle();
    return guarded([&] {
        return sim->eng.time_launch(kind, m, qmask, target, control, slow, (int)reps, us);
    });
}

int dmx_batch_create(dmx_batch** out)
{
    if (!out) {
        dmx::set_error("null output pointer");
        return DMX_EINVAL;
    }
    *out = nullptr;
    return guarded([&] {
        dmx_batch* b = new dmx_batch();
        int rc = b->b.init();
        if (rc) {
            delete b;
            return rc;
        }
        *out = b;
        return DMX_OK;
    });
}

int dmx_batch_destroy(dmx_batch* b)
{
    if (b && b->model) b->model->batches.erase(b);
    delete b;
    return DMX_OK;
}

int dmx_batch_schedule(const dmx_batch* b, uint64_t* group_bytes, uint32_t* groups)
{
    if (!b) return null_handle();
    if (group_bytes) *group_bytes = b->b.group_bytes();
    if (groups) *groups = b->b.groups();
    return DMX_OK;
}

int dmx_batch_set_hermitian(dmx_batch* b, int on)
{
    if (!b) return null_handle();
    return b->b.set_hermitian(on != 0);
}

int dmx_batch_set_group_bytes(dmx_batch* b, uint64_t bytes)
{
    if (!b) return null_handle();
    return b->b.set_group_bytes(bytes);
}

int dmx_batch_set_cost_model(dmx_batch* b, dmx_cost* model)
{
    if (!b) return null_handle();
    if (b->model) b->model->batches.erase(b);
    b->model = model;
    if (model) model->batches.insert(b);
    b->b.set_cost_model(model ? &model->m : nullptr);
    return DMX_OK;
}

int dmx_batch_add_circuit(dmx_batch* b, uint32_t n_qubits, uint32_t n_gates, const int32_t* ops,
                          const uint32_t* qb, const double* params, uint32_t* id)
{
    if (!b) return null_handle();
    if (n_gates > 0 && (!ops || !qb || !params)) {
        dmx::set_error("null gate table");
        return DMX_EINVAL;
    }
    return guarded([&] {
        std::vector<dmx::Gate> gates(n_gates);
        for (uint32_t i = 0; i < n_gates; i++) {
            gates[i].op = ops[i];
            std::memcpy(gates[i].qb, qb + 5 * (size_t)i, sizeof(gates[i].qb));
            gates[i].theta = params[3 * (size_t)i];
            gates[i].phi = params[3 * (size_t)i + 1];
            gates[i].lambda = params[3 * (size_t)i + 2];
        }
        return b->b.add_circuit(n_qubits, std::move(gates), id);
    });
}

int dmx_batch_add_pauli(dmx_batch* b, uint32_t id, const char* pauli, double weight)
{
    if (!b) return null_handle();
    return guarded([&] { return b->b.add_pauli(id, pauli, weight); });
}

int dmx_batch_upload(dmx_batch* b)
{
    if (!b) return null_handle();
    return guarded([&] { return b->b.upload(); });
}

int dmx_batch_set_graph(dmx_batch* b, int on)
{
    if (!b) return null_handle();
    b->b.set_graph(on != 0);
    return DMX_OK;
}

int dmx_batch_set_autotune(dmx_batch* b, int on)
{
    if (!b) return null_handle();
    return b->b.set_autotune(on != 0);
}

int dmx_batch_tile_bits(const dmx_batch* b, uint32_t n_qubits, uint32_t* m)
{
    if (!b || !m) {
        dmx::set_error("null batch handle or output");
        return DMX_EINVAL;
    }
    *m = (uint32_t)b->b.tile_bits((int)n_qubits);
    return DMX_OK;
}

int dmx_batch_run(dmx_batch* b)
{
    if (!b) return null_handle();
    return guarded([&] { return b->b.run(); });
}

int dmx_batch_energies(dmx_batch* b, double* out)
{
    if (!b) return null_handle();
    if (!out) {
        dmx::set_error("null output array");
        return DMX_EINVAL;
    }
    return guarded([&] { return b->b.energies(out); });
}

int dmx_batch_get_dm(dmx_batch* b, uint32_t id, double* re, double* im)
{
    if (!b) return null_handle();
    return guarded([&] { return b->b.get_dm(id, re, im); });
}

int dmx_batch_stats(const dmx_batch* b, double* ms, uint32_t* launches, uint32_t* circuits,
                    uint64_t* gates)
{
    if (!b) return null_handle();
    if (ms) *ms = b->b.last_ms();
    if (launches) *launches = b->b.launches();
    if (circuits) *circuits = b->b.circuits();
    if (gates) *gates = b->b.gates();
    return DMX_OK;
}

int dmx_batch_get_diag(dmx_batch* b, uint32_t id, double* diag)
{
    if (!b) return null_handle();
    return guarded([&] { return b->b.get_diag(id, diag); });
}

int dmx_batch_set_pool_limit(dmx_batch* b, uint64_t bytes)
{
    if (!b) return null_handle();
    return guarded([&] { return b->b.set_pool_limit(bytes); });
}

int dmx_batch_waves(const dmx_batch* b, uint32_t* waves)
{
    if (!b) return null_handle();
    if (waves) *waves = b->b.waves();
    return DMX_OK;
}

int dmx_batch_sweep_bytes(const dmx_batch* b, double* bytes)
{
    if (!b) return null_handle();
    if (bytes) *bytes = b->b.sweep_bytes();
    return DMX_OK;
}

const char* dmx_last_error(void) { return dmx::get_error(); }

}  // extern "C"
