// capi.cpp -- extern "C" entry points of libdmsim_amd.so (include/dmsim_amd.h).
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "dmsim_amd.h"
#include "batch.hpp"
#include "engine.hpp"

struct dmx_sim {
    dmx::Engine eng;
};

struct dmx_batch {
    dmx::Batch b;
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

int dmx_last_run_stats(const dmx_sim* sim, double* ms, uint32_t* launches, uint32_t* sweeps)
{
    if (!sim) return null_handle();
    if (ms) *ms = sim->eng.last_ms();
    if (launches) *launches = sim->eng.last_launches();
    if (sweeps) *sweeps = sim->eng.last_sweeps();
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
    delete b;
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

const char* dmx_last_error(void) { return dmx::get_error(); }

}  // extern "C"
