// oracle/ref_harness.cpp -- C-ABI shim over the reference's own CPU path.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the unmodified
// reference sources where they lie (src/config.hpp, src/util_cpu.h,
// src/dmsim_cpu_omp.hpp) into oracle/_ref/libdmsim_ref.so, only in a
// container that has the reference tree.  tests/golden/gen_golden.py uses it
// to write the golden vectors; the pinning tests use it, when present, to
// compare the oracle restatement with the reference bit for bit.
//
// Two knobs are changed without editing the reference: the measurement seed
// (RAND_SEED is a macro expanded inside measure(), src/dmsim_cpu_omp.hpp:499,
// and config.hpp is include-guarded, so redefining it after the first include
// takes effect) and the per-run timing print (PRINT_MEA_PER_CIRCUIT).
#include "config.hpp"
#undef RAND_SEED
#undef PRINT_MEA_PER_CIRCUIT
static unsigned g_ref_seed = 0;
#define RAND_SEED g_ref_seed
#include "util_cpu.h"
#include "dmsim_cpu_omp.hpp"

using namespace DMSim;

namespace {

// Apply one gate of the table to the fast index of every row of (re, im).
// Ops 0..37 go through the reference's own function table; 38/39 (C1/C2
// matrices, not reachable through enum OP) call C1_GATE/C2_GATE directly.
void ref_apply(const Simulation* sim, ValType* re, ValType* im, int op, const unsigned* qb,
               const double* prm, const double* coef)
{
    if (op == 38) {
        C1_GATE(sim, re, im, coef[0], coef[1], coef[2], coef[3], coef[4], coef[5], coef[6],
                coef[7], qb[0]);
    } else if (op == 39) {
        const double* e = coef;
        C2_GATE(sim, re, im, e[0], e[1], e[2], e[3], e[4], e[5], e[6], e[7], e[8], e[9], e[10],
                e[11], e[12], e[13], e[14], e[15], e[16], e[17], e[18], e[19], e[20], e[21],
                e[22], e[23], e[24], e[25], e[26], e[27], e[28], e[29], e[30], e[31], qb[0],
                qb[1]);
    } else {
        Gate g((enum OP)op, qb[0], qb[1], qb[2], qb[3], qb[4], prm[0], prm[1], prm[2]);
        Gate* local = g.upload(0);
        local->exe_op(const_cast<Simulation*>(sim), re, im);
        free(local);
    }
}

}  // namespace

extern "C" {

// One Simulation(n, n_threads) + append all + upload + sim() [+ measure].
// Only ops 0..37.  Copies dm_real_res/dm_imag_res out.
int ref_sim(int n, int n_threads, int ng, const int* ops, const unsigned* qb, const double* prm,
            double* out_re, double* out_im, unsigned seed, unsigned reps, unsigned* shots)
{
    if (n < 1 || n > 14) return -1;
    Simulation sim((IdxType)n, (IdxType)n_threads);
    for (int i = 0; i < ng; i++) {
        if (ops[i] < 0 || ops[i] > 37) return -2;
        Gate g((enum OP)ops[i], qb[5 * i], qb[5 * i + 1], qb[5 * i + 2], qb[5 * i + 3],
               qb[5 * i + 4], prm[3 * i], prm[3 * i + 1], prm[3 * i + 2]);
        sim.append(&g);
    }
    sim.upload();
    sim.sim();
    const size_t num = (size_t)1 << (2 * n);
    memcpy(out_re, sim.dm_real_res, num * sizeof(double));
    memcpy(out_im, sim.dm_imag_res, num * sizeof(double));
    if (reps > 0) {
        g_ref_seed = seed;
        IdxType* r = sim.measure(reps);
        for (unsigned i = 0; i < reps; i++) shots[i] = r[i];
        delete[] r;
    }
    return 0;
}

// Hand-driven single run from |0><0| that also accepts C1/C2 (ops 38/39):
// forward pass, the reference's block_transpose, backward pass
// (the single-thread branch of simulation_kernel, src/dmsim_cpu_omp.hpp:888-906).
int ref_sim_ext(int n, int ng, const int* ops, const unsigned* qb, const double* prm,
                const double* coef, double* out_re, double* out_im)
{
    if (n < 1 || n > 14) return -1;
    Simulation sim((IdxType)n, 1);
    sim.upload();
    Simulation* s = sim.sim_cpu[0];
    for (int i = 0; i < ng; i++)
        ref_apply(s, s->dm_real, s->dm_imag, ops[i], qb + 5 * i, prm + 3 * i, coef + 32 * i);
    block_transpose(s, s->dm_real_buf, s->dm_imag_buf, s->dm_real, s->dm_imag);
    for (int i = 0; i < ng; i++)
        ref_apply(s, s->dm_real_buf, s->dm_imag_buf, ops[i], qb + 5 * i, prm + 3 * i,
                  coef + 32 * i);
    const size_t num = (size_t)1 << (2 * n);
    memcpy(out_re, s->dm_real_buf, num * sizeof(double));
    memcpy(out_im, s->dm_imag_buf, num * sizeof(double));
    return 0;
}

// The matrix M that one gate applies to each row in the forward pass:
// rows of the identity go in, row r comes out as M e_r, i.e. the buffer holds
// M^T; it is returned as M (row-major, 2^n x 2^n).
int ref_gate_matrix(int n, int op, const unsigned* qb, const double* prm, const double* coef,
                    double* out_re, double* out_im)
{
    if (n < 1 || n > 10) return -1;
    Simulation sim((IdxType)n, 1);
    sim.upload();
    Simulation* s = sim.sim_cpu[0];
    const size_t dim = (size_t)1 << n;
    memset(s->dm_real, 0, dim * dim * sizeof(double));
    memset(s->dm_imag, 0, dim * dim * sizeof(double));
    for (size_t i = 0; i < dim; i++) s->dm_real[i * dim + i] = 1.0;
    ref_apply(s, s->dm_real, s->dm_imag, op, qb, prm, coef);
    for (size_t r = 0; r < dim; r++)
        for (size_t c = 0; c < dim; c++) {
            out_re[r * dim + c] = s->dm_real[c * dim + r];
            out_im[r * dim + c] = s->dm_imag[c * dim + r];
        }
    return 0;
}

}  // extern "C"
