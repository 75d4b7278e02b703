/*
 * dmsim_amd.h -- C ABI of the MI355X density-matrix engine (libdmsim_amd.so).
 *
 * Drop-in boundary for the single-GPU path of pnnl/DM-Sim: every entry point
 * below replaces one member of the reference's `DMSim::Simulation`
 * (src/dmsim_amdgpu_omp.hpp) or of its pybind11 module
 * (src/py_amdgpu_omp_wrapper.cpp:29-88); the comment on each names the
 * reference interface (file:line) it stands for.  Plain pointers and sizes
 * only.  Every function returns DMX_OK (0) or a negative DMX_E* code; the
 * message of the last failure on the calling thread is dmx_last_error().
 * The reference asserts or exit()s in the same situations.
 *
 * State convention (identical to the reference's dm_real_res/dm_imag_res):
 * dmx_get_dm() returns re[i*dim+j] = Re(rho_ji), im[i*dim+j] = Im(rho_ji),
 * i.e. conj(rho) in row-major order, dim = 2^n_qubits; qubit q is bit q of a
 * basis index.
 *
 * A handle is not thread-safe (neither is the reference's object).
 */
#ifndef DMSIM_AMD_H
#define DMSIM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMX_OK 0
#define DMX_EINVAL (-1)  /* bad argument: qubit out of range, n_gpus != 1, ... */
#define DMX_ESTATE (-2)  /* call out of order: run() before upload(), ...      */
#define DMX_EHIP (-3)    /* HIP runtime error                                   */
#define DMX_ENOMEM (-4)  /* device or host allocation failed                    */

/* Op codes: the reference's `enum OP` in order (src/dmsim_amdgpu_omp.hpp:51-57). */
enum dmx_op {
    DMX_U3 = 0, DMX_U2, DMX_U1, DMX_CX, DMX_ID, DMX_X, DMX_Y, DMX_Z, DMX_H, DMX_S,
    DMX_SDG, DMX_T, DMX_TDG, DMX_RX, DMX_RY, DMX_RZ, DMX_CZ, DMX_CY, DMX_SWAP, DMX_CH,
    DMX_CCX, DMX_CSWAP, DMX_CRX, DMX_CRY, DMX_CRZ, DMX_CU1, DMX_CU3, DMX_RXX, DMX_RZZ,
    DMX_RCCX, DMX_RC3X, DMX_C3X, DMX_C3SQRTX, DMX_C4X, DMX_R, DMX_SRN, DMX_W, DMX_RYY,
    DMX_OP_COUNT
};

/* Execution methods (dmx_set_method). */
#define DMX_METHOD_PLAIN 0        /* persistent cooperative gate loop, 2 sweeps per primitive */
#define DMX_METHOD_PLAIN_LAUNCH 1 /* same sweeps, one launch per primitive                    */
#define DMX_METHOD_FUSED 2        /* gates fused into dense 4- and 5-qubit blocks, one sweep each */

typedef struct dmx_sim dmx_sim;

/* Simulation::Simulation(n_qubits, n_gpus), src/dmsim_amdgpu_omp.hpp:211-286.
 * Allocates rho on the current HIP device and sets it to |0..0><0..0|.
 * n_gpus must be 1 (one GPU per process); 1 <= n_qubits <= 16. */
int dmx_create(uint32_t n_qubits, uint32_t n_gpus, dmx_sim** out);

/* Simulation::~Simulation, src/dmsim_amdgpu_omp.hpp:288-316. NULL is a no-op. */
int dmx_destroy(dmx_sim* sim);

/* Simulation::append(Gate*), src/dmsim_amdgpu_omp.hpp:346-359, with the Gate
 * record of src/dmsim_amdgpu_omp.hpp:111-118 passed by value: op, qb[5]
 * (operands first; unused slots ignored), theta, phi, lambda as the static
 * factories place them (:594-782).  Operands must be < n_qubits and
 * distinct. */
int dmx_append(dmx_sim* sim, int32_t op, const uint32_t qb[5], double theta, double phi,
               double lambda);

/* Extension (no OP reaches it in the reference): an arbitrary k-qubit matrix,
 * k = 1..5, row-major 2^k x 2^k split planes, qubits[0] the most significant
 * bit of the matrix index.  k = 1 and k = 2 are the reference's C1_GATE and
 * C2_GATE (src/dmsim_amdgpu_omp.hpp:1019-1137; C2's qubit1 is qubits[0]).
 * k >= 3 runs only under DMX_METHOD_FUSED. */
int dmx_append_matrix(dmx_sim* sim, uint32_t k, const uint32_t* qubits, const double* mat_re,
                      const double* mat_im);

/* Extension with no counterpart in the reference (the DM-Sim paper's Kraus
 * form E(rho) = sum_i K_i rho K_i^dagger, which the v2.5 code does not
 * implement): a single-qubit noise channel given by n_ops 2x2 Kraus operators
 * (k_re/k_im[4*i..4*i+3], row-major).  It acts once on both indices of rho as
 * a 4x4 superoperator sweep; in the fused method it is a barrier no gate is
 * moved across.  Trace preservation (sum_i K_i^dagger K_i = I) is the
 * caller's to ensure. */
int dmx_append_kraus(dmx_sim* sim, uint32_t qubit, uint32_t n_ops, const double* k_re,
                     const double* k_im);

/* Simulation::upload(), src/dmsim_amdgpu_omp.hpp:360-392: one packed copy of
 * the lowered gate table to the device.  A second upload() without
 * clear_circuit() in between is DMX_ESTATE (the reference asserts). */
int dmx_upload(dmx_sim* sim);

/* Simulation::sim() (Python: run), src/dmsim_amdgpu_omp.hpp:405-509: applies
 * the uploaded circuit to the current state, rho <- U rho U^dagger. */
int dmx_run(dmx_sim* sim);

/* Simulation::clear_circuit(), src/dmsim_amdgpu_omp.hpp:511-535: drops the
 * gates, keeps rho (the next run continues from it). */
int dmx_clear_circuit(dmx_sim* sim);

/* Simulation::reset(), src/dmsim_amdgpu_omp.hpp:318-322: clear_circuit() +
 * reset_dm(). */
int dmx_reset(dmx_sim* sim);

/* Simulation::reset_dm(), src/dmsim_amdgpu_omp.hpp:323-344: rho <- |0><0|. */
int dmx_reset_dm(dmx_sim* sim);

/* Simulation::measure(repetition), src/dmsim_amdgpu_omp.hpp:536-564: samples
 * `repetition` basis states from |Re rho_ii| with glibc srand(seed)/rand()
 * (the reference seeds with time(0), src/config.hpp:47); out[repetition].
 * *prob_sum (may be NULL) receives the sum of |Re rho_ii|. */
int dmx_measure(dmx_sim* sim, uint32_t repetition, uint32_t seed, uint32_t* out,
                double* prob_sum);

/* The public result arrays dm_real_res / dm_imag_res
 * (src/dmsim_amdgpu_omp.hpp:473-474), copied on demand into caller buffers of
 * 4^n doubles each.  Either pointer may be NULL. */
int dmx_get_dm(dmx_sim* sim, double* re, double* im);

/* Extension: selected entries of dm_real_res / dm_imag_res without copying
 * the whole matrix (32 GiB per plane, 64 GiB for rho, at n = 16): re[t], im[t] =
 * entry idx[t] (row-major linear index < 4^n, same convention as dmx_get_dm).
 * re or im may be NULL. */
int dmx_get_entries(dmx_sim* sim, uint64_t count, const uint64_t* idx, double* re, double* im);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): the same
 * gather as dmx_get_entries, whose per-call device staging (indices and both
 * outputs) is here a fixed array that each GPU the library runs on holds
 * until the process ends; it calls no hipMalloc/hipFree.  At most 65536
 * entries per call (1.5 MiB of staging), DMX_EINVAL beyond that.  Calls on
 * one GPU share the array: calls that overlap in time on one GPU race. */
int dmx_get_entries_noalloc(dmx_sim* sim, uint64_t count, const uint64_t* idx, double* re,
                            double* im);

/* Re rho_ii, i < 2^n (the diagonal measure() and print_res_sv() read,
 * src/dmsim_amdgpu_omp.hpp:541-543, :565-575). */
int dmx_get_diag(dmx_sim* sim, double* diag);

/* Extension: <Z_q> = sum_i (-1)^{bit q of i} Re rho_ii for q < n; out[n]. */
int dmx_expect_z(dmx_sim* sim, double* out);

/* Public members n_qubits, dim, n_gates (src/dmsim_amdgpu_omp.hpp:786-800). */
int dmx_info(const dmx_sim* sim, uint32_t* n_qubits, uint64_t* dim, uint32_t* n_gates);

/* Simulation::dump(), src/dmsim_amdgpu_omp.hpp:394-402: the circuit as text,
 * one "OP(qb0,qb1,qb2,qb3,qb4,theta,phi,lambda);" line per gate.  Writes at
 * most cap bytes including the terminating NUL; *needed gets the full size. */
int dmx_dump(const dmx_sim* sim, char* buf, size_t cap, size_t* needed);

/* Execution method for subsequent run() calls (default DMX_METHOD_PLAIN). */
int dmx_set_method(dmx_sim* sim, int method);

/* Fusion limit for DMX_METHOD_FUSED: blocks grow to at most k qubits (1..5,
 * default 4); a gate on five qubits always forms a five-qubit block. */
int dmx_set_fusion_max_k(dmx_sim* sim, uint32_t k);

/* Extension with no counterpart in the reference: whether four-qubit blocks of
 * DMX_METHOD_FUSED may also take the X, CX and SWAP gates that follow them as
 * a permutation of their output (on = 1, default; below 12 qubits only when
 * the plan still wins after the permuting sweeps' extra set-up, which matters
 * where a sweep is latency-bound).  Such sweeps run out of place and need a
 * second pair of planes (16*4^n bytes, allocated at the first upload that uses
 * them, or the plan does without); off = 0 keeps every sweep in place.  Used
 * from the next dmx_upload. */
int dmx_set_fusion_permute(dmx_sim* sim, uint32_t on);

/* Extension: with permuting blocks on, run four-qubit blocks as Hermitian
 * sweeps (rho = rho^dagger throughout: only the tiles on and above the
 * diagonal are read and transformed, each result is written with its
 * conjugate transpose; rho's index bits are relabelled between blocks and
 * restored at the end of the run).  mode 0 = off, 1 = from 12 qubits
 * (default), 2 = from 8 qubits.  Out of place: uses the second pair of
 * planes.  Used from the next dmx_upload. */
int dmx_set_fusion_hermitian(dmx_sim* sim, uint32_t mode);

/* Extension: the greedy plan's tile width measured (on = 1, default): without
 * a cost model and without Hermitian sweeps, dmx_upload times up to eight of
 * the plan's block sweeps at every tile width on scratch planes (rho is not
 * touched) and runs every block at the fastest width where it is admissible;
 * off = 0 keeps the default width min(8, n).  dmx_tuned_tile_bits reports the
 * width the last upload chose (0: none timed). */
int dmx_set_fusion_autotune(dmx_sim* sim, uint32_t on);
int dmx_tuned_tile_bits(const dmx_sim* sim, uint32_t* m);

/* Device time of the last run() in milliseconds (HIP events on the run's
 * stream), number of kernel launches it made, and number of full-matrix
 * sweeps it performed. */
int dmx_last_run_stats(const dmx_sim* sim, double* ms, uint32_t* launches, uint32_t* sweeps);

/* The reference's per-run report (src/dmsim_amdgpu_omp.hpp:498-503, printed
 * after every sim() when PRINT_MEA_PER_CIRCUIT is defined, as it is by
 * default, src/config.hpp:25): on = 1 prints the same banner and line
 * ("nqubits:.., ngates:.., ngpus:.., comp:.. ms, comm:.. ms, sim:.. ms,
 * mem:.. MB, mem_per_gpu:.. MB") to stdout after every dmx_run; on = 0 (the
 * C ABI's default) is silent.  The C++ and Python faces turn it on. */
int dmx_set_report(dmx_sim* sim, int on);

/* Extension: how many persistent-loop segments of the last DMX_METHOD_PLAIN
 * run fell back to one launch per primitive (bit-identical results) because
 * the device refused the cooperative launch, as ROCm does while another
 * process holds the device's grid-barrier resource.  The reference's
 * hipLaunchCooperativeKernel call (src/dmsim_amdgpu_omp.hpp:420-433) exits
 * the process in that case. */
int dmx_last_run_fallbacks(const dmx_sim* sim, uint32_t* segments);

/* Extension: time every launch of subsequent runs with a pair of HIP events on
 * the run's stream (on = 1), or stop (on = 0). */
int dmx_set_profiling(dmx_sim* sim, int on);

/* Extension: per-kernel-class statistics of the last profiled run.  Class
 * indices run from 0 to *count - 1 (count may be NULL); name receives the
 * class name (at most cap bytes with the NUL), calls the launches, ms their
 * summed event time and bytes their summed algorithmic bytes (entries each
 * launch must read and write, both planes). */
int dmx_profile_get(dmx_sim* sim, uint32_t cls, char* name, size_t cap, uint32_t* calls,
                    double* ms, double* bytes, uint32_t* count);

/* ---- OpenQASM 2.0 front-end.  Replaces the reference's translator
 * tool/dmsim_qasm.py (Python 2; it writes a Python script of Simulation
 * factory calls): registers are numbered in declaration order on one qubit
 * index space (:266-272), qelib1 gate names map to the factories of the same
 * name with their parameter placement (:29-69), a statement on whole registers
 * is broadcast over their bits (:133-160), user-defined gates are expanded at
 * each call (:240-266); measure, barrier, creg, reset and if statements are
 * accepted and ignored (:105).  Parse errors are DMX_EINVAL with the line in
 * dmx_last_error(). ---- */

/* Parses `source` into a gate table: the first min(cap, *n_gates) gates go to
 * ops[i], qb[5*i..5*i+4], params[3*i..3*i+2] (theta, phi, lambda as in
 * dmx_append); *n_qubits and *n_gates receive the program's totals.  Arrays
 * may be NULL when cap is 0.  Host only. */
int dmx_qasm_parse(const char* source, uint32_t cap, int32_t* ops, uint32_t* qb, double* params,
                   uint32_t* n_qubits, uint32_t* n_gates);

/* Parses `source` and appends its gates to the circuit (as dmx_append would);
 * the program's qubit count must not exceed the simulation's.  *n_gates (may
 * be NULL) receives the number appended. */
int dmx_append_qasm(dmx_sim* sim, const char* source, uint32_t* n_gates);

/* ---- Learned fusion cost model.  Extension with no counterpart in the
 * reference (its planner-free design applies every gate as two sweeps): an
 * FP32 multilayer perceptron, DMX_COST_FEATURES -> 64 -> 64 -> 1 with ReLU,
 * trained and evaluated by HIP kernels, that predicts ln(microseconds) of one
 * kernel launch.  The fused planner prices every block as a fused sweep at
 * each tile width and as its gates' primitive sweeps, and keeps the cheapest;
 * without a model it fuses every block at the default width (greedy). ---- */
#define DMX_COST_FEATURES 14
#define DMX_COST_PARAMS 5185

/* Launch kinds the model prices. */
enum dmx_cost_kind {
    DMX_COST_FUSED = 0, /* dense 4-qubit block, both indices, one sweep */
    DMX_COST_DENSE = 1, /* 2x2 primitive rewriting both entries, one index */
    DMX_COST_DIAG = 2,  /* diagonal primitive (only the bit-set entry changes) */
    DMX_COST_CX = 3,    /* controlled swap */
    DMX_COST_QUAD = 4,  /* 4x4 primitive */
    DMX_COST_FUSED5 = 5, /* dense 5-qubit block, both indices, one sweep */
    DMX_COST_HERM = 6    /* dense 4-qubit block as a Hermitian sweep (m = 8, n >= 8): the
                            tiles on and above the diagonal, out of place, half writes */
};

typedef struct dmx_cost dmx_cost;

int dmx_cost_create(uint32_t seed, dmx_cost** out);
int dmx_cost_destroy(dmx_cost* model);

/* Feature vector (DMX_COST_FEATURES floats) of one launch on n qubits.  fused:
 * m = super-tile column bits (4..min(8, n); five-qubit blocks 5..min(8, n)),
 * qmask = the block's four (five) qubits;
 * sweeps: target (4x4: high qubit), control (cx control, 4x4 low qubit, else
 * -1), slow = 1 for the slow (row) index.  Host only. */
int dmx_cost_features(uint32_t n_qubits, int32_t kind, int32_t m, uint32_t qmask, int32_t target,
                      int32_t control, int32_t slow, float* out);

/* Fits the model to targets y[count] (ln microseconds) from features
 * X[count][DMX_COST_FEATURES]: targets are standardised by their mean and
 * deviation, then `epochs` passes of shuffled minibatches of `batch` with Adam
 * at learning rate lr.  final_loss (may be NULL) = mean squared error on the
 * standardised targets after training. */
int dmx_cost_train(dmx_cost* model, uint32_t count, const float* X, const float* y,
                   uint32_t epochs, uint32_t batch, float lr, uint32_t seed, float* final_loss);

/* out[count] = predictions in the targets' scale. */
int dmx_cost_predict(dmx_cost* model, uint32_t count, const float* X, float* out);

/* out[DMX_COST_PARAMS] = gradient of mean((f(X) - y)^2) over the parameters
 * of the raw network (no target standardisation). */
int dmx_cost_gradient(dmx_cost* model, uint32_t count, const float* X, const float* y,
                      float* out);

/* Parameters (DMX_COST_PARAMS floats: W1[64][F], b1[64], W2[64][64], b2[64],
 * w3[64], b3) and the target standardisation (mean, deviation). */
int dmx_cost_get_params(const dmx_cost* model, float* params, float* mean, float* dev);
int dmx_cost_set_params(dmx_cost* model, const float* params, float mean, float dev);

/* Planner model of a simulation (NULL = greedy plan); used by every later
 * dmx_upload with DMX_METHOD_FUSED, including the one dmx_set_method makes
 * when it switches between the fused and the plain executors.  Destroying the
 * model detaches it from every simulation it was set on (they plan greedily
 * from then on). */
int dmx_set_cost_model(dmx_sim* sim, dmx_cost* model);

/* Fixed super-tile width m (4..8) for every fused block that admits it
 * (m <= n; m = 4 only for blocks holding qubit 0), or 0 = automatic (the cost
 * model's choice, else min(8, n)).  Used from the next dmx_upload; overrides
 * the model's width choice and its fused-or-plain choice. */
int dmx_set_tile_bits(dmx_sim* sim, uint32_t m);

/* What the last upload planned: four-qubit blocks run fused, four-qubit
 * blocks run as primitive sweeps (the model's choice), five-qubit blocks run
 * fused (a five-qubit block the model splits counts as its four-qubit
 * sub-blocks instead), the model's predicted device time of the blocks'
 * launches and its prediction for the same blocks fused at the default width
 * (microseconds; 0 without a model), and tiles[5] = fused four-qubit blocks
 * per tile width m = 4..8 (may be NULL). */
int dmx_plan_info(const dmx_sim* sim, uint32_t* fused_blocks, uint32_t* plain_blocks,
                  uint32_t* fused5_blocks, double* predicted_us, double* greedy_us,
                  uint32_t* tiles);

/* Extension: sweeps of the uploaded fused plan by kind -- four-qubit blocks
 * with a trailing permutation, Hermitian block sweeps, relabelling sweeps --
 * and the grouping the planner chose (0 least growth, 1 frontier, 2 frontier
 * scored by the gates' predicted unfused times, 3 by those of the multi-qubit
 * gates only, 4 by their square roots, 5 structural scores with one-qubit
 * gates weighing four times a multi-qubit gate; with a cost model the one it
 * prices cheapest, else the one with fewer sweeps of 0 and 1). */
int dmx_plan_sweeps(const dmx_sim* sim, uint32_t* perm_blocks, uint32_t* herm_blocks,
                    uint32_t* relabels, uint32_t* grouping);

/* Timing-data generator: average device time (microseconds) of reps
 * back-to-back launches of one kernel configuration (arguments as in
 * dmx_cost_features) on this simulation's rho, after one untimed launch.
 * rho is left transformed: reset it before simulating. */
int dmx_time_launch(dmx_sim* sim, int32_t kind, int32_t m, uint32_t qmask, int32_t target,
                    int32_t control, int32_t slow, uint32_t reps, double* us);

/* ---- Batched executor.  Extension with no counterpart in the reference: many
 * independent circuits of uneven width (4..14 qubits) and depth on one GPU,
 * every rho resident in one device pool, each run a single hipGraph launch
 * (reset, fused block sweeps, Pauli-string energies). ---- */
typedef struct dmx_batch dmx_batch;

int dmx_batch_create(dmx_batch** out);
int dmx_batch_destroy(dmx_batch* b);

/* One circuit as a gate table (ops 0..37 as in dmx_append; qb[5*i..], params
 * [3*i..] = theta, phi, lambda); *id receives its index. */
int dmx_batch_add_circuit(dmx_batch* b, uint32_t n_qubits, uint32_t n_gates, const int32_t* ops,
                          const uint32_t* qb, const double* params, uint32_t* id);

/* A Hamiltonian term of circuit id: pauli = n characters from "IXYZ", qubit 0
 * first, with a real weight. */
int dmx_batch_add_pauli(dmx_batch* b, uint32_t id, const char* pauli, double weight);

/* Plans every circuit, allocates the pool and captures the run into a graph. */
int dmx_batch_upload(dmx_batch* b);

/* Replay the captured graph (on = 1, default) or enqueue the launches eagerly. */
int dmx_batch_set_graph(dmx_batch* b, int on);

/* Tile-width autotuning at upload (on = 1, default): for each qubit count in
 * the batch, the super-tile width (5..min(8, n)) whose batched block launch is
 * fastest on this GPU; on = 0 keeps min(8, n).  Call before dmx_batch_upload. */
int dmx_batch_set_autotune(dmx_batch* b, int on);

/* The super-tile width the batch uses for circuits of n qubits (0: none). */
int dmx_batch_tile_bits(const dmx_batch* b, uint32_t n_qubits, uint32_t* m);

/* Every circuit from |0..0><0..0|; energies become available. */
int dmx_batch_run(dmx_batch* b);

/* E_c = sum of weight * Re Tr(P rho_c) per circuit, out[circuit count]. */
int dmx_batch_energies(dmx_batch* b, double* out);

/* rho of circuit id after the last run, dmx_get_dm's convention (4^n doubles each). */
int dmx_batch_get_dm(dmx_batch* b, uint32_t id, double* re, double* im);

/* Device time of the last run, launches per run, circuits, total API gates. */
int dmx_batch_stats(const dmx_batch* b, double* ms, uint32_t* launches, uint32_t* circuits,
                    uint64_t* gates);

/* Extension: Re rho_ii of circuit id after a run, diag[2^n] (kept for every
 * circuit, also when the batch runs in several waves). */
int dmx_batch_get_diag(dmx_batch* b, uint32_t id, double* diag);

/* Extension: device bytes one pool of circuit states may take (0, the
 * default: the free device memory at upload less 2 GiB, halved when permuting
 * blocks need a second pool).  Circuits that do not fit together run in
 * waves -- consecutive circuits whose states fit, one wave after another
 * inside every run, each reusing the pool -- and dmx_batch_get_dm then serves
 * only the last wave's circuits (DMX_ESTATE for the others; their diagonals
 * and energies are kept).  Call before dmx_batch_upload; dmx_batch_waves
 * reports the number of waves after it.  Circuits take 4..16 qubits.  Below
 * 12 qubits a C4X runs as the gates qelib1 composes it from (batched blocks
 * have at most four qubits); a circuit of 12 or more qubits, whose state
 * alone fills the GPU, runs on its own inside its wave, planned as dmx_upload
 * plans it alone with the fused method at its defaults (Hermitian block
 * sweeps, five-qubit gates kept whole), with the batch's cost model if one is
 * set. */
int dmx_batch_set_pool_limit(dmx_batch* b, uint64_t bytes);
int dmx_batch_waves(const dmx_batch* b, uint32_t* waves);

/* Extension: algorithmic bytes of one batched run's sweeps (read and write
 * of both planes, 32 * 4^n bytes, per batched block; for a circuit run on its
 * own, what its plan's launches move), set at dmx_batch_upload. */
int dmx_batch_sweep_bytes(const dmx_batch* b, double* bytes);

/* Extension: the schedule dmx_batch_upload chose.  Within each wave the
 * batched circuits (below 12 qubits) are packed, largest first, into groups
 * whose states take at most *group_bytes bytes; each group runs all its block
 * steps before the next, so its states stay in the 256 MB Infinity Cache
 * between steps.  With autotuning on (dmx_batch_set_autotune, the default)
 * the budget is the fastest of the whole wave (0) and 16, 32, 64 and 128 MiB,
 * timed at upload; off, 0.  *groups = the number of groups. */
int dmx_batch_schedule(const dmx_batch* b, uint64_t* group_bytes, uint32_t* groups);

/* Extension: fixes the group budget (bytes of state per group, 0 = each wave
 * one group) instead of timing it; before dmx_batch_upload.  The schedule
 * changes no circuit's arithmetic: results are bit-identical under any. */
int dmx_batch_set_group_bytes(dmx_batch* b, uint64_t bytes);

/* Extension: Hermitian block sweeps for the batch's circuits of 10 and 11
 * qubits (on, the default; below 10 in-order blocks are faster in a batch):
 * their blocks run as batched Hermitian sweeps (the tiles
 * on and above the diagonal, each result written with its conjugate
 * transpose) and relabellings, planned as dmx_upload plans the circuit alone
 * with dmx_set_fusion_hermitian(s, 2); off (0): in-order blocks.  Before
 * dmx_batch_upload. */
int dmx_batch_set_hermitian(dmx_batch* b, int on);

/* Extension: the planner network for the batch's circuits of 12 or more
 * qubits (those that run on their own; null: the greedy planner), used at
 * dmx_batch_upload.  Destroying the model detaches it. */
int dmx_batch_set_cost_model(dmx_batch* b, dmx_cost* model);

/* Message of the last failure on this thread ("" if none). */
const char* dmx_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* DMSIM_AMD_H */
