/*
 * oracle/dmsim_oracle.c -- CPU restatement of DM-Sim's density-matrix path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the package's
 * smoke() and bench.py's cpu_baseline leg load it; the product (the HIP
 * library under dm-sim_amd/) never links, loads or calls it.
 *
 * What it restates (reference = pnnl/DM-Sim v2.5, paths relative to its root):
 *   - gate primitives applied to the fast (column) index of every row, one
 *     full sweep per primitive: src/dmsim_cpu_omp.hpp:959-1410 (OP_HEAD family,
 *     C1_GATE :971-992, C2_GATE :995-1088, CX_GATE :1101-1135, X..U1 :1142-1410)
 *   - composite gates as literal sequences of primitives:
 *     src/dmsim_cpu_omp.hpp:2372-2728 (qelib1 decompositions, tool/qelib1.inc)
 *   - op -> gate mapping and parameter placement: src/dmsim_cpu_omp.hpp:2734-2921
 *     and the factories :546-733
 *   - one sim() run = forward pass, conjugate transpose, backward pass
 *     (src/dmsim_cpu_omp.hpp:888-935 single-thread branch, sim() :367-447).
 *     The transpose here is a correct conj-transpose; the reference's serial
 *     emulation (:791-825) reads LDS-emulation cells before writing them and is
 *     only right for the first run from |0><0|.  Continued runs follow the GPU
 *     kernel (src/dmsim_amdgpu_omp.hpp:840-871), which synchronises correctly.
 *   - measure(): |Re rho_ii| prefix sum, glibc srand/rand sampling
 *     (src/dmsim_cpu_omp.hpp:474-515).
 *
 * Storage convention (same as the reference's dm_real_res/dm_imag_res):
 *   re[i*dim + j] = Re(conj(rho))_{ij},  im[i*dim + j] = Im(conj(rho))_{ij}.
 * Arithmetic is written in the reference's per-element evaluation order so a
 * run agrees with the reference bit for bit (x86-64 baseline, no FMA
 * contraction); the pinning tests check exactly that against golden vectors
 * produced from the reference's own sources.
 *
 * Indices are 64-bit (the reference uses 32-bit `unsigned`, src/config.hpp:43,
 * which overflows at n = 15).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Every coefficient set of the reference takes cos and sin of one argument;
 * g++ -O3 compiles each pair to a single glibc sincos() call (ten sites in
 * the compiled reference), which can differ from cos() in the last bit, so
 * the oracle calls sincos() explicitly. */
static void sc(double x, double* s, double* c) { sincos(x, s, c); }

/* Constants exactly as src/config.hpp:55-57 writes them. */
#define ORC_PI 3.14159265358979323846
#define ORC_S2I 0.70710678118654752440

/* Op codes: the reference's enum OP order (src/dmsim_cpu_omp.hpp:43-49),
 * plus two extensions for arbitrary 1- and 2-qubit matrices (C1_GATE/C2_GATE
 * exist in the reference but no OP reaches them). */
enum {
    O_U3, O_U2, O_U1, O_CX, O_ID, O_X, O_Y, O_Z, O_H, O_S,
    O_SDG, O_T, O_TDG, O_RX, O_RY, O_RZ, O_CZ, O_CY, O_SWAP, O_CH,
    O_CCX, O_CSWAP, O_CRX, O_CRY, O_CRZ, O_CU1, O_CU3, O_RXX, O_RZZ, O_RCCX,
    O_RC3X, O_C3X, O_C3SQRTX, O_C4X, O_R, O_SRN, O_W, O_RYY,
    O_C1 = 38, O_C2 = 39, O_NUM = 40
};

typedef int64_t idx_t;

/* A block of `rows` rows, each 2^n complex entries, split planes. */
typedef struct {
    double* re;
    double* im;
    int n;
    idx_t rows;
} rows_t;

static int g_threads = 1;

void orc_set_threads(int t) { g_threads = t > 0 ? t : 1; }

/* Threads of a primitive's own loop: g_threads, or 1 inside the row-wise
 * forward pass (whose rows are already spread over the threads). */
static int team(void)
{
#ifdef _OPENMP
    return omp_in_parallel() ? 1 : g_threads;
#else
    return 1;
#endif
}

/* Pair enumeration within each row (src/dmsim_cpu_omp.hpp:959-968): pair p of
 * a row maps to pos0 = p with a 0 inserted at bit q, pos1 = pos0 + 2^q. */
#define PAIR_LOOP_BEGIN(R, Q)                                                   \
    const idx_t dim_ = (idx_t)1 << (R)->n;                                      \
    const idx_t half_ = dim_ >> 1;                                              \
    const idx_t rows_ = (R)->rows;                                              \
    const idx_t lowmask_ = ((idx_t)1 << (Q)) - 1;                               \
    double* const dr = (R)->re;                                                 \
    double* const di = (R)->im;                                                 \
    _Pragma("omp parallel for collapse(2) num_threads(team()) schedule(static)") \
    for (idx_t row_ = 0; row_ < rows_; row_++)                                  \
    for (idx_t p_ = 0; p_ < half_; p_++) {                                      \
        const idx_t pos0 = row_ * dim_ + (((p_ & ~lowmask_) << 1) | (p_ & lowmask_)); \
        const idx_t pos1 = pos0 + ((idx_t)1 << (Q));
#define PAIR_LOOP_END }

/* ---- 1-qubit primitives: formulas follow src/dmsim_cpu_omp.hpp:971-1410 ---- */

/* general 2x2: e = {e0r,e0i,e1r,e1i,e2r,e2i,e3r,e3i}  (:971-992) */
static void p_c1(rows_t* R, int q, const double* e)
{
    const double e0r = e[0], e0i = e[1], e1r = e[2], e1i = e[3];
    const double e2r = e[4], e2i = e[5], e3r = e[6], e3i = e[7];
    PAIR_LOOP_BEGIN(R, q)
        const double a0 = dr[pos0], b0 = di[pos0], a1 = dr[pos1], b1 = di[pos1];
        dr[pos0] = (e0r * a0) - (e0i * b0) + (e1r * a1) - (e1i * b1);
        di[pos0] = (e0r * b0) + (e0i * a0) + (e1r * b1) + (e1i * a1);
        dr[pos1] = (e2r * a0) - (e2i * b0) + (e3r * a1) - (e3i * b1);
        di[pos1] = (e2r * b0) + (e2i * a0) + (e3r * b1) + (e3i * a1);
    PAIR_LOOP_END
}

static void p_x(rows_t* R, int q) /* :1142-1154 */
{
    PAIR_LOOP_BEGIN(R, q)
        const double a0 = dr[pos0], b0 = di[pos0], a1 = dr[pos1], b1 = di[pos1];
        dr[pos0] = a1; di[pos0] = b1; dr[pos1] = a0; di[pos1] = b0;
    PAIR_LOOP_END
}

static void p_y(rows_t* R, int q) /* :1160-1174 */
{
    PAIR_LOOP_BEGIN(R, q)
        const double a0 = dr[pos0], b0 = di[pos0], a1 = dr[pos1], b1 = di[pos1];
        dr[pos0] = b1; di[pos0] = -a1; dr[pos1] = -b0; di[pos1] = a0;
    PAIR_LOOP_END
}

static void p_z(rows_t* R, int q) /* :1180-1189 */
{
    PAIR_LOOP_BEGIN(R, q)
        const double a1 = dr[pos1], b1 = di[pos1];
        dr[pos1] = -a1; di[pos1] = -b1;
    PAIR_LOOP_END
}

static void p_h(rows_t* R, int q) /* :1195-1209 */
{
    PAIR_LOOP_BEGIN(R, q)
        const double a0 = dr[pos0], b0 = di[pos0], a1 = dr[pos1], b1 = di[pos1];
        dr[pos0] = ORC_S2I * (a0 + a1); di[pos0] = ORC_S2I * (b0 + b1);
        dr[pos1] = ORC_S2I * (a0 - a1); di[pos1] = ORC_S2I * (b0 - b1);
    PAIR_LOOP_END
}

static void p_srn(rows_t* R, int q) /* :1218-1231 */
{
    PAIR_LOOP_BEGIN(R, q)
        const double a0 = dr[pos0], b0 = di[pos0], a1 = dr[pos1], b1 = di[pos1];
        dr[pos0] = 0.5 * (a0 + a1); di[pos0] = 0.5 * (b0 - b1);
        dr[pos1] = 0.5 * (a0 + a1); di[pos1] = 0.5 * (-b0 + b1);
    PAIR_LOOP_END
}

static void p_r(rows_t* R, int q, double phase) /* :1239-1248 */
{
    PAIR_LOOP_BEGIN(R, q)
        const double a1 = dr[pos1], b1 = di[pos1];
        dr[pos1] = -(b1 * phase); di[pos1] = a1 * phase;
    PAIR_LOOP_END
}

static void p_s(rows_t* R, int q) /* :1258-1266 */
{
    PAIR_LOOP_BEGIN(R, q)
        const double a1 = dr[pos1], b1 = di[pos1];
        dr[pos1] = -b1; di[pos1] = a1;
    PAIR_LOOP_END
}

static void p_sdg(rows_t* R, int q) /* :1273-1281 */
{
    PAIR_LOOP_BEGIN(R, q)
        const double a1 = dr[pos1], b1 = di[pos1];
        dr[pos1] = b1; di[pos1] = -a1;
    PAIR_LOOP_END
}

static void p_t(rows_t* R, int q) /* :1288-1296 */
{
    PAIR_LOOP_BEGIN(R, q)
        const double a1 = dr[pos1], b1 = di[pos1];
        dr[pos1] = ORC_S2I * (a1 - b1); di[pos1] = ORC_S2I * (a1 + b1);
    PAIR_LOOP_END
}

static void p_tdg(rows_t* R, int q) /* :1303-1311 */
{
    PAIR_LOOP_BEGIN(R, q)
        const double a1 = dr[pos1], b1 = di[pos1];
        dr[pos1] = ORC_S2I * (a1 + b1); di[pos1] = ORC_S2I * (-a1 + b1);
    PAIR_LOOP_END
}

static void p_rx(rows_t* R, int q, double theta) /* :1335-1351 */
{
    double hs, hc;
    sc(theta / 2.0, &hs, &hc);
    const double xr = hc;
    const double xi = -hs;
    PAIR_LOOP_BEGIN(R, q)
        const double a0 = dr[pos0], b0 = di[pos0], a1 = dr[pos1], b1 = di[pos1];
        dr[pos0] = (xr * a0) - (xi * b1);
        di[pos0] = (xr * b0) + (xi * a1);
        dr[pos1] = -(xi * b0) + (xr * a1);
        di[pos1] = +(xi * a0) + (xr * b1);
    PAIR_LOOP_END
}

static void p_ry(rows_t* R, int q, double theta) /* :1355-1372 */
{
    double hs, hc;
    sc(theta / 2.0, &hs, &hc);
    const double c0 = hc;
    const double c1 = -hs;
    const double c2 = hs;
    const double c3 = hc;
    PAIR_LOOP_BEGIN(R, q)
        const double a0 = dr[pos0], b0 = di[pos0], a1 = dr[pos1], b1 = di[pos1];
        dr[pos0] = (c0 * a0) + (c1 * a1);
        di[pos0] = (c0 * b0) + (c1 * b1);
        dr[pos1] = (c2 * a0) + (c3 * a1);
        di[pos1] = (c2 * b0) + (c3 * b1);
    PAIR_LOOP_END
}

static void p_w(rows_t* R, int q) /* :1379-1391 */
{
    PAIR_LOOP_BEGIN(R, q)
        const double a0 = dr[pos0], b0 = di[pos0], a1 = dr[pos1], b1 = di[pos1];
        dr[pos0] = ORC_S2I * (a0 + b1); di[pos0] = ORC_S2I * (b0 - a1);
        dr[pos1] = ORC_S2I * (b0 + a1); di[pos1] = ORC_S2I * (-a0 + b1);
    PAIR_LOOP_END
}

static void p_u1(rows_t* R, int q, double lambda) /* :1395-1410 */
{
    double ls, lc;
    sc(lambda, &ls, &lc);
    const double c3r = lc;
    const double c3i = ls;
    PAIR_LOOP_BEGIN(R, q)
        const double a1 = dr[pos1], b1 = di[pos1];
        dr[pos1] = (c3r * a1) - (c3i * b1);
        di[pos1] = (c3r * b1) + (c3i * a1);
    PAIR_LOOP_END
}

/* U2/U3 build C1 coefficients (:2384-2415). */
static void p_u2(rows_t* R, int q, double phi, double lambda)
{
    double ls, lc, ps, pc, qs, qc, e[8];
    sc(lambda, &ls, &lc);
    sc(phi, &ps, &pc);
    sc(phi + lambda, &qs, &qc);
    e[0] = ORC_S2I; e[1] = 0;
    e[2] = -ORC_S2I * lc; e[3] = -ORC_S2I * ls;
    e[4] = ORC_S2I * pc; e[5] = ORC_S2I * ps;
    e[6] = ORC_S2I * qc; e[7] = ORC_S2I * qs;
    p_c1(R, q, e);
}

static void p_u3(rows_t* R, int q, double theta, double phi, double lambda)
{
    double hs, hc, ls, lc, ps, pc, qs, qc, e[8];
    sc(theta / 2., &hs, &hc);
    sc(lambda, &ls, &lc);
    sc(phi, &ps, &pc);
    sc(phi + lambda, &qs, &qc);
    e[0] = hc; e[1] = 0;
    e[2] = -lc * hs; e[3] = -ls * hs;
    e[4] = pc * hs; e[5] = ps * hs;
    e[6] = qc * hc; e[7] = qs * hc;
    p_c1(R, q, e);
}

/* ---- 2-qubit primitives ---- */

/* CX (:1101-1135): swap (ctrl=1,tgt=0) <-> (ctrl=1,tgt=1). */
static void p_cx(rows_t* R, int ctrl, int tgt)
{
    const int hi = ctrl > tgt ? ctrl : tgt;
    const int lo = ctrl > tgt ? tgt : ctrl;
    const idx_t dim = (idx_t)1 << R->n;
    const idx_t quarter = dim >> 2;
    const idx_t rows = R->rows;
    const idx_t lo_mask = ((idx_t)1 << lo) - 1;
    const idx_t mid_mask = ((idx_t)1 << hi) - 1;
    double* const dr = R->re;
    double* const di = R->im;
#pragma omp parallel for collapse(2) num_threads(team()) schedule(static)
    for (idx_t row = 0; row < rows; row++)
    for (idx_t k = 0; k < quarter; k++) {
        /* insert zeros at bit lo, then at bit hi */
        idx_t b = ((k & ~lo_mask) << 1) | (k & lo_mask);
        b = ((b & ~mid_mask) << 1) | (b & mid_mask);
        const idx_t base = row * dim + b;
        const idx_t pos0 = base + ((idx_t)1 << ctrl);
        const idx_t pos1 = base + ((idx_t)1 << ctrl) + ((idx_t)1 << tgt);
        const double a0 = dr[pos0], b0 = di[pos0], a1 = dr[pos1], b1 = di[pos1];
        dr[pos0] = a1; di[pos0] = b1; dr[pos1] = a0; di[pos1] = b0;
    }
}

/* C2 (:995-1088): e = 16 complex entries row-major (e[2*(4*r+c)] real part).
 * Quad order: pos0 = base, pos1 = base + 2^q2, pos2 = base + 2^q1,
 * pos3 = base + 2^q1 + 2^q2 (qubit q1 is the high bit of the 4x4 index). */
static void p_c2(rows_t* R, int q1, int q2, const double* e)
{
    const int hi = q1 > q2 ? q1 : q2;
    const int lo = q1 > q2 ? q2 : q1;
    const idx_t dim = (idx_t)1 << R->n;
    const idx_t quarter = dim >> 2;
    const idx_t rows = R->rows;
    const idx_t lo_mask = ((idx_t)1 << lo) - 1;
    const idx_t mid_mask = ((idx_t)1 << hi) - 1;
    double* const dr = R->re;
    double* const di = R->im;
#pragma omp parallel for collapse(2) num_threads(team()) schedule(static)
    for (idx_t row = 0; row < rows; row++)
    for (idx_t k = 0; k < quarter; k++) {
        idx_t b = ((k & ~lo_mask) << 1) | (k & lo_mask);
        b = ((b & ~mid_mask) << 1) | (b & mid_mask);
        const idx_t base = row * dim + b;
        idx_t pos[4];
        pos[0] = base;
        pos[1] = base + ((idx_t)1 << q2);
        pos[2] = base + ((idx_t)1 << q1);
        pos[3] = base + ((idx_t)1 << q1) + ((idx_t)1 << q2);
        double xr[4], xi[4];
        for (int t = 0; t < 4; t++) { xr[t] = dr[pos[t]]; xi[t] = di[pos[t]]; }
        for (int r = 0; r < 4; r++) {
            const double* er = e + 8 * r;
            /* left-to-right accumulation exactly as the reference spells it */
            double s = (er[0] * xr[0]) - (er[1] * xi[0]);
            s = s + (er[2] * xr[1]); s = s - (er[3] * xi[1]);
            s = s + (er[4] * xr[2]); s = s - (er[5] * xi[2]);
            s = s + (er[6] * xr[3]); s = s - (er[7] * xi[3]);
            dr[pos[r]] = s;
        }
        for (int r = 0; r < 4; r++) {
            const double* er = e + 8 * r;
            double s = (er[0] * xi[0]) + (er[1] * xr[0]);
            s = s + (er[2] * xi[1]); s = s + (er[3] * xr[1]);
            s = s + (er[4] * xi[2]); s = s + (er[5] * xr[2]);
            s = s + (er[6] * xi[3]); s = s + (er[7] * xr[3]);
            di[pos[r]] = s;
        }
    }
}

/* ---- composite gates (:2427-2728), same primitive order ---- */

static void g_cu1(rows_t* R, double lambda, int a, int b)
{
    p_u1(R, a, lambda / 2);
    p_cx(R, a, b);
    p_u1(R, b, -lambda / 2);
    p_cx(R, a, b);
    p_u1(R, b, lambda / 2);
}

static void g_ccx(rows_t* R, int a, int b, int c)
{
    p_h(R, c);
    p_cx(R, b, c);
    p_tdg(R, c);
    p_cx(R, a, c);
    p_t(R, c);
    p_cx(R, b, c);
    p_tdg(R, c);
    p_cx(R, a, c);
    p_t(R, b);
    p_t(R, c);
    p_h(R, c);
    p_cx(R, a, b);
    p_t(R, a);
    p_tdg(R, b);
    p_cx(R, a, b);
}

/* C3X / C3SQRTX share the pattern; `ang` is PI/4 or PI/8 (:2632-2697). */
static void g_c3x_like(rows_t* R, double ang, int a, int b, int c, int d)
{
    p_h(R, d); g_cu1(R, -ang, a, d); p_h(R, d);
    p_cx(R, a, b);
    p_h(R, d); g_cu1(R, ang, b, d); p_h(R, d);
    p_cx(R, a, b);
    p_h(R, d); g_cu1(R, -ang, b, d); p_h(R, d);
    p_cx(R, b, c);
    p_h(R, d); g_cu1(R, ang, c, d); p_h(R, d);
    p_cx(R, a, c);
    p_h(R, d); g_cu1(R, -ang, c, d); p_h(R, d);
    p_cx(R, b, c);
    p_h(R, d); g_cu1(R, ang, c, d); p_h(R, d);
    p_cx(R, a, c);
    p_h(R, d); g_cu1(R, -ang, c, d); p_h(R, d);
}

static void apply_gate(rows_t* R, int op, const uint32_t* qb, const double* prm,
                       const double* coef)
{
    const int a = (int)qb[0], b = (int)qb[1], c = (int)qb[2], d = (int)qb[3], e = (int)qb[4];
    const double theta = prm[0], phi = prm[1], lambda = prm[2];
    switch (op) {
    case O_U3: p_u3(R, a, theta, phi, lambda); break;
    case O_U2: p_u2(R, a, phi, lambda); break;
    case O_U1: p_u1(R, a, lambda); break;
    case O_CX: p_cx(R, a, b); break;
    case O_ID: break;
    case O_X: p_x(R, a); break;
    case O_Y: p_y(R, a); break;
    case O_Z: p_z(R, a); break;
    case O_H: p_h(R, a); break;
    case O_S: p_s(R, a); break;
    case O_SDG: p_sdg(R, a); break;
    case O_T: p_t(R, a); break;
    case O_TDG: p_tdg(R, a); break;
    case O_RX: p_rx(R, a, theta); break;
    case O_RY: p_ry(R, a, theta); break;
    case O_RZ: p_u1(R, a, phi); break; /* RZ stores phi, runs U1 (:2419-2423) */
    case O_CZ: p_h(R, b); p_cx(R, a, b); p_h(R, b); break;
    case O_CY: p_sdg(R, b); p_cx(R, a, b); p_s(R, b); break;
    case O_SWAP: p_cx(R, a, b); p_cx(R, b, a); p_cx(R, a, b); break;
    case O_CH:
        p_h(R, b); p_sdg(R, b); p_cx(R, a, b); p_h(R, b); p_t(R, b);
        p_cx(R, a, b); p_t(R, b); p_h(R, b); p_s(R, b); p_x(R, b); p_s(R, a);
        break;
    case O_CCX: g_ccx(R, a, b, c); break;
    case O_CSWAP: p_cx(R, c, b); g_ccx(R, a, b, c); p_cx(R, c, b); break;
    case O_CRX:
        p_u1(R, b, ORC_PI / 2); p_cx(R, a, b); p_u3(R, b, -lambda / 2, 0, 0);
        p_cx(R, a, b); p_u3(R, b, lambda / 2, -ORC_PI / 2, 0);
        break;
    case O_CRY:
        p_u3(R, b, lambda / 2, 0, 0); p_cx(R, a, b);
        p_u3(R, b, -lambda / 2, 0, 0); p_cx(R, a, b);
        break;
    case O_CRZ:
        p_u1(R, b, lambda / 2); p_cx(R, a, b); p_u1(R, b, -lambda / 2); p_cx(R, a, b);
        break;
    case O_CU1: g_cu1(R, lambda, a, b); break;
    case O_CU3: {
        const double t1 = (lambda - phi) / 2;
        const double t2 = theta / 2;
        const double t3 = -(phi + lambda) / 2;
        p_u1(R, a, -t3); p_u1(R, b, t1); p_cx(R, a, b);
        p_u3(R, b, -t2, 0, t3); p_cx(R, a, b); p_u3(R, b, t2, phi, 0);
        break;
    }
    case O_RXX:
        p_u3(R, a, ORC_PI / 2, theta, 0); p_h(R, b); p_cx(R, a, b);
        p_u1(R, b, -theta); p_cx(R, a, b); p_h(R, b);
        p_u2(R, a, -ORC_PI, ORC_PI - theta);
        break;
    case O_RZZ: p_cx(R, a, b); p_u1(R, b, theta); p_cx(R, a, b); break;
    case O_RCCX:
        p_u2(R, c, 0, ORC_PI); p_u1(R, c, ORC_PI / 4); p_cx(R, b, c);
        p_u1(R, c, -ORC_PI / 4); p_cx(R, a, c); p_u1(R, c, ORC_PI / 4);
        p_cx(R, b, c); p_u1(R, c, -ORC_PI / 4); p_u2(R, c, 0, ORC_PI);
        break;
    case O_RC3X:
        p_u2(R, d, 0, ORC_PI); p_u1(R, d, ORC_PI / 4); p_cx(R, c, d);
        p_u1(R, d, -ORC_PI / 4); p_u2(R, d, 0, ORC_PI); p_cx(R, a, d);
        p_u1(R, d, ORC_PI / 4); p_cx(R, b, d); p_u1(R, d, -ORC_PI / 4);
        p_cx(R, a, d); p_u1(R, d, ORC_PI / 4); p_cx(R, b, d);
        p_u1(R, d, -ORC_PI / 4); p_u2(R, d, 0, ORC_PI); p_u1(R, d, ORC_PI / 4);
        p_cx(R, c, d); p_u1(R, d, -ORC_PI / 4); p_u2(R, d, 0, ORC_PI);
        break;
    case O_C3X: g_c3x_like(R, ORC_PI / 4, a, b, c, d); break;
    case O_C3SQRTX: g_c3x_like(R, ORC_PI / 8, a, b, c, d); break;
    case O_C4X: /* qelib1 c4x literally (:2701-2714) */
        p_h(R, e); g_cu1(R, -ORC_PI / 2, d, e); p_h(R, e);
        g_c3x_like(R, ORC_PI / 4, a, b, c, d);
        p_h(R, d); g_cu1(R, ORC_PI / 4, d, e); p_h(R, d);
        g_c3x_like(R, ORC_PI / 4, a, b, c, d);
        g_c3x_like(R, ORC_PI / 8, a, b, c, e);
        break;
    case O_R: p_r(R, a, theta); break;
    case O_SRN: p_srn(R, a); break;
    case O_W: p_w(R, a); break;
    case O_RYY:
        p_rx(R, a, ORC_PI / 2); p_rx(R, b, ORC_PI / 2); p_cx(R, a, b);
        p_u1(R, b, theta); p_cx(R, a, b);
        p_rx(R, a, -ORC_PI / 2); p_rx(R, b, -ORC_PI / 2);
        break;
    case O_C1: p_c1(R, a, coef); break;
    case O_C2: p_c2(R, a, b, coef); break;
    default: break;
    }
}

/* ---------------------------------------------------------------- API ---- */

/* Validates a gate list: op codes and qubit indices (< n; distinct among the
 * qubits the op uses).  Returns 0 or -(index+1) of the first bad gate. */
static int op_arity(int op)
{
    switch (op) {
    case O_CX: case O_CZ: case O_CY: case O_SWAP: case O_CH: case O_CRX: case O_CRY:
    case O_CRZ: case O_CU1: case O_CU3: case O_RXX: case O_RZZ: case O_RYY: case O_C2:
        return 2;
    case O_CCX: case O_CSWAP: case O_RCCX: return 3;
    case O_RC3X: case O_C3X: case O_C3SQRTX: return 4;
    case O_C4X: return 5;
    default: return 1;
    }
}

int orc_op_arity(int op) { return op_arity(op); }

static int check_gates(int n, int ng, const int32_t* ops, const uint32_t* qb)
{
    for (int g = 0; g < ng; g++) {
        if (ops[g] < 0 || ops[g] >= O_NUM) return -(g + 1);
        const int k = op_arity(ops[g]);
        for (int i = 0; i < k; i++) {
            if (qb[5 * g + i] >= (uint32_t)n) return -(g + 1);
            for (int j = 0; j < i; j++)
                if (qb[5 * g + i] == qb[5 * g + j]) return -(g + 1);
        }
    }
    return 0;
}

/* Forward pass: every gate applied to the fast index of each of `rows` rows. */
int orc_forward(int n, idx_t rows, double* re, double* im, int ng, const int32_t* ops,
                const uint32_t* qb, const double* prm, const double* coef)
{
    const int bad = check_gates(n, ng, ops, qb);
    if (bad) return bad;
    rows_t R = {re, im, n, rows};
    for (int g = 0; g < ng; g++)
        apply_gate(&R, ops[g], qb + 5 * g, prm + 3 * g, coef ? coef + 32 * g : NULL);
    return 0;
}

/* The same forward pass in another loop order: each thread applies every
 * gate to one row before taking the next.  Rows do not interact in the
 * forward pass and every entry sees the same operations in the same order,
 * so the result is bit-identical to orc_forward; a row (2^n entries) stays in
 * cache across the gates instead of every gate streaming all rows. */
int orc_forward_rowwise(int n, idx_t rows, double* re, double* im, int ng, const int32_t* ops,
                        const uint32_t* qb, const double* prm, const double* coef)
{
    const int bad = check_gates(n, ng, ops, qb);
    if (bad) return bad;
    const idx_t dim = (idx_t)1 << n;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1)
    for (idx_t r = 0; r < rows; r++) {
        rows_t R = {re + r * dim, im + r * dim, n, 1};
        for (int g = 0; g < ng; g++)
            apply_gate(&R, ops[g], qb + 5 * g, prm + 3 * g, coef ? coef + 32 * g : NULL);
    }
    return 0;
}

/* In-place conjugate transpose of a dim x dim split-plane matrix. */
void orc_conj_transpose(int n, double* re, double* im)
{
    const idx_t dim = (idx_t)1 << n;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (idx_t i = 0; i < dim; i++) {
        im[i * dim + i] = -im[i * dim + i];
        for (idx_t j = i + 1; j < dim; j++) {
            const double ar = re[i * dim + j], ai = im[i * dim + j];
            const double br = re[j * dim + i], bi = im[j * dim + i];
            re[i * dim + j] = br; im[i * dim + j] = -bi;
            re[j * dim + i] = ar; im[j * dim + i] = -ai;
        }
    }
}

/* One sim() run on the state held in (re, im): forward pass, conjugate
 * transpose, backward pass.  Result in place, reference orientation. */
int orc_sim(int n, double* re, double* im, int ng, const int32_t* ops, const uint32_t* qb,
            const double* prm, const double* coef)
{
    const idx_t dim = (idx_t)1 << n;
    int rc = orc_forward(n, dim, re, im, ng, ops, qb, prm, coef);
    if (rc) return rc;
    orc_conj_transpose(n, re, im);
    return orc_forward(n, dim, re, im, ng, ops, qb, prm, coef);
}

/* orc_sim with the row-wise forward pass (bit-identical, faster at large n). */
int orc_sim_rowwise(int n, double* re, double* im, int ng, const int32_t* ops, const uint32_t* qb,
                    const double* prm, const double* coef)
{
    const idx_t dim = (idx_t)1 << n;
    int rc = orc_forward_rowwise(n, dim, re, im, ng, ops, qb, prm, coef);
    if (rc) return rc;
    orc_conj_transpose(n, re, im);
    return orc_forward_rowwise(n, dim, re, im, ng, ops, qb, prm, coef);
}

/* |0..0><0..0| */
void orc_init(int n, double* re, double* im)
{
    const idx_t num = (idx_t)1 << (2 * n);
    memset(re, 0, (size_t)num * sizeof(double));
    memset(im, 0, (size_t)num * sizeof(double));
    re[0] = 1.0;
}

/* measure(repetition) on a result matrix (src/dmsim_cpu_omp.hpp:474-515):
 * linear scan for the j with scan[j] <= r < scan[j+1]; 0 when none matches.
 * Returns the probability sum scan[dim]. */
double orc_measure(int n, const double* re, uint32_t repetition, uint32_t seed, uint32_t* out)
{
    const idx_t dim = (idx_t)1 << n;
    double* scan = (double*)malloc((size_t)(dim + 1) * sizeof(double));
    scan[0] = 0;
    for (idx_t i = 1; i < dim + 1; i++) scan[i] = scan[i - 1] + fabs(re[(i - 1) * dim + (i - 1)]);
    srand(seed);
    for (uint32_t s = 0; s < repetition; s++) {
        const double r = (double)rand() / (double)RAND_MAX;
        uint32_t res = 0;
        for (idx_t j = 0; j < dim; j++)
            if (scan[j] <= r && r < scan[j + 1]) res = (uint32_t)j;
        out[s] = res;
    }
    const double total = scan[dim];
    free(scan);
    return total;
}

/* measure() from a diagonal (Re rho_ii, i < 2^n) instead of the full matrix. */
double orc_measure_diag(int n, const double* diag, uint32_t repetition, uint32_t seed,
                        uint32_t* out)
{
    const idx_t dim = (idx_t)1 << n;
    double* scan = (double*)malloc((size_t)(dim + 1) * sizeof(double));
    scan[0] = 0;
    for (idx_t i = 1; i < dim + 1; i++) scan[i] = scan[i - 1] + fabs(diag[i - 1]);
    srand(seed);
    for (uint32_t s = 0; s < repetition; s++) {
        const double r = (double)rand() / (double)RAND_MAX;
        uint32_t res = 0;
        for (idx_t j = 0; j < dim; j++)
            if (scan[j] <= r && r < scan[j + 1]) res = (uint32_t)j;
        out[s] = res;
    }
    const double total = scan[dim];
    free(scan);
    return total;
}

/* State-vector twin (pure states): psi <- M psi with the same primitives,
 * i.e. the forward pass on a single row.  For a unitary circuit run from
 * |0><0|, the reference result is res[i][j] = conj(psi_i) * psi_j up to
 * rounding. */
int orc_sv_run(int n, double* psi_re, double* psi_im, int ng, const int32_t* ops,
               const uint32_t* qb, const double* prm, const double* coef)
{
    return orc_forward(n, 1, psi_re, psi_im, ng, ops, qb, prm, coef);
}
