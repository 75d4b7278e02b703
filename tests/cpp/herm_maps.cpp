// Host-side check of the Hermitian fused executor's index maps: plans built
// with Hermitian sweeps (dmx::build_fused_plan, herm = true) are executed on
// the CPU step by step with the same per-thread load/store maps the kernels
// use (herm.hpp: herm_load/herm_direct/herm_mirror for k_fused4_herm,
// relabel_load/relabel_store for k_relabel; five-qubit blocks and channels
// in place on the index bits their records name), on a random Hermitian stored
// matrix S, and compared with S <- conj(M) S M^T applied gate by gate with the
// gates' own matrices (dmx::gate_matrix; channels as their Kraus sums
// straight from the operators).  Every store must stay inside the
// planes, every output entry must be written exactly once per step, and the
// final S must agree to 1e-12.  Prints "<cases> cases OK", or the first
// violation and exits nonzero.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <random>
#include <vector>

#include "fuse.hpp"
#include "herm.hpp"
#include "lower.hpp"

using dmx::Gate;
using cd = std::complex<double>;

static int fail(const char* what, int n, int seed)
{
    std::printf("FAIL %s (n=%d seed=%d)\n", what, n, seed);
    return 1;
}

// A channel's Kraus operators: the two 2x2 halves of a random 4x2 isometry V
// (sum K^dagger K = V^dagger V = I).
struct Channel {
    uint32_t q;
    cd K[2][4];
};

// The channel as the engine stores it: OP_KRAUS with the superoperator on the
// entry quad (row bit a, column bit b) of the stored A = conj(rho),
// S[(a,b),(a',b')] = sum_i conj(K_i[a][a']) K_i[b][b'], then its twin for the
// conj-transposed orientation.
static Gate kraus_gate(const Channel& ch)
{
    Gate g;
    g.op = dmx::OP_KRAUS;
    g.k = 1;
    g.qb[0] = ch.q;
    g.mre.assign(32, 0.0);
    g.mim.assign(32, 0.0);
    cd S[16];
    for (int idx = 0; idx < 16; idx++) {
        const int a = idx >> 3, b = (idx >> 2) & 1, a2 = (idx >> 1) & 1, b2 = idx & 1;
        S[idx] = 0.0;
        for (int i = 0; i < 2; i++) S[idx] += std::conj(ch.K[i][2 * a + a2]) * ch.K[i][2 * b + b2];
    }
    for (int idx = 0; idx < 16; idx++) {
        const int a = idx >> 3, b = (idx >> 2) & 1, a2 = (idx >> 1) & 1, b2 = idx & 1;
        const int tdx = (2 * b + a) * 4 + (2 * b2 + a2);
        g.mre[idx] = S[idx].real();
        g.mim[idx] = S[idx].imag();
        g.mre[16 + idx] = S[tdx].real();
        g.mim[16 + idx] = -S[tdx].imag();
    }
    return g;
}

// Brick-wall layers of random two-qubit matrices, then one- and two-qubit
// gates with many X, CX and SWAP (trailing index maps), then a few wider ones;
// with c4x, a five-qubit gate halfway through those (an in-order five-qubit
// step between Hermitian sweeps) and a random channel a quarter and three
// quarters of the way (their Kraus operators in chans).
static std::vector<Gate> circuit(int n, int seed, bool c4x = false,
                                 std::vector<Channel>* chans = nullptr)
{
    static const int kOps[] = {dmx::OP_U3, dmx::OP_RZ, dmx::OP_H, dmx::OP_X, dmx::OP_X,
                               dmx::OP_CX, dmx::OP_CX, dmx::OP_CX, dmx::OP_SWAP, dmx::OP_SWAP,
                               dmx::OP_CU1, dmx::OP_RZZ, dmx::OP_CCX, dmx::OP_RY};
    constexpr int kNum = sizeof(kOps) / sizeof(kOps[0]);
    std::mt19937 rng((uint32_t)seed);
    std::uniform_real_distribution<double> ang(-3.2, 3.2);
    std::normal_distribution<double> g(0.0, 1.0);
    std::vector<Gate> c;
    for (int layer = 0; layer < 3; layer++)
        for (int q = layer & 1; q + 1 < n; q += 2) {
            Gate m;
            m.op = dmx::OP_MAT;
            m.k = 2;
            m.qb[0] = (uint32_t)q;
            m.qb[1] = (uint32_t)q + 1;
            m.mre.resize(16);
            m.mim.resize(16);
            // a unitary: exp(iH) is not needed, any 4x4 from Gram-Schmidt
            std::vector<cd> M(16);
            for (int col = 0; col < 4; col++) {
                std::vector<cd> v(4);
                for (int r = 0; r < 4; r++) v[r] = cd(g(rng), g(rng));
                for (int p = 0; p < col; p++) {
                    cd dot = 0.0;
                    for (int r = 0; r < 4; r++) dot += std::conj(M[r * 4 + p]) * v[r];
                    for (int r = 0; r < 4; r++) v[r] -= dot * M[r * 4 + p];
                }
                double nrm = 0;
                for (int r = 0; r < 4; r++) nrm += std::norm(v[r]);
                for (int r = 0; r < 4; r++) M[r * 4 + col] = v[r] / std::sqrt(nrm);
            }
            for (int i = 0; i < 16; i++) m.mre[i] = M[i].real(), m.mim[i] = M[i].imag();
            c.push_back(m);
        }
    const int n_rand = 90 + 10 * n;
    for (int i = 0; i < n_rand; i++) {
        if (c4x && chans && (i == n_rand / 4 || i == 3 * n_rand / 4)) {
            Channel ch;
            ch.q = (uint32_t)(rng() % n);
            cd V[4][2];
            for (int col = 0; col < 2; col++) {
                for (int r = 0; r < 4; r++) V[r][col] = cd(g(rng), g(rng));
                for (int p = 0; p < col; p++) {
                    cd dot = 0.0;
                    for (int r = 0; r < 4; r++) dot += std::conj(V[r][p]) * V[r][col];
                    for (int r = 0; r < 4; r++) V[r][col] -= dot * V[r][p];
                }
                double nrm = 0;
                for (int r = 0; r < 4; r++) nrm += std::norm(V[r][col]);
                for (int r = 0; r < 4; r++) V[r][col] /= std::sqrt(nrm);
            }
            for (int i2 = 0; i2 < 2; i2++)
                for (int r = 0; r < 2; r++)
                    for (int col = 0; col < 2; col++) ch.K[i2][2 * r + col] = V[2 * i2 + r][col];
            chans->push_back(ch);
            c.push_back(kraus_gate(ch));
        }
        if (c4x && i == n_rand / 2) {
            std::vector<int> q(n);
            for (int t = 0; t < n; t++) q[t] = t;
            std::shuffle(q.begin(), q.end(), rng);
            Gate x;
            x.op = dmx::OP_C4X;
            for (int t = 0; t < 5; t++) x.qb[t] = (uint32_t)q[t];
            c.push_back(x);
        }
        Gate x;
        x.op = kOps[rng() % kNum];
        x.theta = ang(rng), x.phi = ang(rng), x.lambda = ang(rng);
        std::vector<int> q(n);
        for (int t = 0; t < n; t++) q[t] = t;
        std::shuffle(q.begin(), q.end(), rng);
        for (int t = 0; t < dmx::op_arity(x.op); t++) x.qb[t] = (uint32_t)q[t];
        c.push_back(x);
    }
    return c;
}

// S <- conj(M) S M^T for one gate: rows of S take M, columns take conj(M).
static void apply_gate(const Gate& g, int n, std::vector<cd>& S)
{
    int k = 0;
    std::vector<uint32_t> qs;
    std::vector<double> re, im;
    dmx::gate_matrix(g, k, qs, re, im);
    const int d = 1 << k;
    const uint64_t dim = 1ull << n;
    uint64_t gm = 0;
    for (uint32_t q : qs) gm |= 1ull << q;
    std::vector<uint64_t> idx(d);
    std::vector<cd> in(d);
    auto vec = [&](uint64_t stride, uint64_t offset, bool conj) {
        for (uint64_t base = 0; base < dim; base++) {
            if (base & gm) continue;
            for (int l = 0; l < d; l++) {
                uint64_t x = base;
                for (int j = 0; j < k; j++)
                    if ((l >> (k - 1 - j)) & 1) x |= 1ull << qs[j];
                idx[l] = x;
                in[l] = S[offset + x * stride];
            }
            for (int r = 0; r < d; r++) {
                cd s = 0.0;
                for (int l = 0; l < d; l++) {
                    cd m(re[(size_t)r * d + l], im[(size_t)r * d + l]);
                    s += (conj ? std::conj(m) : m) * in[l];
                }
                S[offset + idx[r] * stride] = s;
            }
        }
    };
    for (uint64_t r = 0; r < dim; r++) vec(1, r * dim, false);   // each row: M
    for (uint64_t c = 0; c < dim; c++) vec(dim, c, true);        // each column: conj(M)
}

// S <- sum_i conj(K_i) S K_i^T on qubit ch.q (rows with conj(K_i), columns
// with K_i), straight from the Kraus operators.
static void apply_channel(const Channel& ch, int n, std::vector<cd>& S)
{
    const uint64_t dim = 1ull << n, bit = 1ull << ch.q;
    for (uint64_t x0 = 0; x0 < dim; x0++) {
        if (x0 & bit) continue;
        for (uint64_t y0 = 0; y0 < dim; y0++) {
            if (y0 & bit) continue;
            cd in[2][2], out[2][2] = {};
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++) in[a][b] = S[(x0 | (a ? bit : 0)) * dim + (y0 | (b ? bit : 0))];
            for (int i = 0; i < 2; i++)
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++)
                        for (int a2 = 0; a2 < 2; a2++)
                            for (int b2 = 0; b2 < 2; b2++)
                                out[a][b] += std::conj(ch.K[i][2 * a + a2]) * in[a2][b2] *
                                             ch.K[i][2 * b + b2];
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++) S[(x0 | (a ? bit : 0)) * dim + (y0 | (b ? bit : 0))] = out[a][b];
        }
    }
}

// One channel step (k_super) in place: the record's 4x4 on the entry quad of
// row bit q0 and column bit q0, quad index 2 * (row bit) + (column bit).
static void run_channel(const dmx::PrimOp& p, int n, std::vector<cd>& S)
{
    const uint64_t dim = 1ull << n, bit = 1ull << p.q0;
    for (uint64_t x0 = 0; x0 < dim; x0++) {
        if (x0 & bit) continue;
        for (uint64_t y0 = 0; y0 < dim; y0++) {
            if (y0 & bit) continue;
            cd in[4], out[4] = {};
            for (int j = 0; j < 4; j++) in[j] = S[(x0 | ((j >> 1) ? bit : 0)) * dim + (y0 | ((j & 1) ? bit : 0))];
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) out[i] += cd(p.c[2 * (4 * i + j)], p.c[2 * (4 * i + j) + 1]) * in[j];
            for (int i = 0; i < 4; i++) S[(x0 | ((i >> 1) ? bit : 0)) * dim + (y0 | ((i & 1) ? bit : 0))] = out[i];
        }
    }
}

// One in-order five-qubit step (k_fused5) in place: S <- conj(U) S U^T on
// the index bits of blk.qmask, local bit t of U on its t-th lowest bit.
static void run_fused5(const dmx::FusedBlock5& blk, int n, std::vector<cd>& S)
{
    const uint64_t dim = 1ull << n;
    int qs[5], t = 0;
    for (uint32_t m = blk.qmask; m; m &= m - 1) qs[t++] = __builtin_ctz(m);
    auto dep = [&](int l) {
        uint64_t x = 0;
        for (int u = 0; u < 5; u++)
            if ((l >> u) & 1) x |= 1ull << qs[u];
        return x;
    };
    cd in[32];
    auto vec = [&](uint64_t stride, uint64_t offset, bool conj) {
        for (uint64_t base = 0; base < dim; base++) {
            if (base & blk.qmask) continue;
            for (int l = 0; l < 32; l++) in[l] = S[offset + (base | dep(l)) * stride];
            for (int r = 0; r < 32; r++) {
                cd s = 0.0;
                for (int l = 0; l < 32; l++) {
                    const cd u(blk.ur[r * 32 + l], blk.ui[r * 32 + l]);
                    s += (conj ? std::conj(u) : u) * in[l];
                }
                S[offset + (base | dep(r)) * stride] = s;
            }
        }
    };
    for (uint64_t r = 0; r < dim; r++) vec(1, r * dim, false);  // each row: U
    for (uint64_t c = 0; c < dim; c++) vec(dim, c, true);       // each column: conj(U)
}

// One k_fused4_herm launch, thread by thread.
static bool run_herm(const dmx::FusedBlock& blk, int n, const std::vector<cd>& in,
                     std::vector<cd>& out, const char** why)
{
    dmx::HermGeom g;
    g.init(blk, n);
    const uint64_t num = 1ull << (2 * n);
    std::vector<int> hits(num, 0);
    std::vector<cd> lds(16 * dmx::kHermPitch);
    // the launch's schedule (herm_slot) takes every item exactly once
    for (uint32_t grid : {128u, 256u, 4096u}) {
        std::vector<int> seen(g.items(), 0);
        for (uint32_t b = 0; b < grid; b++) {
            uint64_t w = 0;
            for (uint64_t k = 0; dmx::herm_slot(g, b, grid, k, w); k++) {
                if (w >= g.items()) return *why = "schedule item out of range", false;
                seen[w]++;
            }
        }
        for (int v : seen)
            if (v != 1) return *why = "schedule takes an item other than once", false;
    }
    std::vector<dmx::HermThread> th(256);
    for (int tid = 0; tid < 256; tid++) th[tid].init(g, tid);
    for (uint64_t w = 0; w < g.items(); w++) {
        const dmx::HermItem it = dmx::herm_item(g, w);
        for (int tid = 0; tid < 256; tid++)
            for (int rep = 0; rep < 8; rep++) {
                int a;
                const int lc = th[tid].lc;
                const uint64_t off = dmx::herm_load(g, th[tid], it, rep, a);
                if (off + 1 >= num) return *why = "load out of range", false;
                lds[dmx::herm_lds(a, lc)] = in[off];
                lds[dmx::herm_lds(a, lc + 1)] = in[off + 1];
            }
        for (int tb = 0; tb < 16; tb++) {
            if (!dmx::herm_tile_on(it, tb)) continue;
            cd T[16][16], X[16][16];
            for (int a = 0; a < 16; a++)
                for (int j = 0; j < 16; j++) T[a][j] = lds[dmx::herm_lds(a, g.lds_col(tb, j))];
            for (int a = 0; a < 16; a++)  // X = T U^T
                for (int j = 0; j < 16; j++) {
                    cd s = 0.0;
                    for (int k = 0; k < 16; k++) s += T[a][k] * cd(blk.ur[j * 16 + k], blk.ui[j * 16 + k]);
                    X[a][j] = s;
                }
            for (int a = 0; a < 16; a++)  // T' = conj(U) X
                for (int j = 0; j < 16; j++) {
                    cd s = 0.0;
                    for (int k = 0; k < 16; k++) s += cd(blk.ur[a * 16 + k], -blk.ui[a * 16 + k]) * X[k][j];
                    lds[dmx::herm_lds(a, g.lds_col(tb, j))] = s;
                }
        }
        for (int tid = 0; tid < 256; tid++)
            for (int rep = 0; rep < 8; rep++) {
                uint64_t off;
                int a, c0, c1, a0, a1, col;
                if (dmx::herm_direct(g, th[tid], it, rep, off, a, c0, c1)) {
                    if (off + 1 >= num || (off & 1)) return *why = "direct store out of range", false;
                    out[off] = lds[dmx::herm_lds(a, c0)];
                    out[off + 1] = lds[dmx::herm_lds(a, c1)];
                    hits[off]++, hits[off + 1]++;
                }
                if (dmx::herm_mirror(g, th[tid], it, rep, off, a0, a1, col)) {
                    if (off + 1 >= num || (off & 1)) return *why = "mirror store out of range", false;
                    out[off] = std::conj(lds[dmx::herm_lds(a0, col)]);
                    out[off + 1] = std::conj(lds[dmx::herm_lds(a1, col)]);
                    hits[off]++, hits[off + 1]++;
                }
            }
    }
    // every entry the next step reads written once (all of them unless the
    // next step is a Hermitian sweep, which reads (x, y) with x & ~Q' <= y & ~Q'),
    // none twice; the others are left stale and poisoned, so a read of one
    // shows in the result
    const uint64_t dim = 1ull << n;
    const uint64_t keep = (dim - 1) & ~(uint64_t)blk.next_q;
    for (uint64_t i = 0; i < num; i++) {
        const bool needed = blk.next_q == 0 || ((i / dim) & keep) <= ((i % dim) & keep);
        if (hits[i] > 1 || (needed && hits[i] != 1))
            return *why = "entry the next step reads not written exactly once", false;
        if (!hits[i]) out[i] = cd(std::nan(""), std::nan(""));
    }
    return true;
}

// One k_relabel launch, thread by thread.
static bool run_relabel(const dmx::FusedBlock& blk, int n, const std::vector<cd>& in,
                        std::vector<cd>& out, const char** why)
{
    dmx::RelabelGeom g;
    g.init(blk, n);
    const uint64_t num = 1ull << (2 * n);
    std::vector<int> hits(num, 0);
    std::vector<cd> lds(16 * dmx::kHermPitch);
    std::vector<dmx::RelabelThread> th(256);
    for (int tid = 0; tid < 256; tid++) th[tid].init(g, tid);
    for (uint64_t w = 0; w < g.items(); w++) {
        const dmx::RelabelItem it = dmx::relabel_item(g, w);
        for (int tid = 0; tid < 256; tid++)
            for (int rep = 0; rep < 8; rep++) {
                int a;
                const int lc = th[tid].lc;
                const uint64_t off = dmx::relabel_load(g, th[tid], it, tid, rep, a);
                if (off + 1 >= num) return *why = "relabel load out of range", false;
                lds[a * dmx::kHermPitch + lc] = in[off];
                lds[a * dmx::kHermPitch + lc + 1] = in[off + 1];
            }
        for (int tid = 0; tid < 256; tid++)
            for (int rep = 0; rep < 8; rep++) {
                int a;
                const uint64_t off = dmx::relabel_store(g, th[tid], it, tid, rep, a);
                if (off + 1 >= num || (off & 1)) return *why = "relabel store out of range", false;
                out[off] = lds[a * dmx::kHermPitch + th[tid].s0];
                out[off + 1] = lds[a * dmx::kHermPitch + th[tid].s1];
                hits[off]++, hits[off + 1]++;
            }
    }
    for (uint64_t i = 0; i < num; i++)
        if (hits[i] != 1) return *why = "relabel entry not written exactly once", false;
    return true;
}

int main()
{
    int cases = 0;
    for (int n = 8; n <= 10; n++)
        for (int seed = 0; seed < (n == 8 ? 5 : n == 9 ? 3 : 2); seed++) {
            // the last seed of each width: with a C4X halfway
            const bool c4x = seed == (n == 8 ? 4 : n == 9 ? 2 : 1);
            std::vector<Channel> chans;
            const std::vector<Gate> c = circuit(n, 100 * n + seed, c4x, &chans);
            dmx::FusedPlan* plan = nullptr;
            dmx::PlanInfo info;
            if (dmx::build_fused_plan(c, n, 4, nullptr, 0, true, &plan, &info, true) != 0 || !plan)
                return fail("build_fused_plan", n, seed);
            std::vector<dmx::PlanStepView> steps;
            plan->describe(steps);
            if (info.herm_blocks == 0 || info.relabels != 1 || info.perm_blocks == 0)
                return fail("plan has no Hermitian sweeps, relabel or trailing maps", n, seed);
            // a Hermitian sweep writes half exactly when another follows it
            int fused5 = 0;
            for (size_t i = 0; i < steps.size(); i++) {
                fused5 += steps[i].kind == dmx::PlanStepView::FUSED5;
                if (steps[i].kind != dmx::PlanStepView::HERM) continue;
                const bool next_herm = i + 1 < steps.size() && steps[i + 1].kind == dmx::PlanStepView::HERM;
                if ((steps[i].block->next_q != 0) != next_herm)
                    return fail("a Hermitian sweep writes half unless a non-Hermitian step follows", n, seed);
            }
            if (c4x != (fused5 == 1))
                return fail("the C4X is not the plan's one five-qubit step", n, seed);
            int n_chan = 0;
            for (const auto& st : steps) n_chan += st.kind == dmx::PlanStepView::CHANNEL;
            if (n_chan != (int)chans.size()) return fail("a channel is not a step of its own", n, seed);
            const uint64_t dim = 1ull << n;
            std::mt19937 rng(17u + (uint32_t)seed);
            std::normal_distribution<double> g(0.0, 1.0);
            std::vector<cd> S(dim * dim);
            for (uint64_t i = 0; i < dim; i++)
                for (uint64_t j = i; j < dim; j++) {
                    const cd v = i == j ? cd(g(rng), 0.0) : cd(g(rng), g(rng));
                    S[i * dim + j] = v;
                    S[j * dim + i] = std::conj(v);
                }
            std::vector<cd> want = S, cur = S, nxt(dim * dim);
            size_t next_chan = 0;
            for (const Gate& gt : c) {
                if (gt.op == dmx::OP_KRAUS)
                    apply_channel(chans[next_chan++], n, want);
                else
                    apply_gate(gt, n, want);
            }
            for (const auto& st : steps) {
                const char* why = "";
                if (st.kind == dmx::PlanStepView::HERM) {
                    if (!run_herm(*st.block, n, cur, nxt, &why)) return fail(why, n, seed);
                } else if (st.kind == dmx::PlanStepView::RELABEL) {
                    if (!run_relabel(*st.block, n, cur, nxt, &why)) return fail(why, n, seed);
                } else if (st.kind == dmx::PlanStepView::FUSED5) {
                    nxt = cur;
                    run_fused5(*st.block5, n, nxt);
                } else if (st.kind == dmx::PlanStepView::CHANNEL) {
                    nxt = cur;
                    run_channel(*st.prim, n, nxt);
                } else {
                    return fail("unexpected step kind in a Hermitian plan", n, seed);
                }
                cur.swap(nxt);
            }
            double err = 0;
            bool finite = true;
            for (uint64_t i = 0; i < dim * dim; i++) {
                const double d = std::abs(cur[i] - want[i]);
                finite = finite && std::isfinite(d);
                err = std::max(err, d);
            }
            if (!finite || !(err < 1e-12)) {
                std::printf("max |d| = %g\n", err);
                return fail("result differs from gate-by-gate application", n, seed);
            }
            delete plan;
            cases++;
        }
    std::printf("%d cases OK\n", cases);
    return 0;
}
