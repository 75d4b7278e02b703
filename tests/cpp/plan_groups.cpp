// Validity of the fused planner's grouping (dmx::plan_blocks) on seeded random
// circuits, with and without trailing permutations, checked on the host:
// every gate lands in exactly one block, every block holds its matrix
// members' qubits within its width, its trailing members are X, CX or SWAP
// and its stored index map (perm_L, perm_b) is what they do to an index, no
// gate crosses a channel, and the blocks' execution order (which may swap
// commuting gates) applies the same operator as program order -- both orders
// are applied to a random state with the gates' own matrices
// (dmx::gate_matrix) and compared.  Prints "<cases> cases OK", or the first
// violation and exits nonzero.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "fuse.hpp"
#include "lower.hpp"

using dmx::Gate;
using cd = std::complex<double>;

static int fail(const char* what, int n, int k, int seed)
{
    std::printf("FAIL %s (n=%d max_k=%d seed=%d)\n", what, n, k, seed);
    return 1;
}

// What a block's trailing permutation gates do to index x.
static uint32_t permute(const std::vector<Gate>& c, const std::vector<int>& mem, int from,
                        uint32_t x)
{
    for (size_t t = from; t < mem.size(); t++) {
        const Gate& g = c[mem[t]];
        const uint32_t a = g.qb[0], b = g.qb[1];
        if (g.op == dmx::OP_X) {
            x ^= 1u << a;
        } else if (g.op == dmx::OP_CX) {
            if ((x >> a) & 1u) x ^= 1u << b;
        } else {
            const uint32_t ba = (x >> a) & 1u, bb = (x >> b) & 1u;
            x = (x & ~((1u << a) | (1u << b))) | (ba << b) | (bb << a);
        }
    }
    return x;
}

// Random d x d unitary (Gram-Schmidt on Gaussian columns), or a random
// diagonal of phases.
static void random_matrix(std::mt19937& rng, int d, bool diagonal, std::vector<double>& re,
                          std::vector<double>& im)
{
    std::normal_distribution<double> g(0.0, 1.0);
    std::vector<cd> M((size_t)d * d, 0.0);
    if (diagonal) {
        for (int i = 0; i < d; i++) M[(size_t)i * d + i] = std::polar(1.0, g(rng));
    } else {
        for (int c = 0; c < d; c++) {
            std::vector<cd> v(d);
            for (int r = 0; r < d; r++) v[r] = cd(g(rng), g(rng));
            for (int p = 0; p < c; p++) {
                cd dot = 0.0;
                for (int r = 0; r < d; r++) dot += std::conj(M[(size_t)r * d + p]) * v[r];
                for (int r = 0; r < d; r++) v[r] -= dot * M[(size_t)r * d + p];
            }
            double nrm = 0;
            for (int r = 0; r < d; r++) nrm += std::norm(v[r]);
            nrm = std::sqrt(nrm);
            for (int r = 0; r < d; r++) M[(size_t)r * d + c] = v[r] / nrm;
        }
    }
    re.resize(M.size());
    im.resize(M.size());
    for (size_t i = 0; i < M.size(); i++) re[i] = M[i].real(), im[i] = M[i].imag();
}

static std::vector<Gate> random_circuit(int n, int count, int seed, bool channels)
{
    static const int kOps[] = {dmx::OP_U3, dmx::OP_RZ, dmx::OP_U1, dmx::OP_RX, dmx::OP_X,
                               dmx::OP_H, dmx::OP_T, dmx::OP_CX, dmx::OP_CX, dmx::OP_CX,
                               dmx::OP_CZ, dmx::OP_CU1, dmx::OP_RZZ, dmx::OP_RXX, dmx::OP_CRX,
                               dmx::OP_CCX, dmx::OP_C3X, dmx::OP_CSWAP, dmx::OP_ID,
                               dmx::OP_MAT, dmx::OP_MAT};
    constexpr int kNumOps = sizeof(kOps) / sizeof(kOps[0]);
    std::mt19937 rng((uint32_t)seed);
    std::uniform_real_distribution<double> ang(-3.2, 3.2);
    std::vector<Gate> c;
    auto pick = [&](int k, uint32_t* qb) {
        std::vector<int> q(n);
        for (int i = 0; i < n; i++) q[i] = i;
        std::shuffle(q.begin(), q.end(), rng);
        for (int i = 0; i < k; i++) qb[i] = (uint32_t)q[i];
    };
    for (int i = 0; i < count; i++) {
        Gate g;
        if (channels && rng() % 100 < 3) {  // a one-qubit channel: its records are not read here
            g.op = dmx::OP_KRAUS;
            pick(1, g.qb);
            c.push_back(g);
            continue;
        }
        g.op = kOps[rng() % kNumOps];
        g.theta = ang(rng), g.phi = ang(rng), g.lambda = ang(rng);
        if (g.op == dmx::OP_MAT) {
            g.k = 1 + (int)(rng() % (uint32_t)std::min(5, n));
            pick(g.k, g.qb);
            random_matrix(rng, 1 << g.k, rng() % 2 == 0, g.mre, g.mim);
        } else {
            pick(dmx::op_arity(g.op), g.qb);
        }
        c.push_back(g);
    }
    return c;
}

static uint32_t mask_of(const Gate& g)
{
    const int k = g.op == dmx::OP_MAT ? g.k : dmx::op_arity(g.op);
    uint32_t m = 0;
    for (int i = 0; i < k; i++) m |= 1u << g.qb[i];
    return m;
}

// psi <- G psi for one gate (matrix index bit k-1-j <-> qubits[j]).
static void apply(const Gate& g, std::vector<cd>& psi)
{
    int k = 0;
    std::vector<uint32_t> qs;
    std::vector<double> re, im;
    dmx::gate_matrix(g, k, qs, re, im);
    const int d = 1 << k;
    uint64_t gm = 0;
    for (uint32_t q : qs) gm |= 1ull << q;
    std::vector<cd> in(d), out(d);
    std::vector<uint64_t> idx(d);
    for (uint64_t base = 0; base < psi.size(); base++) {
        if (base & gm) continue;
        for (int l = 0; l < d; l++) {
            uint64_t x = base;
            for (int j = 0; j < k; j++)
                if ((l >> (k - 1 - j)) & 1) x |= 1ull << qs[j];
            idx[l] = x;
            in[l] = psi[x];
        }
        for (int r = 0; r < d; r++) {
            cd s = 0.0;
            for (int l = 0; l < d; l++) s += cd(re[(size_t)r * d + l], im[(size_t)r * d + l]) * in[l];
            out[r] = s;
        }
        for (int l = 0; l < d; l++) psi[idx[l]] = out[l];
    }
}

int main()
{
    int cases = 0;
    for (int perm = 0; perm < 3; perm++)  // in place, permuting, Hermitian-sweep rule
    for (int n = 4; n <= 12; n++)
        for (int max_k = 4; max_k <= 5; max_k++)
            for (int seed = 0; seed < 6; seed++) {
                const bool channels = seed % 2 == 1;
                const std::vector<Gate> c = random_circuit(n, 40 + 60 * seed, 1000 * n + seed,
                                                           channels);
                dmx::BlockPlan bp;
                bool fusable = false;
                if (dmx::plan_blocks(c, n, max_k, bp, fusable, perm >= 1, perm == 2) != 0 ||
                    !fusable)
                    return fail("plan_blocks", n, max_k, seed);
                if (bp.dense.size() != bp.blocks.size()) return fail("dense counts", n, max_k, seed);
                // execution sequence of circuit indices (channels included)
                std::vector<int> seq;
                for (const auto& o : bp.order) {
                    if (o.first == 3) {
                        seq.push_back(o.second);
                        continue;
                    }
                    const bool five = o.first == 2;
                    const uint32_t qmask = five ? bp.blocks5[o.second].qmask
                                                : bp.blocks[o.second].qmask;
                    const int width = __builtin_popcount(qmask);
                    if (width != (five ? 5 : 4) || (qmask >> n) != 0)
                        return fail("block width", n, max_k, seed);
                    if (five && max_k < 5) {  // only a five-qubit gate makes one
                        bool wide = false;
                        for (int gi : bp.members5[o.second])
                            wide |= __builtin_popcount(mask_of(c[gi])) == 5;
                        if (!wide) return fail("five-qubit block without a five-qubit gate",
                                               n, max_k, seed);
                    }
                    const auto& mem = five ? bp.members5[o.second] : bp.members[o.second];
                    const int dense = five ? (int)mem.size() : bp.dense[o.second];
                    for (int t = 0; t < (int)mem.size(); t++) {
                        const int gi = mem[t];
                        const int op = c[gi].op;
                        if (t < dense && (mask_of(c[gi]) & ~qmask) != 0)
                            return fail("member outside its block", n, max_k, seed);
                        if (t >= dense && op != dmx::OP_X && op != dmx::OP_CX && op != dmx::OP_SWAP)
                            return fail("trailing member not a permutation gate", n, max_k, seed);
                        seq.push_back(gi);
                    }
                    if (!five) {
                        const dmx::FusedBlock& fb = bp.blocks[o.second];
                        if ((fb.perm_on != 0) != (dense < (int)mem.size()) || (perm == 0 && fb.perm_on))
                            return fail("permutation flag", n, max_k, seed);
                        if (fb.perm_on) {
                            if (perm == 1 && fb.perm_L[0] != 1u)
                                return fail("permutation moves bit 0", n, max_k, seed);
                            if (perm == 2) {  // the block's bits must map onto a set of bits
                                uint32_t img = 0;
                                for (uint32_t m = fb.qmask; m; m &= m - 1)
                                    img |= fb.perm_L[__builtin_ctz(m)];
                                if (__builtin_popcount(img) != 4)
                                    return fail("Hermitian map of the block's bits", n, max_k, seed);
                            }
                            std::mt19937 r2((uint32_t)(seed + 31 * n));
                            for (int trial = 0; trial < 64; trial++) {
                                const uint32_t x = (uint32_t)r2() & ((1u << n) - 1);
                                uint32_t y = fb.perm_b;
                                for (int q = 0; q < n; q++)
                                    if ((x >> q) & 1u) y ^= fb.perm_L[q];
                                if (y != permute(c, mem, dense, x))
                                    return fail("stored index map", n, max_k, seed);
                            }
                        }
                    }
                }
                std::vector<int> seen(c.size(), 0);
                for (int gi : seq) seen[gi]++;
                for (size_t gi = 0; gi < c.size(); gi++)
                    if (seen[gi] != (c[gi].op == dmx::OP_ID ? 0 : 1))
                        return fail("gate placed other than once", n, max_k, seed);
                // no gate on either side of a channel moves across it
                int last_channel = -1, max_before = -1;
                for (int gi : seq) {
                    if (c[gi].op == dmx::OP_KRAUS) {
                        if (max_before > gi) return fail("gate moved past a channel", n, max_k,
                                                         seed);
                        last_channel = gi;
                    } else if (gi < last_channel) {
                        return fail("gate moved before a channel", n, max_k, seed);
                    }
                    max_before = std::max(max_before, gi);
                }
                // the same operator in both orders (channels left out: no gate crosses one)
                std::mt19937 rng(7u + (uint32_t)cases);
                std::normal_distribution<double> g(0.0, 1.0);
                std::vector<cd> a((size_t)1 << n);
                for (cd& v : a) v = cd(g(rng), g(rng));
                std::vector<cd> b = a;
                for (size_t gi = 0; gi < c.size(); gi++)
                    if (c[gi].op != dmx::OP_KRAUS && c[gi].op != dmx::OP_ID) apply(c[gi], a);
                for (int gi : seq)
                    if (c[gi].op != dmx::OP_KRAUS) apply(c[gi], b);
                double err = 0, nrm = 0;
                for (size_t i = 0; i < a.size(); i++) {
                    err = std::max(err, std::abs(a[i] - b[i]));
                    nrm = std::max(nrm, std::abs(a[i]));
                }
                if (!(err <= 1e-10 * nrm)) return fail("execution order changes the operator",
                                                       n, max_k, seed);
                cases++;
            }
    std::printf("%d cases OK\n", cases);
    return 0;
}
