// Validity of the fused planner's grouping (dmx::plan_blocks) on seeded random
// circuits, checked on the host: every gate lands in exactly one block, the
// gates on each qubit keep their program order in the blocks' execution
// order, no gate crosses a channel, and every block holds its members'
// qubits within its width.  Prints "<cases> cases OK", or the first violation
// and exits nonzero.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "fuse.hpp"
#include "lower.hpp"

using dmx::Gate;

static int fail(const char* what, int n, int k, int seed)
{
    std::printf("FAIL %s (n=%d max_k=%d seed=%d)\n", what, n, k, seed);
    return 1;
}

static std::vector<Gate> random_circuit(int n, int count, int seed, bool channels)
{
    std::mt19937 rng((uint32_t)seed);
    std::vector<Gate> c;
    auto pick = [&](int k, uint32_t* qb) {
        std::vector<int> q(n);
        for (int i = 0; i < n; i++) q[i] = i;
        std::shuffle(q.begin(), q.end(), rng);
        for (int i = 0; i < k; i++) qb[i] = (uint32_t)q[i];
    };
    for (int i = 0; i < count; i++) {
        Gate g;
        const int r = (int)(rng() % 100);
        if (channels && r < 3) {  // a one-qubit channel: its records are not read here
            g.op = dmx::OP_KRAUS;
            pick(1, g.qb);
        } else if (r < 45) {
            g.op = dmx::OP_U3;
            pick(1, g.qb);
        } else if (r < 80) {
            g.op = dmx::OP_CX;
            pick(2, g.qb);
        } else if (r < 88) {
            g.op = dmx::OP_CCX;
            pick(3, g.qb);
        } else if (r < 92) {
            g.op = dmx::OP_ID;
            pick(1, g.qb);
        } else {  // dense matrix on 2..min(5, n) qubits
            g.op = dmx::OP_MAT;
            g.k = 2 + (int)(rng() % (uint32_t)(std::min(5, n) - 1));
            pick(g.k, g.qb);
            const int d = 1 << g.k;
            g.mre.assign((size_t)d * d, 0.0);
            g.mim.assign((size_t)d * d, 0.0);
            for (int j = 0; j < d; j++) g.mre[(size_t)j * d + j] = 1.0;
        }
        g.theta = 0.1 * i;
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

int main()
{
    int cases = 0;
    for (int n = 4; n <= 12; n++)
        for (int max_k = 4; max_k <= 5; max_k++)
            for (int seed = 0; seed < 6; seed++) {
                const bool channels = seed % 2 == 1;
                const std::vector<Gate> c = random_circuit(n, 40 + 60 * seed, 1000 * n + seed,
                                                           channels);
                dmx::BlockPlan bp;
                bool fusable = false;
                if (dmx::plan_blocks(c, n, max_k, bp, fusable) != 0 || !fusable)
                    return fail("plan_blocks", n, max_k, seed);
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
                        for (int gi : bp.members5[o.second]) wide |= c[gi].k == 5;
                        if (!wide) return fail("five-qubit block without a five-qubit gate",
                                               n, max_k, seed);
                    }
                    const auto& mem = five ? bp.members5[o.second] : bp.members[o.second];
                    for (int gi : mem) {
                        if ((mask_of(c[gi]) & ~qmask) != 0)
                            return fail("member outside its block", n, max_k, seed);
                        seq.push_back(gi);
                    }
                }
                std::vector<int> seen(c.size(), 0);
                for (int gi : seq) seen[gi]++;
                for (size_t gi = 0; gi < c.size(); gi++)
                    if (seen[gi] != (c[gi].op == dmx::OP_ID ? 0 : 1))
                        return fail("gate placed other than once", n, max_k, seed);
                // per-qubit program order, and no gate on either side of a channel moves
                std::vector<int> last(n, -1);
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
                    for (uint32_t b = mask_of(c[gi]); b; b &= b - 1) {
                        const int q = __builtin_ctz(b);
                        if (gi < last[q]) return fail("per-qubit order", n, max_k, seed);
                        last[q] = gi;
                    }
                }
                cases++;
            }
    std::printf("%d cases OK\n", cases);
    return 0;
}
