// fuse.cpp -- fusion planner and fused executor (host side).
//
// Planner: gates are taken in circuit order; a gate joins an existing block
// when the union of qubits stays within the width limit and no later block
// touches its qubits (so moving it there only reorders gates on disjoint
// qubits, which commute).  Among the admissible blocks the one that grows the
// least is chosen.  A block of at most four qubits is padded to four with the
// lowest qubits it does not use (identity there); a five-qubit block (a
// five-qubit gate such as C4X, or max_k = 5) runs through the 32 x 32 kernel.
// A block's matrix is the product of the member gates' matrices, each
// composed from the gate's own primitive sweeps (so C4X, C3SQRTX, R keep the
// reference's semantics).
//
// Execution choice: without a cost model every block runs as one fused sweep
// at the default tile width.  With one, every block is priced (one batched
// inference over all candidates) as a fused sweep at each tile width and as
// its member gates' primitive sweeps on both indices, and the cheapest
// prediction is taken.
#include "fuse.hpp"

#include <algorithm>
#include <complex>
#include <cstring>
#include <exception>
#include <map>
#include <queue>
#include <system_error>
#include <thread>
#include <tuple>

#include "dmsim_amd.h"
#include "engine.hpp"
#include "fused.hpp"
#include "mlp.hpp"
#include "sweep.hpp"

namespace dmx {

namespace {

using cd = std::complex<double>;

// A run is a sequence of steps: a fused four- or five-qubit block (one sweep
// each); a plain segment, the primitive sweeps of a block the cost model runs
// unfused, applied to the fast index and then to the slow index (valid for
// complex-linear gates, whose row and column actions commute); or a channel
// (one superoperator sweep, in the record for rho's current orientation).
// Hermitian plans add two out-of-place steps: a four-qubit block through
// k_fused4_herm (its record rewritten for the bit order rho has then), and a
// relabelling of rho's index bits (k_relabel) back to the circuit's order
// before any other step and at the end of the run.
enum StepType {
    ST_FUSED4 = 0,
    ST_PLAIN = 1,
    ST_FUSED5 = 2,
    ST_CHANNEL = 3,
    ST_HERM = 4,
    ST_RELABEL = 5
};

struct Step {
    int type;
    int index;  // block index (blocks or blocks5), or first primitive of the segment
    int count;  // primitives in the segment
};

struct Plan : FusedPlan {
    int n = 0;
    std::vector<FusedBlock> blocks;
    std::vector<FusedBlock5> blocks5;
    std::vector<PrimOp> prims;
    std::vector<Step> steps;
    FusedBlock* d_blocks = nullptr;
    FusedBlock5* d_blocks5 = nullptr;
    PrimOp* d_prims = nullptr;
    int max_blocks = kFusedGridCap;

    ~Plan() override
    {
        if (d_blocks) (void)hipFree(d_blocks);
        if (d_blocks5) (void)hipFree(d_blocks5);
        if (d_prims) (void)hipFree(d_prims);
    }

    template <class T>
    static int put(T** dst, const std::vector<T>& src, hipStream_t stream)
    {
        if (src.empty()) return DMX_OK;
        hipError_t e = hipMalloc(dst, src.size() * sizeof(T));
        if (e != hipSuccess) {
            set_error(std::string("hipMalloc(fused plan): ") + hipGetErrorString(e));
            return e == hipErrorOutOfMemory ? DMX_ENOMEM : DMX_EHIP;
        }
        e = hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice,
                           stream);
        if (e != hipSuccess) {
            set_error(std::string("hipMemcpyAsync(fused plan): ") + hipGetErrorString(e));
            return DMX_EHIP;
        }
        return DMX_OK;
    }

    bool needs_alt() const override
    {
        for (const Step& st : steps)
            if ((st.type == ST_FUSED4 && blocks[st.index].perm_on) || st.type == ST_HERM ||
                st.type == ST_RELABEL)
                return true;
        return false;
    }

    int autotune_tiles(int nq, double* re, double* im, hipStream_t stream, int* chosen,
                       PlanInfo& info) override
    {
        *chosen = 0;
        std::vector<int> sample;
        for (const Step& st : steps)
            if (st.type == ST_FUSED4 && sample.size() < 8) sample.push_back(st.index);
        if (sample.empty()) return DMX_OK;
        bool all_q0 = true;
        for (int b : sample) all_q0 = all_q0 && (blocks[b].qmask & 1u);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipError_t e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        int best_m = 0;
        float best_ms = 0.f;
        for (int m = all_q0 ? 4 : 5; m <= std::min(8, nq) && e == hipSuccess; m++) {
            for (int rep = 0; rep < 4 && e == hipSuccess; rep++) {  // rep 0 warms up
                if (rep == 1) e = hipEventRecord(e0, stream);
                for (int b : sample)
                    if (e == hipSuccess)
                        e = launch_fused4(re, im, d_blocks + b, nq, m, max_blocks, stream);
            }
            if (e == hipSuccess) e = hipEventRecord(e1, stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e == hipSuccess && (best_m == 0 || ms < best_ms)) {
                best_m = m;
                best_ms = ms;
            }
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (e != hipSuccess) {
            set_error(std::string("tile-width timing: ") + hipGetErrorString(e));
            return DMX_EHIP;
        }
        for (int m = 4; m <= 8; m++) info.tile_hist[m] = 0;
        for (const Step& st : steps)
            if (st.type == ST_FUSED4) {
                FusedBlock& fb = blocks[st.index];
                if (tile_bits_ok(nq, best_m, fb.qmask)) fb.m = (uint32_t)best_m;
                info.tile_hist[fb.m]++;
            }
        *chosen = best_m;
        return DMX_OK;
    }

    void describe(std::vector<PlanStepView>& out) const override
    {
        out.clear();
        for (const Step& st : steps) {
            PlanStepView v;
            v.kind = st.type == ST_HERM      ? PlanStepView::HERM
                     : st.type == ST_RELABEL ? PlanStepView::RELABEL
                     : st.type == ST_FUSED4  ? PlanStepView::FUSED4
                     : st.type == ST_FUSED5  ? PlanStepView::FUSED5
                     : st.type == ST_CHANNEL ? PlanStepView::CHANNEL
                                             : PlanStepView::OTHER;
            v.block = (st.type == ST_HERM || st.type == ST_RELABEL || st.type == ST_FUSED4)
                          ? &blocks[st.index]
                          : nullptr;
            v.block5 = st.type == ST_FUSED5 ? &blocks5[st.index] : nullptr;
            v.prim = st.type == ST_CHANNEL ? &prims[st.index] : nullptr;
            out.push_back(v);
        }
    }

    double bytes(int nq) const override
    {
        const double sweep = 32.0 * (double)(1ull << (2 * nq));
        double b = 0;
        for (const Step& st : steps) {
            if (st.type == ST_HERM) {
                b += herm_bytes(nq, blocks[st.index].next_q != 0);
            } else if (st.type == ST_RELABEL || st.type == ST_FUSED4 || st.type == ST_FUSED5) {
                b += sweep;
            } else if (st.type == ST_CHANNEL) {
                b += prim_bytes(prims[st.index], nq);
            } else {
                for (int t = st.index; t < st.index + st.count; t++)
                    b += 2.0 * prim_bytes(prims[t], nq);
            }
        }
        return b;
    }

    int upload(hipStream_t stream) override
    {
        int rc = put(&d_blocks, blocks, stream);
        if (!rc) rc = put(&d_blocks5, blocks5, stream);
        return rc ? rc : put(&d_prims, prims, stream);
    }

    int execute(double*& re, double*& im, double*& re2, double*& im2, int nq, bool transposed,
                hipStream_t stream, LaunchProfiler& prof, uint32_t* launches,
                uint32_t* sweeps) override
    {
        const double sweep_bytes = 32.0 * (double)(1ull << (2 * nq));
        uint32_t la = 0;
        for (const Step& st : steps) {
            hipError_t e = hipSuccess;
            if (st.type == ST_HERM || st.type == ST_RELABEL) {
                if (!re2 || !im2) {
                    set_error("internal: an out-of-place step needs the second pair of planes");
                    return DMX_EINVAL;
                }
                const bool h = st.type == ST_HERM;
                prof.begin(h ? KC_FUSED4H : KC_RELABEL,
                           h ? herm_bytes(nq, blocks[st.index].next_q != 0) : sweep_bytes);
                e = h ? launch_fused4_herm(re, im, re2, im2, d_blocks + st.index, nq, max_blocks,
                                           stream)
                      : launch_relabel(re, im, re2, im2, d_blocks + st.index, nq, max_blocks,
                                       stream);
                prof.end();
                std::swap(re, re2);
                std::swap(im, im2);
                la++;
            } else if (st.type == ST_FUSED4 && blocks[st.index].perm_on) {
                if (!re2 || !im2) {
                    set_error("internal: a permuting block needs the second pair of planes");
                    return DMX_EINVAL;
                }
                prof.begin(KC_FUSED4P, sweep_bytes);
                e = launch_fused4_perm(re, im, re2, im2, d_blocks + st.index, nq,
                                       (int)blocks[st.index].m, max_blocks, stream);
                prof.end();
                std::swap(re, re2);
                std::swap(im, im2);
                la++;
            } else if (st.type == ST_FUSED4) {
                prof.begin(KC_FUSED4, sweep_bytes);
                e = launch_fused4(re, im, d_blocks + st.index, nq, (int)blocks[st.index].m,
                                  max_blocks, stream);
                prof.end();
                la++;
            } else if (st.type == ST_CHANNEL) {
                const int t = st.index + (transposed ? 1 : 0);  // record for this orientation
                prof.begin(KC_QUAD, prim_bytes(prims[t], nq));
                e = launch_plain_one(re, im, d_prims + t, prims[t], nq, false, kSweepGridCap,
                                     stream);
                prof.end();
                la++;
            } else if (st.type == ST_FUSED5) {
                prof.begin(KC_FUSED5, sweep_bytes);
                e = launch_fused5(re, im, d_blocks5 + st.index, nq, (int)blocks5[st.index].m,
                                  max_blocks, stream);
                prof.end();
                la++;
            } else {
                for (int side = 0; side < 2 && e == hipSuccess; side++)
                    for (int t = st.index; t < st.index + st.count && e == hipSuccess; t++) {
                        prof.begin(prim_class(prims[t]), prim_bytes(prims[t], nq));
                        e = launch_plain_one(re, im, d_prims + t, prims[t], nq, side == 1,
                                             kSweepGridCap, stream);
                        prof.end();
                        la++;
                    }
            }
            if (e != hipSuccess) {
                set_error(std::string("fused run launch: ") + hipGetErrorString(e));
                return DMX_EHIP;
            }
        }
        *launches = la;
        *sweeps = la;
        return DMX_OK;
    }
};

uint32_t gate_mask(const Gate& g)
{
    const int k = g.op == OP_MAT ? g.k : op_arity(g.op);
    uint32_t m = 0;
    for (int i = 0; i < k; i++) m |= 1u << g.qb[i];
    return m;
}

// Left-multiplies the D x D block matrix U (D = 2^popcount(qmask)) by gate G
// embedded on the block's local bits.  G's index bit (kg-1-j) belongs to
// qubit gq[j].
void apply_embedded(std::vector<cd>& U, uint32_t qmask, const std::vector<uint32_t>& gq,
                    const std::vector<double>& gre, const std::vector<double>& gim)
{
    const int D = 1 << __builtin_popcount(qmask);
    int local[32];
    {
        int t = 0;
        for (int q = 0; q < 32; q++)
            if (qmask & (1u << q)) local[q] = t++;
    }
    const int kg = (int)gq.size();
    const int dg = 1 << kg;
    uint32_t gl = 0;  // gate's local bit mask
    for (int jj = 0; jj < kg; jj++) gl |= 1u << local[gq[jj]];
    auto sub = [&](int x) {
        int s = 0;
        for (int jj = 0; jj < kg; jj++)
            if (x & (1 << local[gq[jj]])) s |= 1 << (kg - 1 - jj);
        return s;
    };
    std::vector<cd> out((size_t)D * D, 0.0);
    for (int xo = 0; xo < D; xo++) {
        const int so = sub(xo);
        for (int xi = 0; xi < D; xi++) {
            if ((xo & ~gl) != (xi & ~gl)) continue;
            const int si = sub(xi);
            const cd gv(gre[(size_t)so * dg + si], gim[(size_t)so * dg + si]);
            if (gv == 0.0) continue;
            for (int c = 0; c < D; c++) out[(size_t)xo * D + c] += gv * U[(size_t)xi * D + c];
        }
    }
    U.swap(out);
}

// A block under construction: qubit mask and member gates (circuit indices,
// in application order).
struct PB {
    uint32_t mask;            // qubits of the block's matrix gates
    std::vector<int> gates;
    int dense = -1;           // gates[0, dense) form the matrix, the rest are a trailing
                              // permutation; -1: all of them
};
using Seq = std::vector<std::pair<int, int>>;  // (0, index into blocks) or (3, channel gate)

// Index-bit order of rho along a Hermitian plan.  sigma[q] is the index bit
// holding logical qubit q, sinv its inverse (the identity at the start and
// after every relabelling).
struct BitOrder {
    std::vector<int> sigma, sinv;
    explicit BitOrder(int n) : sigma(n), sinv(n)
    {
        for (int q = 0; q < n; q++) sigma[q] = sinv[q] = q;
    }
    bool home() const
    {
        for (size_t q = 0; q < sigma.size(); q++)
            if (sigma[q] != (int)q) return false;
        return true;
    }
    void go_home()
    {
        for (size_t q = 0; q < sigma.size(); q++) sigma[q] = sinv[q] = (int)q;
    }
    uint32_t phys(uint32_t logical) const
    {
        uint32_t o = 0;
        for (size_t q = 0; q < sigma.size(); q++)
            if ((logical >> q) & 1u) o |= 1u << sigma[q];
        return o;
    }
};

// Linear part of a block's trailing permutation (its X, CX and SWAP gates
// after the matrix gates) by columns on logical qubits: Lp[q] = pi(2^q) ^
// pi(0); the identity without trailing gates.  n <= 16.
void trailing_linear(const std::vector<Gate>& circuit, const PB& pb, int n, uint32_t* Lp)
{
    auto pi = [&](uint32_t x) {
        if (pb.dense < 0) return x;
        for (size_t t = pb.dense; t < pb.gates.size(); t++) {
            const Gate& g = circuit[pb.gates[t]];
            const uint32_t a = g.qb[0], c = g.qb[1];
            if (g.op == OP_X) {
                x ^= 1u << a;
            } else if (g.op == OP_CX) {
                x ^= ((x >> a) & 1u) << c;
            } else {  // SWAP
                const uint32_t d = ((x >> a) ^ (x >> c)) & 1u;
                x ^= (d << a) | (d << c);
            }
        }
        return x;
    };
    const uint32_t b0 = pi(0);
    for (int q = 0; q < 16; q++) Lp[q] = q < n ? pi(1u << q) ^ b0 : 0u;
}

// A Hermitian sweep of the four-qubit block on logical qubits Q with trailing
// map pi (linear part Lp, by columns, on logical qubits) under bit order bo:
// false if pi(span Q) is no set of bits (the block then runs in order);
// otherwise *qphys = the index bits the sweep reads the block on and the order
// it writes: the qubits spanning pi(span Q) at bits 0..3 (ascending), then the
// images of the super-tile's four low free bits, then the rest by their
// current bit.
bool herm_step(int n, const BitOrder& bo, uint32_t Q, const uint32_t* Lp, uint32_t* qphys,
               BitOrder* next)
{
    uint32_t img = 0;
    for (uint32_t m = Q; m; m &= m - 1) img |= Lp[__builtin_ctz(m)];
    if (__builtin_popcount(img) != 4) return false;
    const uint32_t qp = bo.phys(Q);
    *qphys = qp;
    if (!next) return true;
    uint32_t cm = qp;
    for (int q = 0; q < n && __builtin_popcount(cm) < 8; q++) cm |= 1u << q;
    std::vector<int> order;
    std::vector<char> placed(n, 0);
    auto place = [&](int q) {
        if (!placed[q]) {
            placed[q] = 1;
            order.push_back(q);
        }
    };
    for (int q = 0; q < n; q++)
        if ((img >> q) & 1u) place(q);
    for (int p = 0; p < n; p++)
        if (((cm & ~qp) >> p) & 1u) {
            const uint32_t t = Lp[bo.sinv[p]];
            if (__builtin_popcount(t) == 1) place(__builtin_ctz(t));
        }
    for (int p = 0; p < n; p++) place(bo.sinv[p]);
    for (int i = 0; i < n; i++) {
        next->sigma[order[i]] = i;
        next->sinv[i] = order[i];
    }
    return true;
}

// Grouping 1, in program order: each gate joins the open block that must take
// the fewest new qubits (the latest such block on ties) among the blocks at or
// after the last one touching its qubits -- the blocks in between do not touch
// them, so the gate commutes back to that point -- or opens a new block.
void group_least_growth(const std::vector<Gate>& circuit, int kmax, std::vector<PB>& pbs, Seq& seq)
{
    int frozen = 0;  // blocks before a channel take no later gate
    for (int gi = 0; gi < (int)circuit.size(); gi++) {
        const Gate& g = circuit[gi];
        if (g.op == OP_ID) continue;
        if (g.op == OP_KRAUS) {
            seq.push_back({3, gi});
            frozen = (int)pbs.size();
            continue;
        }
        const uint32_t S = gate_mask(g);
        const int k_here = __builtin_popcount(S);  // <= 5: widest gate or matrix
        int jstar = frozen - 1;  // last block touching the gate's qubits
        for (int b = (int)pbs.size() - 1; b >= frozen; b--)
            if (pbs[b].mask & S) {
                jstar = b;
                break;
            }
        int best = -1, best_g = 99;
        for (int b = std::max(jstar, frozen); b < (int)pbs.size(); b++) {
            const int have = __builtin_popcount(pbs[b].mask);
            const int w = __builtin_popcount(pbs[b].mask | S);
            if (w <= std::max(std::max(kmax, k_here), have) && w - have <= best_g) {
                best = b;
                best_g = w - have;
            }
        }
        if (best >= 0) {
            pbs[best].mask |= S;
            pbs[best].gates.push_back(gi);
        } else {
            seq.push_back({0, (int)pbs.size()});
            pbs.push_back({S, {gi}});
        }
    }
}

// Per-qubit commutation classes of a gate, from its own matrix: bit q of zm
// (xm) is set when the gate is block-diagonal in the Z (X) basis of qubit q,
// i.e. of the form sum_k P_k(q) (x) G_k with P_k the projectors of that basis.
// Two gates whose every shared qubit has a common class commute: on each
// shared qubit both are sums over the same projectors, and what remains acts
// on disjoint qubits.
void commute_classes(const Gate& g, uint32_t& zm, uint32_t& xm)
{
    zm = xm = 0;
    int k = 0;
    std::vector<uint32_t> qs;
    std::vector<double> re, im;
    gate_matrix(g, k, qs, re, im);
    const int D = 1 << k;
    constexpr double kTol = 1e-12;
    auto at = [&](int r, int c) { return cd(re[(size_t)r * D + c], im[(size_t)r * D + c]); };
    for (int j = 0; j < k; j++) {
        const int b = 1 << (k - 1 - j);  // matrix index bit of operand j
        bool z = true, x = true;
        for (int r = 0; r < D && (z || x); r++)
            for (int c = 0; c < D && (z || x); c++) {
                if (((r ^ c) & b) == 0) continue;
                if (std::abs(at(r, c)) > kTol) z = false;
                // (H_b M H_b)[r][c] = 1/2 sum_{a,d} (-1)^(r_b a + d c_b) M[r_a][c_d]
                const int r0 = r & ~b, c0 = c & ~b;
                const double sr = (r & b) ? -1.0 : 1.0, sc = (c & b) ? -1.0 : 1.0;
                const cd h = 0.5 * (at(r0, c0) + sc * at(r0, c0 | b) + sr * at(r0 | b, c0) +
                                    sr * sc * at(r0 | b, c0 | b));
                if (std::abs(h) > kTol) x = false;
            }
        if (z) zm |= 1u << qs[j];
        if (x) xm |= 1u << qs[j];
    }
}

// Grouping 2, over the gates' dependency order: a gate depends on every
// earlier gate that shares a qubit and does not commute with it (classes
// above).  Blocks are built one at a time from the ready gates (all their
// dependencies grouped).  A block on qubit set Q absorbs, repeatedly, every
// ready gate inside Q; Q starts as one ready gate's qubits and may take a
// further ready gate's qubits while it stays within kmax.  The absorption
// order is an order of the circuit's gates that differs from program order
// only by swaps of commuting gates, so the blocks' product is the circuit's.
// Every Q reached this way (best first, memoised, at most kMaxExplored per
// search) is a candidate, scored by the gates it absorbs: 100 per multi-qubit
// gate, 1 per one-qubit gate (those join almost any block later).  With
// allow_perm, a chosen four-qubit block also takes the X, CX and SWAP gates
// that become ready after it, as the permutation of its output (they follow
// all of its matrix gates, so the order is still a valid one).  The best few
// candidates (kLook, kLookDeep from 12 qubits), each with the permutation
// gates it would take, are compared by
// their score plus the best score of the block that would follow them; the
// smaller Q wins ties.  A gate is looked for at most kWindow places past the
// first ungrouped gate of each qubit.
class FrontierGrouper {
public:
    FrontierGrouper(const std::vector<Gate>& circuit, int n, int kmax, bool allow_perm, bool herm)
        : circuit_(circuit), n_(n), kmax_(kmax), herm_(herm),
          look_(n >= kHermAutoQubits ? kLookDeep : kLook), lists_(n), head_(n, 0),
          mask_(circuit.size(), 0), zm_(circuit.size(), 0), xm_(circuit.size(), 0),
          pos0_(circuit.size(), 0), done_(circuit.size(), 0), mark_(circuit.size(), 0),
          listed_(circuit.size(), 0), base_(n), work_(n)
    {
        // X, CX and SWAP join as index maps only where their own matrices are
        // exactly those permutations (operand 0 = the matrix index's high bit)
        if (allow_perm && n <= 16) {
            perm_op_[OP_X] = is_exact_permutation(OP_X, [](int l) { return l ^ 1; });
            perm_op_[OP_CX] = is_exact_permutation(OP_CX, [](int l) { return l & 2 ? l ^ 1 : l; });
            perm_op_[OP_SWAP] = is_exact_permutation(OP_SWAP, [](int l) {
                return ((l & 1) << 1) | ((l >> 1) & 1);
            });
        }
    }

    // Groups circuit indices seg (no channels, no identities) into pbs/seq.
    void run(const std::vector<int>& seg, std::vector<PB>& pbs, Seq& seq)
    {
        for (auto& l : lists_) l.clear();
        std::fill(head_.begin(), head_.end(), 0);
        pos_.clear();
        for (int gi : seg) {
            const uint32_t m = gate_mask(circuit_[gi]);
            mask_[gi] = m;
            classes_of(circuit_[gi], zm_[gi], xm_[gi]);
            pos0_[gi] = (int)pos_.size();
            for (uint32_t b = m; b; b &= b - 1) {
                const int q = __builtin_ctz(b);
                pos_.push_back((int)lists_[q].size());
                lists_[q].push_back(gi);
            }
        }
        // lookahead costs about kLook + 1 searches per block: skipped for very
        // long segments to bound planning time
        const bool small = seg.size() <= kLookMaxGates;
        const int look = small ? look_ : 0;
        const int cap = small ? kMaxExplored : kMaxExploredLong;
        size_t left = seg.size();
        while (left > 0) {
            std::vector<Cand> c1;
            search({}, cap, c1);
            std::sort(c1.begin(), c1.end(), ranked);
            const Cand* best = &c1[0];
            const bool perms = perm_op_[OP_X] || perm_op_[OP_CX] || perm_op_[OP_SWAP];
            // a candidate with the permutation gates it would take after it
            auto with_perms = [&](const Cand& c, PB& pb, uint32_t& touched) {
                pb = PB{0, c.taken};
                for (int gi : pb.gates) pb.mask |= mask_[gi];
                touched = pb.mask;
                if (perms) absorb_permutations(pb, touched);
            };
            PB pb;
            uint32_t touched = 0;
            with_perms(*best, pb, touched);
            if (look > 0 && pb.gates.size() < left) {
                long best_total = -1;
                for (int i = 0; i < (int)c1.size() && i < look; i++) {
                    PB cand;
                    uint32_t t = 0;
                    with_perms(c1[i], cand, t);
                    long total = c1[i].score;
                    for (size_t k = c1[i].taken.size(); k < cand.gates.size(); k++)
                        total += weight(cand.gates[k]);
                    if (cand.gates.size() == left) {
                        total += 1L << 40;  // finishes the segment
                    } else {
                        std::vector<Cand> c2;
                        search(cand.gates, kMaxExplored2, c2);
                        long s2 = 0;
                        for (const Cand& c : c2) s2 = std::max(s2, c.score);
                        total += s2;
                    }
                    if (total > best_total) {
                        best_total = total;
                        pb = std::move(cand);
                        touched = t;
                    }
                }
            }
            for (int gi : pb.gates) done_[gi] = 1;
            for (uint32_t b = touched; b; b &= b - 1) {
                const int q = __builtin_ctz(b);
                while (head_[q] < (int)lists_[q].size() && done_[lists_[q][head_[q]]]) head_[q]++;
            }
            left -= pb.gates.size();
            seq.push_back({0, (int)pbs.size()});
            pbs.push_back(std::move(pb));
        }
    }

private:
    struct Cand {
        uint32_t Q;
        long score;               // 100 per multi-qubit gate absorbed, 1 per one-qubit gate
        std::vector<int> taken;   // the search's starting set, then the gates absorbed
    };

    // Flags of the ungrouped gates before a place in a qubit's list: some
    // gate, one not Z-diagonal there, one not X-diagonal, one neither.
    enum : uint8_t { F_ANY = 1, F_NZ = 2, F_NX = 4, F_GEN = 8 };

    static bool ranked(const Cand& a, const Cand& b)
    {
        if (a.score != b.score) return a.score > b.score;
        const int wa = __builtin_popcount(a.Q), wb = __builtin_popcount(b.Q);
        if (wa != wb) return wa < wb;
        return a.Q < b.Q;
    }

    void classes_of(const Gate& g, uint32_t& zm, uint32_t& xm)
    {
        if (g.op == OP_MAT) {
            commute_classes(g, zm, xm);
            return;
        }
        const auto key = std::make_tuple((int)g.op, g.theta, g.phi, g.lambda);
        auto it = class_cache_.find(key);
        if (it == class_cache_.end()) {
            if (class_cache_.size() >= kClassCacheMax) class_cache_.clear();  // distinct angles
            // classes with operand j on qubit j, then kept per operand position
            Gate p = g;
            const int k = op_arity(g.op);
            for (int j = 0; j < k; j++) p.qb[j] = (uint32_t)j;
            uint32_t z = 0, x = 0;
            commute_classes(p, z, x);
            it = class_cache_.emplace(key, std::make_pair(z, x)).first;
        }
        zm = xm = 0;
        const int k = op_arity(g.op);
        for (int j = 0; j < k; j++) {
            if ((it->second.first >> j) & 1u) zm |= 1u << g.qb[j];
            if ((it->second.second >> j) & 1u) xm |= 1u << g.qb[j];
        }
    }

    bool taken(int gi) const { return done_[gi] || mark_[gi] == stamp_; }

    long weight(int gi) const
    {
        if (weights_) return (*weights_)[gi];
        return __builtin_popcount(mask_[gi]) > 1 ? 100 : 1;
    }

public:
    // Gate weights for the candidates' scores (circuit index -> weight); the
    // default is 100 per multi-qubit gate and 1 per one-qubit gate.
    void set_weights(const std::vector<long>* w) { weights_ = w; }

private:

    // Whether op's matrix (operands on qubits 0, 1) is exactly the permutation
    // local index l -> f(l).
    template <class F>
    static bool is_exact_permutation(int op, F f)
    {
        Gate g;
        g.op = op;
        for (int j = 0; j < 5; j++) g.qb[j] = (uint32_t)j;
        int k = 0;
        std::vector<uint32_t> qs;
        std::vector<double> re, im;
        gate_matrix(g, k, qs, re, im);
        const int D = 1 << k;
        for (int r = 0; r < D; r++)
            for (int c = 0; c < D; c++)
                if (re[(size_t)r * D + c] != (f(c) == r ? 1.0 : 0.0) || im[(size_t)r * D + c] != 0.0)
                    return false;
        return true;
    }

    // A permutation gate joins a block's index map only if the map keeps the
    // sweep's stores whole: bit 0 stays in place (column pairs), and no
    // condition bit inside a contiguous column run of the block's super-tile
    // (bits below r) moves a bit outside it.
    bool keeps_runs(int gi, int r) const
    {
        const Gate& g = circuit_[gi];
        const int a = (int)g.qb[0], b = (int)g.qb[1];
        switch (g.op) {
        case OP_X: return true;
        case OP_CX: return a != 0 && !(a < r && b >= r);
        case OP_SWAP: return a != 0 && b != 0 && ((a < r) == (b < r));
        default: return false;
        }
    }

    // Hermitian sweeps: the index map must send the span of the block's bits
    // (img: where the trailing gates so far have moved them) onto a set of
    // index bits.  X and SWAP always do (SWAP moves img); CX does unless its
    // control is in img and its target is not.
    bool keeps_bits(int gi, uint32_t& img) const
    {
        const Gate& g = circuit_[gi];
        const uint32_t a = 1u << g.qb[0], b = 1u << g.qb[1];
        switch (g.op) {
        case OP_X: return true;
        case OP_CX: return !((img & a) && !(img & b));
        case OP_SWAP:
            if (((img & a) != 0) != ((img & b) != 0)) img ^= a | b;
            return true;
        default: return false;
        }
    }

    // Appends to a four-qubit block the permutation gates that are ready
    // after it (repeatedly), as its trailing index map; touched gains their
    // qubits.
    void absorb_permutations(PB& pb, uint32_t& touched)
    {
        uint32_t qm = pb.mask;
        for (int q = 0; q < n_ && __builtin_popcount(qm) < 4; q++) qm |= 1u << q;
        if (__builtin_popcount(qm) != 4) return;
        uint32_t img = qm;
        uint32_t cm = qm;  // super-tile column bits at the default width
        for (int q = 0; q < n_ && __builtin_popcount(cm) < default_tile_bits(n_); q++)
            cm |= 1u << q;
        const int r = __builtin_ctz(~cm);
        const size_t dense = pb.gates.size();
        const uint32_t all = n_ >= 32 ? ~0u : (1u << n_) - 1;
        ++stamp_;
        for (int gi : pb.gates) mark_[gi] = stamp_;
        for (bool grew = true; grew;) {
            grew = false;
            for (int q = 0; q < n_; q++) prefix(q, work_[q]);
            std::vector<int> ready;
            ready_on(all, ready);
            for (int h : ready) {
                const int op = circuit_[h].op;
                if (op >= OP_COUNT || !perm_op_[op] || !(herm_ ? keeps_bits(h, img) : keeps_runs(h, r)))
                    continue;
                mark_[h] = stamp_;
                pb.gates.push_back(h);
                touched |= mask_[h];
                grew = true;
            }
        }
        if (pb.gates.size() > dense) pb.dense = (int)dense;
    }

    int window_end(int q) const
    {
        return std::min((int)lists_[q].size(), head_[q] + kWindow + 1);
    }

    // f[i] = flags of the ungrouped gates before window place i of qubit q.
    void prefix(int q, std::vector<uint8_t>& f) const
    {
        const int end = window_end(q);
        f.assign(end - head_[q] + 1, 0);
        uint8_t acc = 0;
        for (int i = head_[q]; i < end; i++) {
            f[i - head_[q]] = acc;
            const int h = lists_[q][i];
            if (!taken(h)) {
                const bool z = (zm_[h] >> q) & 1u, x = (xm_[h] >> q) & 1u;
                acc |= F_ANY | (z ? 0 : F_NZ) | (x ? 0 : F_NX) | (z || x ? 0 : F_GEN);
            }
        }
        f[end - head_[q]] = acc;
    }

    // No ungrouped gate before gi on any of its qubits fails to commute with
    // it (flags from work_).
    bool ready(int gi) const
    {
        int j = 0;
        for (uint32_t b = mask_[gi]; b; b &= b - 1, j++) {
            const int q = __builtin_ctz(b);
            const int p = pos_[pos0_[gi] + j] - head_[q];
            if (p > kWindow) return false;
            const uint8_t acc = work_[q][p];
            const bool z = (zm_[gi] >> q) & 1u, x = (xm_[gi] >> q) & 1u;
            const uint8_t stop = z && x ? F_GEN : z ? F_NZ : x ? F_NX : F_ANY;
            if (acc & stop) return false;
        }
        return true;
    }

    bool general_on(int h, int q) const { return (((zm_[h] | xm_[h]) >> q) & 1u) == 0; }

    // Ready, untaken gates with a qubit in `qs`, appended once each (listed_
    // stamps).
    void ready_on(uint32_t qs, std::vector<int>& out)
    {
        for (uint32_t b = qs; b; b &= b - 1) {
            const int q = __builtin_ctz(b);
            const int end = window_end(q);
            for (int i = head_[q]; i < end; i++) {
                const int h = lists_[q][i];
                if (taken(h)) continue;
                if (listed_[h] != stamp_ && ready(h)) {
                    listed_[h] = stamp_;
                    out.push_back(h);
                }
                if (general_on(h, q)) break;  // nothing later on q passes it
            }
        }
    }

    // Absorbs, under the current marks, every ready gate inside Q; appends
    // them to `absorbed` in absorption order and marks them.
    void closure(uint32_t Q, std::vector<int>& absorbed)
    {
        for (bool changed = true; changed;) {
            changed = false;
            for (uint32_t b = Q; b; b &= b - 1) prefix(__builtin_ctz(b), work_[__builtin_ctz(b)]);
            for (uint32_t b = Q; b; b &= b - 1) {
                const int q = __builtin_ctz(b);
                const int end = window_end(q);
                for (int i = head_[q]; i < end; i++) {
                    const int h = lists_[q][i];
                    if (taken(h)) continue;
                    if ((mask_[h] & ~Q) == 0 && ready(h)) {
                        mark_[h] = stamp_;
                        absorbed.push_back(h);
                        changed = true;
                    } else if (general_on(h, q)) {
                        break;
                    }
                }
            }
        }
    }

    // Every candidate block reachable once the gates `from` are grouped.
    // Children Q | (a ready gate's qubits) continue from their parent's
    // absorbed gates (absorption only grows with Q).
    void search(const std::vector<int>& from, int cap, std::vector<Cand>& out)
    {
        const uint32_t all = n_ >= 32 ? ~0u : (1u << n_) - 1;
        ++stamp_;
        for (int gi : from) mark_[gi] = stamp_;
        for (int q = 0; q < n_; q++) prefix(q, base_[q]);
        work_ = base_;
        std::vector<int> base_ready;
        ready_on(all, base_ready);
        std::map<uint32_t, int> seen;
        // best first: (priority = parent's score + the joining gate's weight, Q, parent)
        struct Node {
            long prio;
            uint32_t Q;
            int parent;
        };
        auto worse = [](const Node& a, const Node& b) {
            if (a.prio != b.prio) return a.prio < b.prio;
            const int wa = __builtin_popcount(a.Q), wb = __builtin_popcount(b.Q);
            if (wa != wb) return wa > wb;
            return a.Q > b.Q;
        };
        std::priority_queue<Node, std::vector<Node>, decltype(worse)> open(worse);
        for (int gi : base_ready) open.push({weight(gi), mask_[gi], -1});
        uint32_t touched = 0;  // qubits whose work_ flags differ from base_
        while (!open.empty() && (int)out.size() < cap) {
            const Node nd = open.top();
            open.pop();
            const uint32_t Q = nd.Q;
            const int parent = nd.parent;
            if (seen.count(Q)) continue;
            Cand c{Q, parent < 0 ? 0 : out[parent].score,
                   parent < 0 ? from : out[parent].taken};
            ++stamp_;
            for (int gi : c.taken) mark_[gi] = stamp_;
            for (uint32_t b = touched; b; b &= b - 1) work_[__builtin_ctz(b)] = base_[__builtin_ctz(b)];
            touched = Q;
            const size_t before = c.taken.size();
            closure(Q, c.taken);
            for (size_t i = before; i < c.taken.size(); i++) c.score += weight(c.taken[i]);
            seen[Q] = (int)out.size();
            const int limit = std::max(kmax_, __builtin_popcount(Q));
            if (__builtin_popcount(Q) < limit) {
                // ready gates now: those of the search's start off Q (their
                // qubits saw no change), and a rescan of Q's qubits
                for (uint32_t b = Q; b; b &= b - 1) prefix(__builtin_ctz(b), work_[__builtin_ctz(b)]);
                std::vector<int> r;
                for (int gi : base_ready)
                    if (!taken(gi) && (mask_[gi] & Q) == 0) {
                        listed_[gi] = stamp_;
                        r.push_back(gi);
                    }
                ready_on(Q, r);
                for (int gi : r) {
                    const uint32_t Q2 = Q | mask_[gi];
                    if (Q2 != Q && __builtin_popcount(Q2) <= limit && !seen.count(Q2))
                        open.push({c.score + weight(gi), Q2, (int)out.size()});
                }
            }
            out.push_back(std::move(c));
        }
    }

    static constexpr int kMaxExplored = 128;
    static constexpr int kMaxExplored2 = 32;  // lookahead searches
    static constexpr int kLook = 4;
    // From 12 qubits a sweep costs about a millisecond and more: a deeper
    // lookahead (about twice the planning time) saves sweeps -- 192 -> 183 on
    // the 2130-gate 14-qubit brick-wall circuit, 292 -> 284 on a 4000-gate
    // 16-qubit one.
    static constexpr int kLookDeep = 8;
    static constexpr int kWindow = 64;
    static constexpr int kMaxExploredLong = 32;  // segments above kLookMaxGates
    static constexpr size_t kLookMaxGates = 1u << 16;
    static constexpr size_t kClassCacheMax = 4096;
    const std::vector<Gate>& circuit_;
    const int n_, kmax_;
    const bool herm_;  // trailing gates must keep the block's bits a set of index bits
    const int look_;   // candidates compared with the block that would follow them
    std::vector<std::vector<int>> lists_;  // per qubit: its gates in program order
    std::vector<int> head_;                // per qubit: first gate not yet grouped
    std::vector<uint32_t> mask_, zm_, xm_; // per circuit index: qubits, commutation classes
    std::vector<int> pos0_, pos_;          // gate's place in each of its qubits' lists
    std::vector<uint8_t> done_;            // grouped into a committed block
    std::vector<uint64_t> mark_;           // == stamp_: taken in the current evaluation
    std::vector<uint64_t> listed_;         // == stamp_: already in a ready list
    uint64_t stamp_ = 0;
    std::vector<std::vector<uint8_t>> base_, work_;  // per qubit window flags (prefix)
    // classes per operand position (bit j = operand j) of the gates seen, by
    // op and angles (matrix gates are not cached)
    std::map<std::tuple<int, double, double, double>, std::pair<uint32_t, uint32_t>> class_cache_;
    bool perm_op_[OP_COUNT] = {};  // ops that may join a block's index map
    const std::vector<long>* weights_ = nullptr;
};

void group_frontier(const std::vector<Gate>& circuit, int n, int kmax, bool allow_perm, bool herm,
                    std::vector<PB>& pbs, Seq& seq, const std::vector<long>* weights = nullptr)
{
    FrontierGrouper fg(circuit, n, kmax, allow_perm, herm);
    fg.set_weights(weights);
    std::vector<int> seg;
    for (int gi = 0; gi < (int)circuit.size(); gi++) {
        const Gate& g = circuit[gi];
        if (g.op == OP_ID) continue;
        if (g.op == OP_KRAUS) {  // a channel closes the segment: no gate crosses it
            fg.run(seg, pbs, seq);
            seg.clear();
            seq.push_back({3, gi});
            continue;
        }
        seg.push_back(gi);
    }
    fg.run(seg, pbs, seq);
}

// With a cost model: the grouping the network prices cheapest (defined below
// with the pricing helpers).
int choose_grouping(const std::vector<Gate>& circuit, int n, int kmax, bool allow_perm, bool herm,
                    CostModel& model, std::vector<PB>& pbs, Seq& seq, uint32_t& which);

// Sweeps a grouping costs: a five-qubit block counts as about 1.5 four-qubit
// sweeps (its MFMA work is eight times a four-qubit block's per entry).
double grouping_cost(const std::vector<PB>& pbs)
{
    double c = 0;
    for (const PB& pb : pbs) c += __builtin_popcount(pb.mask) > 4 ? 1.5 : 1.0;
    return c;
}

}  // namespace

int plan_blocks(const std::vector<Gate>& circuit, int n, int max_k, BlockPlan& out, bool& fusable,
                bool allow_perm, bool herm, CostModel* model, uint32_t* grouping)
{
    out = BlockPlan{};
    fusable = false;
    if (n < 4) return DMX_OK;  // too few qubits for a 4-qubit block: plain sweeps
    const int kmax = std::min(std::max(max_k, 1), std::min(5, n));
    for (const Gate& g : circuit)
        if (g.op == OP_SRN) return DMX_OK;  // not complex-linear: plain sweeps
    fusable = true;
    // without a model: both groupings, the one with fewer sweeps kept (the
    // first on ties); with one: the cheapest of three as the network prices them
    std::vector<PB> pbs, pbs2;
    Seq seq, seq2;
    uint32_t which = 0;
    if (model) {
        const int rc = choose_grouping(circuit, n, kmax, allow_perm, herm, *model, pbs, seq, which);
        if (rc) return rc;
    } else {
        group_least_growth(circuit, kmax, pbs, seq);
        group_frontier(circuit, n, kmax, allow_perm, herm, pbs2, seq2);
        if (grouping_cost(pbs2) < grouping_cost(pbs)) {
            pbs.swap(pbs2);
            seq.swap(seq2);
            which = 1;
        }
    }
    if (grouping) *grouping = which;
    const uint32_t all = (1u << n) - 1;
    std::vector<std::pair<int, int>> where(pbs.size());  // pbs index -> order entry
    for (size_t bi = 0; bi < pbs.size(); bi++) {
        const PB& pb = pbs[bi];
        uint32_t qmask = pb.mask;
        for (int q = 0; q < n && __builtin_popcount(qmask) < 4; q++) qmask |= 1u << q;
        const int D = 1 << __builtin_popcount(qmask);
        const int dense = pb.dense < 0 ? (int)pb.gates.size() : pb.dense;
        std::vector<cd> U((size_t)D * D, 0.0);
        for (int i = 0; i < D; i++) U[(size_t)i * D + i] = 1.0;
        for (int t = 0; t < dense; t++) {
            const int gi = pb.gates[t];
            int k = 0;
            std::vector<uint32_t> gq;
            std::vector<double> gre, gim;
            gate_matrix(circuit[gi], k, gq, gre, gim);
            apply_embedded(U, qmask, gq, gre, gim);
        }
        if (D == 16) {
            FusedBlock fb;
            std::memset(&fb, 0, sizeof(fb));
            fb.qmask = qmask & all;
            fb.m = (uint32_t)default_tile_bits(n);
            for (int i = 0; i < 256; i++) {
                fb.ur[i] = U[i].real();
                fb.ui[i] = U[i].imag();
            }
            if (dense < (int)pb.gates.size()) {
                // trailing permutation pi(x) = L x + b: b = pi(0), column q = pi(2^q) + b
                auto pi = [&](uint32_t x) {
                    for (size_t t = dense; t < pb.gates.size(); t++) {
                        const Gate& g = circuit[pb.gates[t]];
                        const uint32_t a = g.qb[0], c = g.qb[1];
                        if (g.op == OP_X) {
                            x ^= 1u << a;
                        } else if (g.op == OP_CX) {
                            x ^= ((x >> a) & 1u) << c;
                        } else {  // SWAP
                            const uint32_t d = ((x >> a) ^ (x >> c)) & 1u;
                            x ^= (d << a) | (d << c);
                        }
                    }
                    return x;
                };
                fb.perm_on = 1;
                fb.perm_b = pi(0);
                for (int q = 0; q < n; q++) fb.perm_L[q] = pi(1u << q) ^ fb.perm_b;
                if (!herm && fb.perm_L[0] != 1u) {  // k_fused4_perm relies on bit 0 staying
                    set_error("internal: a block's permutation moves index bit 0");
                    return DMX_EINVAL;
                }
            }
            where[bi] = {0, (int)out.blocks.size()};
            out.blocks.push_back(fb);
            out.members.push_back(pb.gates);
            out.dense.push_back(dense);
        } else {
            FusedBlock5 fb;
            std::memset(&fb, 0, sizeof(fb));
            fb.qmask = qmask & all;
            fb.m = (uint32_t)default_tile_bits5(n);
            if (!tile_bits5_ok(n, (int)fb.m, fb.qmask)) fb.m = (uint32_t)std::min(8, n);
            for (int i = 0; i < 1024; i++) {
                fb.ur[i] = U[i].real();
                fb.ui[i] = U[i].imag();
            }
            where[bi] = {2, (int)out.blocks5.size()};
            out.blocks5.push_back(fb);
            out.members5.push_back(pb.gates);
        }
    }
    for (const auto& e : seq) out.order.push_back(e.first == 3 ? e : where[e.second]);
    return DMX_OK;
}

namespace {

int cost_kind(const PrimOp& p)
{
    switch (prim_class(p)) {
    case KC_QUAD: return CK_QUAD;
    case KC_CX: return CK_CX;
    case KC_PAIR_POS1: return CK_DIAG;
    default: return CK_DENSE;
    }
}

// Predicted time of each gate run unfused: its primitive sweeps on both
// indices (us[gi]; a gate with no primitive sweeps -- a 3..5-qubit matrix --
// gets a very large value, an identity or channel 0).
int unfused_gate_times(const std::vector<Gate>& circuit, int n, CostModel& model,
                       std::vector<double>& us)
{
    std::vector<std::vector<float>> feats;
    std::vector<size_t> first(circuit.size() + 1, 0);
    std::vector<char> none(circuit.size(), 0);
    float f[kMlpFeatures];
    for (size_t gi = 0; gi < circuit.size(); gi++) {
        first[gi] = feats.size();
        const Gate& g = circuit[gi];
        if (g.op == OP_ID || g.op == OP_KRAUS) continue;
        std::vector<PrimOp> ps;
        if (!lower_gate(g, ps)) {
            none[gi] = 1;
            continue;
        }
        for (int side = 0; side < 2; side++)
            for (const PrimOp& p : ps) {
                cost_features(n, cost_kind(p), 0, 0, p.q0, p.q1, side, f);
                feats.emplace_back(f, f + kMlpFeatures);
            }
    }
    first[circuit.size()] = feats.size();
    std::vector<float> t;
    const int rc = model.price(feats, t);
    if (rc) return rc;
    us.assign(circuit.size(), 0.0);
    for (size_t gi = 0; gi < circuit.size(); gi++) {
        if (none[gi]) {
            us[gi] = 1e12;
            continue;
        }
        for (size_t i = first[gi]; i < first[gi + 1]; i++) us[gi] += t[i];
    }
    return DMX_OK;
}

// Predicted time of a grouping: per block the cheaper of its fused sweep (a
// Hermitian sweep, priced on the index bits it reads the block on after the
// relabelling of the sweeps before it; or k_fused4 at the default width;
// k_fused5 for five-qubit blocks) and its gates run unfused.
int price_grouping(const std::vector<Gate>& circuit, const std::vector<PB>& pbs, const Seq& seq,
                   int n, bool herm, const std::vector<double>& gate_us, CostModel& model,
                   double& total)
{
    std::vector<std::vector<float>> feats;
    std::vector<int> row(pbs.size(), -1);
    float f[kMlpFeatures];
    BitOrder bo(n);
    for (const auto& e : seq) {
        if (e.first != 0) continue;  // a channel: on the bit that holds its qubit, order kept
        const PB& pb = pbs[e.second];
        uint32_t q = pb.mask;
        for (int b = 0; b < n && __builtin_popcount(q) < 4; b++) q |= 1u << b;
        if (__builtin_popcount(q) == 5) {
            // in order on the bits that hold its qubits, order kept
            const uint32_t qp = bo.phys(q);
            int m = default_tile_bits5(n);
            if (!tile_bits5_ok(n, m, qp)) m = std::min(8, n);
            cost_features(n, CK_FUSED5, m, qp, 0, -1, 0, f);
        } else if (herm) {
            // priced on the index bits the sweep reads the block on, after
            // the relabelling of the Hermitian sweeps before it
            uint32_t Lp[16];
            trailing_linear(circuit, pb, n, Lp);
            uint32_t qphys = 0;
            BitOrder nb(n);
            if (herm_step(n, bo, q, Lp, &qphys, &nb)) {
                cost_features(n, CK_HERM, 8, qphys, 0, -1, 0, f);
                bo = nb;
            } else {
                cost_features(n, CK_FUSED, default_tile_bits(n), q, 0, -1, 0, f);
                bo.go_home();
            }
        } else {
            cost_features(n, CK_FUSED, default_tile_bits(n), q, 0, -1, 0, f);
        }
        row[e.second] = (int)feats.size();
        feats.emplace_back(f, f + kMlpFeatures);
    }
    std::vector<float> t;
    const int rc = model.price(feats, t);
    if (rc) return rc;
    total = 0;
    for (size_t b = 0; b < pbs.size(); b++) {
        if (row[b] < 0) continue;
        const double fused = (double)t[row[b]];
        double unfused = 0;
        for (int gi : pbs[b].gates) unfused += gate_us[gi];
        total += __builtin_popcount(pbs[b].mask) > 4 ? fused : std::min(fused, unfused);
    }
    return DMX_OK;
}

}  // namespace

namespace {

int choose_grouping(const std::vector<Gate>& circuit, int n, int kmax, bool allow_perm, bool herm,
                    CostModel& model, std::vector<PB>& pbs, Seq& seq, uint32_t& which)
{
    std::vector<double> gate_us;
    int rc = unfused_gate_times(circuit, n, model, gate_us);
    if (rc) return rc;
    // further candidates: the frontier search scoring each block by the
    // predicted unfused time of the gates it takes (hundredths of a
    // microsecond), by that time for its multi-qubit gates only (one-qubit
    // gates join almost any block later), by its square root (a milder
    // spread), and by the structural scores with one-qubit gates weighing
    // four times a multi-qubit gate.  The search is greedy: the weights that
    // give the fewest blocks depend on the circuit (one-qubit weight 1 ->
    // 400: 183 -> 174 blocks on a 2130-gate 14-qubit brick-wall circuit,
    // 284 -> 302 on a 4000-gate 16-qubit U3-heavy one), so the network picks.
    std::vector<long> w(circuit.size(), 1), wm(circuit.size(), 1), ws(circuit.size(), 1),
        wh(circuit.size(), 1);
    for (size_t gi = 0; gi < circuit.size(); gi++) {
        const double us = std::min(gate_us[gi], 1e9);
        const bool multi = __builtin_popcount(gate_mask(circuit[gi])) > 1;
        w[gi] = std::max<long>(1, std::lround(us * 100.0));
        wm[gi] = multi ? w[gi] : 1;
        ws[gi] = std::max<long>(1, std::lround(std::sqrt(us) * 100.0));
        wh[gi] = multi ? 100 : 400;
    }
    constexpr int kCands = 6;
    std::vector<PB> cand[kCands];
    Seq cseq[kCands];
    // The frontier searches are independent of one another and take nearly
    // all of the planning time (0.15-0.3 s each on a few thousand gates with
    // the deep lookahead): one host thread each, joined before the pricing.
    // A thread that cannot be started runs its search here instead.
    const std::vector<long>* wts[kCands] = {nullptr, nullptr, &w, &wm, &ws, &wh};
    std::exception_ptr err[kCands];
    std::vector<std::thread> pool;
    pool.reserve(kCands - 1);  // no reallocation (and no bad_alloc) once threads run
    for (int c = 1; c < kCands; c++) {
        auto job = [&, c] {
            try {
                group_frontier(circuit, n, kmax, allow_perm, herm, cand[c], cseq[c], wts[c]);
            } catch (...) {
                err[c] = std::current_exception();
            }
        };
        try {
            pool.emplace_back(job);
        } catch (const std::system_error&) {
            job();
        }
    }
    try {
        group_least_growth(circuit, kmax, cand[0], cseq[0]);
    } catch (...) {
        err[0] = std::current_exception();
    }
    for (std::thread& t : pool) t.join();
    for (const std::exception_ptr& e : err)
        if (e) std::rethrow_exception(e);
    double best = 0;
    which = 0;
    for (uint32_t c = 0; c < kCands; c++) {
        double t = 0;
        rc = price_grouping(circuit, cand[c], cseq[c], n, herm, gate_us, model, t);
        if (rc) return rc;
        if (c == 0 || t < best ||
            (t == best && grouping_cost(cand[c]) < grouping_cost(cand[which]))) {
            best = t;
            which = c;
        }
    }
    pbs.swap(cand[which]);
    seq.swap(cseq[which]);
    return DMX_OK;
}

// Learned choice per block: fused at tile width m (4..min(8, n)) or its
// member gates' primitive sweeps on both indices.  Sets blocks[b].m, fills
// plain[b] with the primitives of blocks that run unfused.
// A candidate replaces the default only if predicted at least this much faster
// (the model's median error on held-out launches is about 10 %).
constexpr double kMargin = 0.92;

int choose_executions(const std::vector<Gate>& circuit, int n, BlockPlan& bp, CostModel& model,
                      std::vector<std::vector<PrimOp>>& plain, PlanInfo& info, bool herm)
{
    const int nb = (int)bp.blocks.size();
    // Hermitian plans: one fused option per block, the Hermitian sweep at width
    // 8 priced on the index bits it reads the block on (hq, following the
    // relabelling along the plan; 0 for a block that runs in order)
    std::vector<uint32_t> hq(nb, 0);
    if (herm) {
        BitOrder bo(n);
        for (const auto& o : bp.order) {
            if (o.first != 0) continue;  // five-qubit blocks and channels keep the order
            const FusedBlock& fb = bp.blocks[o.second];
            uint32_t Lp[16];
            for (int q = 0; q < 16; q++) Lp[q] = q < n ? (fb.perm_on ? fb.perm_L[q] : 1u << q) : 0u;
            BitOrder next(n);
            if (herm_step(n, bo, fb.qmask, Lp, &hq[o.second], &next))
                bo = next;
            else
                bo.go_home();
        }
    }
    const int m_hi = herm ? 4 : std::min(8, n);
    const int n_m = m_hi - 4 + 1;
    std::vector<std::vector<float>> feats;
    std::vector<size_t> first(nb);
    std::vector<int> n_plain(nb, -1);  // feature rows of the unfused option, -1 if none
    plain.assign(nb, {});
    float f[kMlpFeatures];
    for (int b = 0; b < nb; b++) {
        first[b] = feats.size();
        for (int m = 4; m <= m_hi; m++) {
            if (herm && hq[b]) {
                cost_features(n, CK_HERM, 8, hq[b], 0, -1, 0, f);
            } else if (herm) {
                cost_features(n, CK_FUSED, default_tile_bits(n), bp.blocks[b].qmask, 0, -1, 0, f);
            } else {
                // an inadmissible width is priced like a valid one and never chosen
                const int mm = tile_bits_ok(n, m, bp.blocks[b].qmask) ? m : m_hi;
                cost_features(n, CK_FUSED, mm, bp.blocks[b].qmask, 0, -1, 0, f);
            }
            feats.emplace_back(f, f + kMlpFeatures);
        }
        std::vector<PrimOp> ps;
        bool ok = true;
        for (int gi : bp.members[b]) ok = ok && lower_gate(circuit[gi], ps);
        if (!ok) continue;
        for (int side = 0; side < 2; side++)
            for (const PrimOp& p : ps) {
                cost_features(n, cost_kind(p), 0, 0, p.q0, p.q1, side, f);
                feats.emplace_back(f, f + kMlpFeatures);
            }
        n_plain[b] = 2 * (int)ps.size();
        plain[b] = std::move(ps);
    }
    std::vector<float> us;
    int rc = model.price(feats, us);
    if (rc) return rc;
    const int m_default = herm ? 4 : default_tile_bits(n);  // herm: the one option, slot 0
    for (int b = 0; b < nb; b++) {
        const float* t = us.data() + first[b];
        // the width the network prices fastest (one kernel at different
        // widths: no margin, its errors there are alike)
        int best_m = m_default;
        double best = t[m_default - 4];
        for (int i = 0; i < n_m && !herm; i++)
            if (tile_bits_ok(n, 4 + i, bp.blocks[b].qmask) && t[i] < best) {
                best_m = 4 + i;
                best = t[i];
            }
        if (herm) best_m = 8;
        info.greedy_us += t[m_default - 4];
        bool use_plain = false;
        if (n_plain[b] >= 0) {
            double s = 0;
            for (int i = 0; i < n_plain[b]; i++) s += t[n_m + i];
            if (s < kMargin * best) {
                best = s;
                use_plain = true;
            }
        }
        info.predicted_us += best;
        if (use_plain) {
            info.plain_blocks++;
        } else {
            bp.blocks[b].m = (uint32_t)best_m;
            plain[b].clear();
        }
    }
    return DMX_OK;
}

// Learned choice per five-qubit block: one k_fused5 sweep at tile width m
// (5..min(8, n)), or its member gates re-planned into four-qubit blocks that
// are priced and chosen as above.  Sets blocks5[b].m; sub[b].used marks the
// blocks that run split, with their four-qubit plan.
struct SubPlan {
    bool used = false;
    BlockPlan bp;
    std::vector<std::vector<PrimOp>> plain;
};

int choose_fused5(const std::vector<Gate>& circuit, int n, BlockPlan& bp, CostModel& model,
                  std::vector<SubPlan>& sub, PlanInfo& info)
{
    const int nb = (int)bp.blocks5.size();
    sub.assign(nb, SubPlan{});
    if (nb == 0) return DMX_OK;
    const int m_hi = std::min(8, n);
    const int n_m = m_hi - 5 + 1;
    std::vector<std::vector<float>> feats;
    float f[kMlpFeatures];
    for (int b = 0; b < nb; b++)
        for (int m = 5; m <= m_hi; m++) {
            const int mm = tile_bits5_ok(n, m, bp.blocks5[b].qmask) ? m : m_hi;
            cost_features(n, CK_FUSED5, mm, bp.blocks5[b].qmask, 0, -1, 0, f);
            feats.emplace_back(f, f + kMlpFeatures);
        }
    std::vector<float> us;
    int rc = model.price(feats, us);
    if (rc) return rc;
    for (int b = 0; b < nb; b++) {
        const float* t = us.data() + (size_t)b * n_m;
        const int md = (int)bp.blocks5[b].m;  // the default width plan_blocks chose
        int best_m = md;
        double best = t[md - 5];
        for (int i = 0; i < n_m; i++)
            if (tile_bits5_ok(n, 5 + i, bp.blocks5[b].qmask) && t[i] < best) {
                best_m = 5 + i;
                best = t[i];
            }
        info.greedy_us += t[md - 5];
        bp.blocks5[b].m = (uint32_t)best_m;
        std::vector<Gate> gates;
        for (int gi : bp.members5[b]) gates.push_back(circuit[gi]);
        SubPlan sp;
        bool fusable = false;
        rc = plan_blocks(gates, n, 4, sp.bp, fusable);
        if (rc) return rc;
        if (fusable && sp.bp.blocks5.empty()) {  // a five-qubit gate cannot be split
            PlanInfo si;
            rc = choose_executions(gates, n, sp.bp, model, sp.plain, si, false);
            if (rc) return rc;
            if (si.predicted_us < kMargin * best) {
                best = si.predicted_us;
                info.plain_blocks += si.plain_blocks;
                sp.used = true;
                sub[b] = std::move(sp);
            }
        }
        info.predicted_us += best;
    }
    return DMX_OK;
}

// Rewrites a plan's four-qubit blocks as Hermitian sweeps.  sigma[q] is the
// index bit that holds logical qubit q (the identity at the start).  A block
// on logical qubits Q with trailing map pi (logical, x -> Lp x ^ bp) reads
// rho in the order sigma and writes it in a new order sigma': the qubits
// spanning pi(span Q) at bits 0..3 (ascending), then the images of the
// super-tile's four low bits, then the rest by their current bit.  Its
// record gets the block's bits and matrix in the current order and the map
// phi = sigma' o pi o sigma^-1 on index bits.  A five-qubit block and a
// channel run in order on the bits that hold their qubits, in any order.
// Before any other step and at the end, a relabelling step restores the
// circuit's order.
void to_hermitian(Plan& plan, int n, PlanInfo& info)
{
    std::vector<Step> steps;
    BitOrder bo(n);
    auto move_bits = [&](const std::vector<int>& to, uint32_t v) {
        uint32_t o = 0;
        for (int q = 0; q < n; q++)
            if ((v >> q) & 1u) o |= 1u << to[q];
        return o;
    };
    auto go_home = [&] {
        if (bo.home()) return;
        FusedBlock fb;
        std::memset(&fb, 0, sizeof(fb));
        fb.perm_on = 1;
        for (int p = 0; p < n; p++) fb.perm_L[p] = 1u << bo.sinv[p];  // bit p holds qubit sinv[p]
        steps.push_back({ST_RELABEL, (int)plan.blocks.size(), 0});
        plan.blocks.push_back(fb);
        info.relabels++;
        bo.go_home();
    };
    for (const Step& st : plan.steps) {
        if (st.type == ST_FUSED5 && !bo.home()) {
            // an in-order five-qubit sweep runs on whatever bits hold its
            // qubits: its record gets them and its matrix in the current
            // order, and the order stays (no relabelling before or after)
            FusedBlock5 h = plan.blocks5[st.index];
            const uint32_t qphys = bo.phys(h.qmask);
            int tphys[5], t = 0;
            for (uint32_t m = h.qmask; m; m &= m - 1)
                tphys[t++] = __builtin_popcount(qphys & ((1u << bo.sigma[__builtin_ctz(m)]) - 1u));
            auto remap = [&](int a) {
                int o = 0;
                for (int u = 0; u < 5; u++)
                    if ((a >> u) & 1) o |= 1 << tphys[u];
                return o;
            };
            const FusedBlock5& b = plan.blocks5[st.index];
            for (int a = 0; a < 32; a++)
                for (int c = 0; c < 32; c++) {
                    h.ur[remap(a) * 32 + remap(c)] = b.ur[a * 32 + c];
                    h.ui[remap(a) * 32 + remap(c)] = b.ui[a * 32 + c];
                }
            h.qmask = qphys;
            if (!tile_bits5_ok(n, (int)h.m, qphys))
                h.m = (uint32_t)(tile_bits5_ok(n, default_tile_bits5(n), qphys) ? default_tile_bits5(n)
                                                                                 : std::min(8, n));
            steps.push_back({ST_FUSED5, (int)plan.blocks5.size(), 0});
            plan.blocks5.push_back(h);
            continue;
        }
        if (st.type == ST_CHANNEL && !bo.home()) {
            // a channel's two records (one per orientation) act on the same
            // row and column bit, which a relabelling moves alike: they get
            // the bit that holds their qubit, and the order stays
            const int first = (int)plan.prims.size();
            for (int t = 0; t < 2; t++) {
                PrimOp p = plan.prims[st.index + t];
                p.q0 = bo.sigma[p.q0];
                p.q1 = bo.sigma[p.q1];
                plan.prims.push_back(p);
            }
            steps.push_back({ST_CHANNEL, first, 2});
            continue;
        }
        if (st.type != ST_FUSED4) {
            go_home();
            steps.push_back(st);
            continue;
        }
        const FusedBlock b = plan.blocks[st.index];
        uint32_t Lp[16] = {};
        for (int q = 0; q < n; q++) Lp[q] = b.perm_on ? b.perm_L[q] : 1u << q;
        const uint32_t bp = b.perm_on ? b.perm_b : 0u;
        uint32_t qphys = 0;
        BitOrder nb(n);
        if (!herm_step(n, bo, b.qmask, Lp, &qphys, &nb)) {  // in-order kernel
            go_home();
            steps.push_back(st);
            continue;
        }
        const std::vector<int>& sigma = bo.sigma;
        const std::vector<int>& sinv = bo.sinv;
        const std::vector<int>& so = nb.sigma;
        FusedBlock h;
        std::memset(&h, 0, sizeof(h));
        h.qmask = qphys;
        h.m = 8;
        h.perm_on = 1;
        for (int p = 0; p < n; p++) h.perm_L[p] = move_bits(so, Lp[sinv[p]]);
        h.perm_b = move_bits(so, bp);
        // local bit t of the block's matrix (t-th lowest logical qubit) is
        // local bit tphys[t] of the kernel's (t-th lowest index bit of qphys)
        int tphys[4], t = 0;
        for (uint32_t m = b.qmask; m; m &= m - 1)
            tphys[t++] = __builtin_popcount(qphys & ((1u << sigma[__builtin_ctz(m)]) - 1u));
        auto remap = [&](int a) {
            int o = 0;
            for (int u = 0; u < 4; u++)
                if ((a >> u) & 1) o |= 1 << tphys[u];
            return o;
        };
        for (int a = 0; a < 16; a++)
            for (int c = 0; c < 16; c++) {
                h.ur[remap(a) * 16 + remap(c)] = b.ur[a * 16 + c];
                h.ui[remap(a) * 16 + remap(c)] = b.ui[a * 16 + c];
            }
        steps.push_back({ST_HERM, (int)plan.blocks.size(), 0});
        plan.blocks.push_back(h);
        info.herm_blocks++;
        bo = nb;
    }
    go_home();
    // a Hermitian sweep followed by another writes only what that one reads
    for (size_t i = 0; i + 1 < steps.size(); i++)
        if (steps[i].type == ST_HERM && steps[i + 1].type == ST_HERM)
            plan.blocks[steps[i].index].next_q = plan.blocks[steps[i + 1].index].qmask;
    plan.steps.swap(steps);
}

// The width of every five-qubit sweep of a plan priced by the network on the
// index bits it runs on (a Hermitian plan moves them; the grouping priced
// the circuit's own).
int reprice_fused5_widths(Plan& plan, int n, CostModel& model)
{
    const int m_hi = std::min(8, n);
    std::vector<std::vector<float>> feats;
    std::vector<int> which;
    float f[kMlpFeatures];
    for (const Step& st : plan.steps) {
        if (st.type != ST_FUSED5) continue;
        which.push_back(st.index);
        for (int m = 5; m <= m_hi; m++) {
            // an inadmissible width is priced like a valid one and never chosen
            const int mm = tile_bits5_ok(n, m, plan.blocks5[st.index].qmask) ? m : m_hi;
            cost_features(n, CK_FUSED5, mm, plan.blocks5[st.index].qmask, 0, -1, 0, f);
            feats.emplace_back(f, f + kMlpFeatures);
        }
    }
    if (which.empty()) return DMX_OK;
    std::vector<float> us;
    const int rc = model.price(feats, us);
    if (rc) return rc;
    const int n_m = m_hi - 5 + 1;
    for (size_t i = 0; i < which.size(); i++) {
        FusedBlock5& b = plan.blocks5[which[i]];
        const float* t = us.data() + i * n_m;
        int best_m = (int)b.m;
        for (int k = 0; k < n_m; k++)
            if (tile_bits5_ok(n, 5 + k, b.qmask) && t[k] < t[best_m - 5]) best_m = 5 + k;
        b.m = (uint32_t)best_m;
    }
    return DMX_OK;
}

}  // namespace

int build_fused_plan(const std::vector<Gate>& circuit, int n, int max_k, CostModel* model,
                     int tile_bits, bool allow_perm, FusedPlan** out, PlanInfo* info, bool herm)
{
    *out = nullptr;
    PlanInfo pi;
    BlockPlan bp;
    bool fusable = false;
    herm = herm && allow_perm && n >= kHermMinQubits && n <= 16;
    int rc = plan_blocks(circuit, n, max_k, bp, fusable, allow_perm, herm,
                         tile_bits > 0 ? nullptr : model, &pi.grouping);
    if (rc || !fusable) return rc;
    std::vector<std::vector<PrimOp>> plain(bp.blocks.size());
    std::vector<SubPlan> sub(bp.blocks5.size());
    if (tile_bits > 0) {
        for (FusedBlock& fb : bp.blocks)
            if (tile_bits_ok(n, tile_bits, fb.qmask)) fb.m = (uint32_t)tile_bits;
        for (FusedBlock5& fb : bp.blocks5)
            if (tile_bits5_ok(n, tile_bits, fb.qmask)) fb.m = (uint32_t)tile_bits;
    } else if (model) {
        rc = choose_executions(circuit, n, bp, *model, plain, pi, herm);
        if (rc) return rc;
        rc = choose_fused5(circuit, n, bp, *model, sub, pi);
        if (rc) return rc;
    }
    Plan* plan = new Plan();
    plan->n = n;
    plan->blocks = std::move(bp.blocks);
    plan->blocks5 = std::move(bp.blocks5);
    for (const auto& o : bp.order) {
        if (o.first == 3) {
            const int first = (int)plan->prims.size();
            lower_gate(circuit[o.second], plan->prims);  // the two orientation records
            plan->steps.push_back({ST_CHANNEL, first, 2});
            pi.channels++;
            continue;
        }
        if (o.first == 2 && sub[o.second].used) {  // the block's four-qubit plan
            SubPlan& sp = sub[o.second];
            for (const auto& so : sp.bp.order) {
                const int j = so.second;
                if (sp.plain[j].empty()) {
                    plan->steps.push_back({ST_FUSED4, (int)plan->blocks.size(), 0});
                    plan->blocks.push_back(sp.bp.blocks[j]);
                    pi.fused_blocks++;
                    pi.tile_hist[sp.bp.blocks[j].m]++;
                } else {
                    const int first = (int)plan->prims.size();
                    plan->prims.insert(plan->prims.end(), sp.plain[j].begin(), sp.plain[j].end());
                    plan->steps.push_back({ST_PLAIN, first, (int)plan->prims.size() - first});
                }
            }
        } else if (o.first == 2) {
            plan->steps.push_back({ST_FUSED5, o.second, 0});
            pi.fused5_blocks++;
        } else if (plain[o.second].empty()) {
            plan->steps.push_back({ST_FUSED4, o.second, 0});
            pi.fused_blocks++;
            pi.perm_blocks += plan->blocks[o.second].perm_on ? 1 : 0;
            pi.tile_hist[plan->blocks[o.second].m]++;
        } else {
            const int first = (int)plan->prims.size();
            plan->prims.insert(plan->prims.end(), plain[o.second].begin(), plain[o.second].end());
            plan->steps.push_back({ST_PLAIN, first, (int)plan->prims.size() - first});
        }
    }
    if (herm) to_hermitian(*plan, n, pi);
    if (herm && model && tile_bits <= 0) {
        rc = reprice_fused5_widths(*plan, n, *model);
        if (rc) {
            delete plan;
            return rc;
        }
    }
    if (model && tile_bits <= 0 && max_k > 4 && !plan->blocks5.empty()) {
        // Whole-plan check: five-qubit blocks pay off only when they replace
        // enough four-qubit sweeps, which no single block's split can show.
        // The requested width stays unless the narrower plan is predicted
        // faster by the margin.  (At max_k = 4 the only five-qubit blocks are
        // five-qubit gates, which every plan keeps: nothing to compare.)
        FusedPlan* alt = nullptr;
        PlanInfo ai;
        rc = build_fused_plan(circuit, n, 4, model, 0, allow_perm, &alt, &ai, herm);
        if (rc) {
            delete plan;
            return rc;
        }
        if (alt && ai.predicted_us < kMargin * pi.predicted_us) {
            delete plan;
            *out = alt;
            if (info) *info = ai;
            return DMX_OK;
        }
        delete alt;
    }
    *out = plan;
    if (info) *info = pi;
    return DMX_OK;
}

}  // namespace dmx
