// herm.hpp -- index maps of the Hermitian block sweep (k_fused4_herm) and of
// the bit relabelling (k_relabel), shared by the kernels in fused.hip and the
// host-side check of those maps (tests/cpp/herm_maps.cpp).
//
// Hermitian sweep on n = 8..16 qubits: the block's tiles are the 16 x 16
// entries over its four index bits (qmask) with the other row and column bits
// fixed (free parts r, c).  A super-tile is 16 rows (one free row part) x 256
// columns: the block's bits plus the four lowest free bits ("low" bits),
// i.e. 16 tiles c = 16 c_hi + l.  Free parts are compared as compressed
// indices (bits outside qmask, in order), so the low bits are a free part's
// four lowest.  Only super-tiles with r >> 4 <= c_hi are work items; on an
// item with r >> 4 == c_hi, tiles l < r & 15 lie below the diagonal and are
// skipped, tile l == r & 15 is the diagonal one.
// Output map phi(x) = L x ^ b (L by its columns): the host chooses it so that
// L maps the span of qmask onto bits 0..3; f(a) = L(a on the block's bits) is
// then a 4-bit bijection, and entry (row a, column j) of a tile with free
// parts (r, c) lands at (R0 ^ f(a), C0 ^ f(j)) with R0 = phi(r), C0 = phi(c).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fused.hpp"

namespace dmx {

// Scatter the low bits of v into the positions of `mask` (ascending).
__host__ __device__ inline uint64_t deposit_bits(uint64_t v, uint64_t mask)
{
    uint64_t out = 0;
    for (uint64_t m = mask; m; m &= m - 1) {
        const uint64_t bit = m & (~m + 1);
        if (v & 1) out |= bit;
        v >>= 1;
    }
    return out;
}

// LDS row pitch (doubles) of the 16 x 256 staging planes.
constexpr int kHermPitch = 257;

// LDS place (doubles from the plane's start) of super-tile row a, column x in
// the Hermitian sweep: the column's high nibble is XORed into its low nibble
// (a bijection on 0..255).  Where the block's bits sit at super-tile column
// positions 4..7 (the block above the four low free bits), a tile's 16
// columns are 16 doubles apart and the transform's result writes -- one row,
// 16 columns per instruction -- would hit 8 of the 32 double-wide LDS banks;
// the XOR spreads them over all of them.
__host__ __device__ inline int herm_lds(int a, int x)
{
    return a * kHermPitch + (x ^ ((x >> 4) & 15));
}

struct HermGeom {
    int n;
    uint64_t dim;
    uint32_t qmask;    // the block's index bits
    uint32_t cmask;    // super-tile column bits: qmask and the four lowest free bits
    uint32_t lowmask;  // those four free bits
    uint32_t pos_q;    // positions of qmask's bits inside the 8-bit super-tile column index
    uint32_t pos_f;    // positions of the low bits there
    uint64_t row_free, col_free;  // free row bits; column bits outside cmask
    uint32_t H;        // column groups 2^(n-8)
    uint32_t L[16];
    uint32_t pb;
    uint64_t itab;     // f^-1, four bits per entry
    // Entries the next Hermitian sweep reads: (x, y) with x & ~Q' <= y & ~Q'
    // (Q' = next_q, its index bits; compressing away bits keeps this order),
    // i.e. the tiles on and above its diagonal.  For an output tile at rows
    // R0 ^ f(a), columns C0 ^ f(j) that comparison is decided by the high
    // parts R0 & hi_keep vs C0 & hi_keep unless they are equal; then both
    // the tile and its mirror are written.  half = false: write every entry.
    bool half;
    uint32_t hi_keep;  // bits >= 4 outside next_q

    __host__ __device__ void init(const FusedBlock& blk, int nq)
    {
        n = nq;
        dim = 1ull << n;
        qmask = blk.qmask;
        cmask = qmask;
        for (int q = 0; q < n && __builtin_popcount(cmask) < 8; q++) cmask |= 1u << q;
        lowmask = cmask & ~qmask;
        pos_q = pos_f = 0;
        int p = 0;
        for (uint32_t b = cmask; b; b &= b - 1, p++) {
            if ((b & (~b + 1)) & qmask)
                pos_q |= 1u << p;
            else
                pos_f |= 1u << p;
        }
        row_free = (dim - 1) & ~(uint64_t)qmask;
        col_free = (dim - 1) & ~(uint64_t)cmask;
        H = 1u << (n - 8);
#pragma unroll
        for (int q = 0; q < 16; q++) L[q] = blk.perm_L[q];
        pb = blk.perm_b;
        half = blk.next_q != 0;
        hi_keep = (uint32_t)(dim - 1) & ~15u & ~blk.next_q;
        itab = 0;
        for (uint32_t a = 0; a < 16; a++)
            itab |= (uint64_t)a << (4 * (lin((uint32_t)deposit_bits(a, qmask)) & 15u));
    }

    __host__ __device__ uint32_t lin(uint32_t x) const
    {
        uint32_t y = 0;
#pragma unroll
        for (int q = 0; q < 16; q++) y ^= ((x >> q) & 1u) ? L[q] : 0u;
        return y;
    }

    __host__ __device__ uint32_t inv(uint32_t p) const { return (uint32_t)(itab >> (4 * p)) & 15u; }

    __host__ __device__ uint64_t items() const { return 16ull * ((uint64_t)H * (H + 1) / 2); }

    // LDS column of entry jq of tile l
    __host__ __device__ int lds_col(int l, uint32_t jq) const
    {
        return (int)deposit_bits((uint64_t)l, pos_f) + (int)deposit_bits(jq, pos_q);
    }
};

// Per-thread constants of a launch (thread tid of 256).
struct HermThread {
    int lc;              // loads: LDS column pair lc, lc + 1 ...
    uint64_t coloff;     // ... whose column bits these are
    uint64_t rowoff[8];  // load rep: row bits of LDS row (tid >> 7) + 2 rep, times dim
    int k2;              // stores: column pair k2, k2 + 1 of a 16-column output run
    int mid;             // direct stores: tile l; mirror stores: output row within the run
    int top;             // direct: row slot top + 2 rep; mirror: tile top + 2 rep
    uint32_t Ld;         // L(low bits of tile mid)
    int fd;              // LDS column of tile mid's first entry
    uint32_t Lm[8];      // L(low bits of tile top + 2 rep)
    int fm[8];           // LDS column of that tile's first entry
    int pq0, pq1, pq2, pq3;  // positions of the block's bits in the LDS column index

    __host__ __device__ void init(const HermGeom& g, int tid)
    {
        lc = (2 * tid) & 255;
        coloff = deposit_bits((uint64_t)lc, g.cmask);
        k2 = 2 * (tid & 7);
        mid = (tid >> 3) & 15;
        top = tid >> 7;
#pragma unroll
        for (int rep = 0; rep < 8; rep++) {
            rowoff[rep] = deposit_bits((uint64_t)(top + 2 * rep), g.qmask) * g.dim;
            Lm[rep] = g.lin((uint32_t)deposit_bits((uint64_t)(top + 2 * rep), g.lowmask));
            fm[rep] = (int)deposit_bits((uint64_t)(top + 2 * rep), g.pos_f);
        }
        Ld = g.lin((uint32_t)deposit_bits((uint64_t)mid, g.lowmask));
        fd = (int)deposit_bits((uint64_t)mid, g.pos_f);
        uint32_t m = g.pos_q;  // four bits
        pq0 = __builtin_ctz(m);
        m &= m - 1;
        pq1 = __builtin_ctz(m);
        m &= m - 1;
        pq2 = __builtin_ctz(m);
        m &= m - 1;
        pq3 = __builtin_ctz(m);
    }

    // LDS column offset of the block-bit part jq (0..15) of a tile entry
    __host__ __device__ int dq(uint32_t jq) const
    {
        return (int)(((jq & 1u) << pq0) | (((jq >> 1) & 1u) << pq1) |
                     (((jq >> 2) & 1u) << pq2) | (((jq >> 3) & 1u) << pq3));
    }
};

// One work item: super-tile of free row part (r_hi << 4 | r_l), column group c_hi.
struct HermItem {
    uint32_t r_l;
    bool diag;
    uint64_t rphys;  // the rows' free part as index bits
    uint64_t cbase;  // the columns' bits outside cmask
    uint32_t R0;     // phi(rphys)
    uint32_t Cb;     // phi(cbase)
};

__host__ __device__ inline HermItem herm_item(const HermGeom& g, uint64_t w)
{
    HermItem it;
    it.r_l = (uint32_t)(w & 15);
    const uint64_t pidx = w >> 4;
    uint32_t r_hi = 0, c_hi = 0;
    if (g.H > 1) {
        // the triangle r_hi <= c_hi folded into a (H + 1) x H/2 rectangle: row a
        // holds (a, a..H-1) then (H-1-a, H-1-a..H-1)
        const uint32_t a = (uint32_t)(pidx / (g.H + 1)), bc = (uint32_t)(pidx % (g.H + 1));
        if (bc < g.H - a) {
            r_hi = a;
            c_hi = a + bc;
        } else {
            r_hi = g.H - 1 - a;
            c_hi = r_hi + (bc - (g.H - a));
        }
    }
    it.diag = r_hi == c_hi;
    it.rphys = deposit_bits(((uint64_t)r_hi << 4) | it.r_l, g.row_free);
    it.cbase = deposit_bits(c_hi, g.col_free);
    it.R0 = g.lin((uint32_t)it.rphys) ^ g.pb;
    it.Cb = g.lin((uint32_t)it.cbase) ^ g.pb;
    return it;
}

// Schedule: the 16 items of one (r_hi, c_hi) pair -- whose mirror stores
// fill the same 256 rows x 2 KiB of the output together -- run on one XCD at
// the same time.  Workgroups are dealt to the 8 XCDs round-robin (XCD =
// block % 8); the grid is a multiple of 128, so each XCD's workgroups
// local = block / 8 come in groups of 16.  Slot s = local + k * (grid / 8)
// of an XCD (k-th pass of its loop) takes pair (s / 16) * 8 + XCD, item
// s % 16 of it; w is that pair's item.  False once the pairs are used up
// (then for every later k as well).
__host__ __device__ inline bool herm_slot(const HermGeom& g, uint32_t block, uint32_t grid,
                                         uint64_t k, uint64_t& w)
{
    const uint64_t s = (uint64_t)(block >> 3) + k * (uint64_t)(grid >> 3);
    const uint64_t pair = (s >> 4) * 8 + (block & 7);
    if (pair >= g.items() / 16) return false;
    w = pair * 16 + (s & 15);
    return true;
}

// Load `rep` (0..7): global offset (from the plane's start) of a column pair
// and its LDS place (row a, columns th.lc, th.lc + 1).
__host__ __device__ inline uint64_t herm_load(const HermGeom& g, const HermThread& th,
                                             const HermItem& it, int rep, int& a)
{
    a = th.top + 2 * rep;
    return it.rphys * g.dim + it.cbase + th.coloff + th.rowoff[rep];
}

// Whether wave-tile tb of the item is transformed (and stored).
__host__ __device__ inline bool herm_tile_on(const HermItem& it, int tb)
{
    return !(it.diag && tb < (int)it.r_l);
}

// Direct store `rep` (0..7): the tile in its own place.  Output row
// (R0 & ~15) | s, columns (C0 & ~15) | k2, k2 + 1; sources LDS row a,
// columns c0, c1.  Returns false where the thread stores nothing.
__host__ __device__ inline bool herm_direct(const HermGeom& g, const HermThread& th,
                                           const HermItem& it, int rep, uint64_t& off, int& a,
                                           int& c0, int& c1)
{
    if (it.diag && th.mid < (int)it.r_l) return false;
    const uint32_t C0 = it.Cb ^ th.Ld;
    if (g.half && (it.R0 & g.hi_keep) > (C0 & g.hi_keep)) return false;  // the next sweep reads the mirror
    const uint32_t tc = C0 & 15u;
    c0 = th.fd + th.dq(g.inv((uint32_t)th.k2 ^ tc));
    c1 = th.fd + th.dq(g.inv((uint32_t)(th.k2 + 1) ^ tc));
    const uint32_t s = (uint32_t)(th.top + 2 * rep);
    a = (int)g.inv(s ^ (it.R0 & 15u));
    off = (uint64_t)((it.R0 & ~15u) | s) * g.dim + ((C0 & ~15u) | (uint32_t)th.k2);
    return true;
}

// Mirror store `rep` (0..7): the conjugate transpose of tile top + 2 rep in
// the mirror tile's place.  Output row (C0 & ~15) | mid, columns
// (R0 & ~15) | k2, k2 + 1; sources LDS rows a0, a1, column col (conjugated).
__host__ __device__ inline bool herm_mirror(const HermGeom& g, const HermThread& th,
                                           const HermItem& it, int rep, uint64_t& off, int& a0,
                                           int& a1, int& col)
{
    if (it.diag && th.top + 2 * rep <= (int)it.r_l) return false;
    const uint32_t C0 = it.Cb ^ th.Lm[rep];
    if (g.half && (C0 & g.hi_keep) > (it.R0 & g.hi_keep)) return false;  // the next sweep reads the tile
    col = th.fm[rep] + th.dq(g.inv((uint32_t)th.mid ^ (C0 & 15u)));
    a0 = (int)g.inv((uint32_t)th.k2 ^ (it.R0 & 15u));
    a1 = (int)g.inv((uint32_t)(th.k2 + 1) ^ (it.R0 & 15u));
    off = (uint64_t)((C0 & ~15u) | (uint32_t)th.mid) * g.dim + ((it.R0 & ~15u) | (uint32_t)th.k2);
    return true;
}

// ---- relabelling: entry (x, y) -> (p(x), p(y)) for a bit permutation p ----

struct RelabelGeom {
    int n;
    uint64_t dim;
    uint32_t P[16], Pinv[16];  // bit q -> the single bit p(q); the inverse
    uint32_t S;                // staged column bits: 0..3, p^-1(0..3), then the lowest others
    uint32_t pS;               // p(S)
    uint64_t comp;             // column bits outside S

    __host__ __device__ void init(const FusedBlock& blk, int nq)
    {
        n = nq;
        dim = 1ull << n;
#pragma unroll
        for (int q = 0; q < 16; q++) P[q] = q < n ? blk.perm_L[q] : (1u << q);
        // inverse without indexing by a computed position (keeps both tables in registers)
#pragma unroll
        for (int t = 0; t < 16; t++) {
            uint32_t v = 0;
#pragma unroll
            for (int q = 0; q < 16; q++) v |= P[q] == (1u << t) ? (1u << q) : 0u;
            Pinv[t] = v;
        }
        S = 0xFu | map(Pinv, 0xFu);
        for (int q = 0; q < n && __builtin_popcount(S) < 8; q++) S |= 1u << q;
        pS = map(P, S);
        comp = (dim - 1) & ~(uint64_t)S;
    }

    __host__ __device__ static uint32_t map(const uint32_t* T, uint32_t x)
    {
        uint32_t y = 0;
#pragma unroll
        for (int q = 0; q < 16; q++) y |= ((x >> q) & 1u) ? T[q] : 0u;
        return y;
    }

    __host__ __device__ static uint32_t extract(uint32_t x, uint32_t m)
    {
        uint32_t out = 0;
        int k = 0;
        for (uint32_t b = m; b; b &= b - 1, k++)
            if (x & b & (~b + 1)) out |= 1u << k;
        return out;
    }

    // items: 16 consecutive rows x the 256 columns of one column group
    __host__ __device__ uint64_t items() const { return (dim >> 4) * (1ull << (n - 8)); }
};

// Per-thread constants of a relabelling launch.
struct RelabelThread {
    int lc;           // LDS column pair of loads
    uint64_t incol;   // its input column bits
    uint32_t ocol0;   // output column bits of this thread's store pair
    int s0, s1;       // LDS columns holding them
    uint32_t pa[8];   // p(row (tid >> 7) + 2 rep of the 16-row group)

    __host__ __device__ void init(const RelabelGeom& g, int tid)
    {
        lc = (2 * tid) & 255;
        incol = deposit_bits((uint64_t)lc, g.S);
        ocol0 = (uint32_t)deposit_bits((uint64_t)lc, g.pS);
        s0 = (int)RelabelGeom::extract(RelabelGeom::map(g.Pinv, ocol0), g.S);
        s1 = (int)RelabelGeom::extract(RelabelGeom::map(g.Pinv, ocol0 | 1u), g.S);
#pragma unroll
        for (int rep = 0; rep < 8; rep++)
            pa[rep] = RelabelGeom::map(g.P, (uint32_t)((tid >> 7) + 2 * rep));
    }
};

// One item: 16 consecutive input rows (rg) x the 256 columns of group cg.
struct RelabelItem {
    uint64_t rg;    // first row
    uint64_t cg;    // the columns' bits outside S
    uint32_t prg;   // p(rg)
    uint32_t pcg;   // p(cg)
};

__host__ __device__ inline RelabelItem relabel_item(const RelabelGeom& g, uint64_t w)
{
    RelabelItem it;
    it.rg = (w >> (g.n - 8)) << 4;
    it.cg = deposit_bits(w & ((1ull << (g.n - 8)) - 1), g.comp);
    it.prg = RelabelGeom::map(g.P, (uint32_t)it.rg);
    it.pcg = RelabelGeom::map(g.P, (uint32_t)it.cg);
    return it;
}

// Load `rep` (0..7): global offset of a column pair; LDS row a (columns th.lc,
// th.lc + 1).
__host__ __device__ inline uint64_t relabel_load(const RelabelGeom& g, const RelabelThread& th,
                                                const RelabelItem& it, int tid, int rep, int& a)
{
    a = (tid >> 7) + 2 * rep;
    return (it.rg | (uint64_t)a) * g.dim + it.cg + th.incol;
}

// Store `rep`: global offset of an output column pair; sources LDS row a,
// columns th.s0, th.s1.
__host__ __device__ inline uint64_t relabel_store(const RelabelGeom& g, const RelabelThread& th,
                                                 const RelabelItem& it, int tid, int rep, int& a)
{
    a = (tid >> 7) + 2 * rep;
    return (uint64_t)(it.prg | th.pa[rep]) * g.dim + (it.pcg | th.ocol0);
}

}  // namespace dmx
