// fused.hip -- one sweep of rho applies a dense 4-qubit block on both sides,
// rho <- U rho U^dagger, i.e. A <- conj(U) A U^T for the stored A = conj(rho),
// with FP64 MFMA (v_mfma_f64_16x16x4f64) on gfx950.
//
// Tile: the 16 x 16 entries whose row and column indices range over the
// block's four qubits Q (other bits fixed).  A workgroup stages a super-tile
// in LDS: those 16 rows x 2^m columns, the columns spanning Q plus the lowest
// m - 4 free column bits.  At m = 8 the four lowest column bits are always
// inside the super-tile, so every global access is a contiguous run of at
// least 16 doubles.  Threads move column pairs, so column bit 0 must be in
// the super-tile (m > 4, or qubit 0 in Q).
// The 2^(m-4) tiles of a super-tile are dealt round-robin to the four waves;
// each tile is transformed by one wave:
//   X  = T U^T      (12 MFMAs: complex product from three real products)
//   T' = conj(U) X  (12 MFMAs)
// Two workgroups per CU (64 KiB of LDS each at m = 8) overlap one another's
// load, transform and store phases.
// MFMA lane layout on gfx950 (measured): A[i][k] in lane i+16k, B[k][j] in
// lane j+16k, D[4r+g][j] in lane j+16g register r -- so register c of X is
// exactly the B operand of the second product's K-chunk c, and the block
// coefficients U[j][4c+k] held in lane j+16k serve as B (U^T) in the first
// product and as A (conj(U)) in the second.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fused.hpp"
#include "herm.hpp"

namespace dmx {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kPitch = 257;  // LDS row pitch in doubles (odd: spreads rows over banks)

// LDS place of super-tile row a, column x (x < 2^M): the column's high nibble
// XORed into its low nibble (a bijection on 0..2^M-1).  A block on qubits
// above the super-tile's low free bits keeps a tile's 16 columns 16 doubles
// apart, so the transform's result writes (one row, 16 columns) would hit 8
// of the 32 double-wide banks; the XOR spreads them.
__device__ __forceinline__ int lds_at(int a, int x)
{
    return a * kPitch + (x ^ ((x >> 4) & 15));
}

__device__ __forceinline__ d4 mfma(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

typedef double dv2 __attribute__((ext_vector_type(2)));

// Nontemporal 16-byte load and store (rho is streamed once per sweep).
__device__ __forceinline__ dv2 ld_stream(const double* p)
{
    return __builtin_nontemporal_load(reinterpret_cast<const dv2*>(p));
}

__device__ __forceinline__ void st_stream(double* p, dv2 v)
{
    __builtin_nontemporal_store(v, reinterpret_cast<dv2*>(p));
}

template <bool NT>
__device__ __forceinline__ dv2 ld_v2(const double* p)
{
    if constexpr (NT) return ld_stream(p);
    return *reinterpret_cast<const dv2*>(p);
}

template <bool NT>
__device__ __forceinline__ void st_v2(double* p, dv2 v)
{
    if constexpr (NT)
        st_stream(p, v);
    else
        *reinterpret_cast<dv2*>(p) = v;
}

// Scatter the low bits of v into the positions of `mask` (ascending).
__device__ __forceinline__ uint64_t deposit(uint64_t v, uint64_t mask)
{
    uint64_t out = 0;
    for (uint64_t m = mask; m; m &= m - 1) {
        const uint64_t bit = m & (~m + 1);
        if (v & 1) out |= bit;
        v >>= 1;
    }
    return out;
}

}  // namespace

// Super-tiles st_first, st_first + st_stride, ... of one block sweep; s_re and
// s_im are the workgroup's 16 x kPitch LDS planes.  M = number of super-tile
// column bits (8 for n >= 8, n below that).
// PERM: the block has a trailing permutation pi; super-tiles are read from
// (re, im) and written, entry (i, j) to (pi(i), pi(j)), into (re_out, im_out).
// Otherwise re_out/im_out are re/im (in place).
template <int M, bool PERM = false, bool PIPE = false>
__device__ __forceinline__ void fused_sweep(double* re, double* im, double* re_out,
                                            double* im_out, const FusedBlock* __restrict__ blk,
                                            int n, uint64_t st_first, uint64_t st_stride,
                                            double* __restrict__ s_re, double* __restrict__ s_im)
{
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int j = lane & 15;   // MFMA row/column within the 16x16 tile
    const int kk = lane >> 4;  // MFMA K index / row group

    const uint64_t dim = 1ull << n;
    const uint32_t qmask = blk->qmask;  // the block's 4 qubits
    // super-tile column bits: the block's qubits plus the lowest free bits, M in all
    uint32_t cmask = qmask;
    for (int q = 0; q < n && __builtin_popcount(cmask) < M; q++) cmask |= 1u << q;
    constexpr int m = M;
    constexpr int ntiles = 1 << (M - 4);

    // positions (inside the super-tile column index) of the block's qubits and
    // of the free "batch" bits
    uint32_t pos_q = 0, pos_f = 0;
    {
        int p = 0;
        for (uint32_t b = cmask; b; b &= b - 1, p++) {
            const uint32_t bit = b & (~b + 1);
            if (bit & qmask)
                pos_q |= 1u << p;
            else
                pos_f |= 1u << p;
        }
    }

    // Complex products with three real multiplications each (P1 = Ar Br,
    // P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi); re = P1 - P2, im = P3 - P1 - P2).
    // First product B = U^T, second product A = conj(U): both read U[j][4c+k]
    // in lane j+16k.
    double ur[4], ui[4], upi[4], nui[4], umi[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int idx = j * 16 + 4 * c + kk;
        ur[c] = blk->ur[idx];
        ui[c] = blk->ui[idx];
        upi[c] = ur[c] + ui[c];
        nui[c] = -ui[c];
        umi[c] = ur[c] - ui[c];
    }
    // super-tile column index of tile element b (its Q part)
    uint32_t lcq_a[4];
#pragma unroll
    for (int c = 0; c < 4; c++) lcq_a[c] = (uint32_t)deposit(4 * c + kk, pos_q);
    const uint32_t lcq_d = (uint32_t)deposit(j, pos_q);

    const uint64_t row_free = (dim - 1) & ~(uint64_t)qmask;
    const uint64_t col_free = (dim - 1) & ~(uint64_t)cmask;
    const int cg_bits = n - m;
    const uint64_t n_st = 1ull << (n - 4 + cg_bits);
    constexpr int per_plane = 16 << M;
    constexpr int iters = per_plane >= 512 ? per_plane / 512 : 1;
    const bool active = 2 * tid < per_plane;

    // this thread's fixed super-tile column (the same for every row it moves)
    const int lc = (2 * tid) & ((1 << m) - 1);
    const uint64_t coloff = deposit((uint64_t)lc, cmask);
    uint64_t rowoff[iters];
#pragma unroll
    for (int it = 0; it < iters; it++)
        rowoff[it] = deposit((uint64_t)((2 * tid + 512 * it) >> m), qmask) * dim;

    auto base_of = [&](uint64_t st) {
        const uint64_t rg = st >> cg_bits;
        const uint64_t cg = st & ((1ull << cg_bits) - 1);
        return deposit(rg, row_free) * dim + deposit(cg, col_free) + coloff;
    };

    // pi's linear part: L x = XOR of the columns of L for the set bits of x
    // (columns loaded once; n <= 16)
    uint32_t L[PERM ? 16 : 1];
    if (PERM) {
#pragma unroll
        for (int q = 0; q < (PERM ? 16 : 1); q++) L[q] = blk->perm_L[q];
    }
    auto lin = [&](uint32_t x) {
        uint32_t y = 0;
#pragma unroll
        for (int q = 0; q < (PERM ? 16 : 1); q++) y ^= ((x >> q) & 1u) ? L[q] : 0u;
        return y;
    };
    // this thread's fixed row and column offsets inside a super-tile, mapped
    // by L (pi(base ^ off) = L base ^ L off ^ b)
    uint32_t l_row[PERM ? iters : 1], l_col = 0;
    if (PERM) {
#pragma unroll
        for (int it = 0; it < (PERM ? iters : 1); it++)
            l_row[it] = lin((uint32_t)deposit((uint64_t)((2 * tid + 512 * it) >> m), qmask));
        l_col = lin((uint32_t)coloff);
    }

    // PIPE (rho streamed from HBM): software pipeline, the next super-tile's
    // loads are issued once this one is in LDS and land in registers during
    // the transform and stores.  Without it (rho in the Infinity Cache, few
    // super-tiles per workgroup) each super-tile is loaded when its turn comes.
    double2 vr[iters], vi[iters];
    auto load = [&](uint64_t g) {
#pragma unroll
        for (int it = 0; it < iters; it++) {
            vr[it] = *reinterpret_cast<const double2*>(re + g + rowoff[it]);
            vi[it] = *reinterpret_cast<const double2*>(im + g + rowoff[it]);
        }
    };
    if (PIPE && active && st_first < n_st) load(base_of(st_first));
    for (uint64_t st = st_first; st < n_st; st += st_stride) {
        const uint64_t g0 = base_of(st);
        if (active) {
            if (!PIPE) load(g0);
#pragma unroll
            for (int it = 0; it < iters; it++) {
                const int a = (2 * tid + 512 * it) >> m;
                s_re[lds_at(a, lc)] = vr[it].x;
                s_re[lds_at(a, lc + 1)] = vr[it].y;
                s_im[lds_at(a, lc)] = vi[it].x;
                s_im[lds_at(a, lc + 1)] = vi[it].y;
            }
        }
        __syncthreads();
        if (PIPE && active && st + st_stride < n_st) load(base_of(st + st_stride));

        for (int tb = wave; tb < ntiles; tb += 4) {
            const uint32_t base = (uint32_t)deposit((uint64_t)tb, pos_f);
            double tr[4], ti[4], ts[4];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                tr[c] = s_re[lds_at(j, (int)(base + lcq_a[c]))];
                ti[c] = s_im[lds_at(j, (int)(base + lcq_a[c]))];
                ts[c] = tr[c] + ti[c];
            }
            // X = T U^T
            d4 p1 = {0, 0, 0, 0}, p2 = {0, 0, 0, 0}, p3 = {0, 0, 0, 0};
#pragma unroll
            for (int c = 0; c < 4; c++) {
                p1 = mfma(tr[c], ur[c], p1);
                p2 = mfma(ti[c], ui[c], p2);
                p3 = mfma(ts[c], upi[c], p3);
            }
            const d4 xr = p1 - p2;
            const d4 xi = p3 - p1 - p2;
            const d4 xs = xr + xi;
            // T' = conj(U) X, conj(U) = ur + i*(-ui)
            d4 q1 = {0, 0, 0, 0}, q2 = {0, 0, 0, 0}, q3 = {0, 0, 0, 0};
#pragma unroll
            for (int c = 0; c < 4; c++) {
                q1 = mfma(ur[c], xr[c], q1);
                q2 = mfma(nui[c], xi[c], q2);
                q3 = mfma(umi[c], xs[c], q3);
            }
            const d4 yr = q1 - q2;
            const d4 yi = q3 - q1 - q2;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int a = 4 * r + kk;
                s_re[lds_at(a, (int)(base + lcq_d))] = yr[r];
                s_im[lds_at(a, (int)(base + lcq_d))] = yi[r];
            }
        }
        __syncthreads();

        if (active && !PERM) {
#pragma unroll
            for (int it = 0; it < iters; it++) {
                const int a = (2 * tid + 512 * it) >> m;
                *reinterpret_cast<double2*>(re + g0 + rowoff[it]) =
                    make_double2(s_re[lds_at(a, lc)], s_re[lds_at(a, lc + 1)]);
                *reinterpret_cast<double2*>(im + g0 + rowoff[it]) =
                    make_double2(s_im[lds_at(a, lc)], s_im[lds_at(a, lc + 1)]);
            }
        } else if (active) {
            const uint64_t rg = st >> cg_bits;
            const uint64_t cg = st & ((1ull << cg_bits) - 1);
            const uint32_t pr = lin((uint32_t)deposit(rg, row_free)) ^ blk->perm_b;
            // pi(column pair) = (c, c ^ 1) as pi keeps bit 0: store in order
            const uint32_t pc = lin((uint32_t)deposit(cg, col_free)) ^ blk->perm_b ^ l_col;
            const int sw = (int)(pc & 1u);
            const uint64_t dcol = pc & ~1u;
#pragma unroll
            for (int it = 0; it < iters; it++) {
                const int a = (2 * tid + 512 * it) >> m;
                const uint64_t off = (uint64_t)(pr ^ l_row[it]) * dim + dcol;
                const double r0 = s_re[lds_at(a, lc)], r1 = s_re[lds_at(a, lc + 1)];
                const double i0 = s_im[lds_at(a, lc)], i1 = s_im[lds_at(a, lc + 1)];
                *reinterpret_cast<double2*>(re_out + off) =
                    sw ? make_double2(r1, r0) : make_double2(r0, r1);
                *reinterpret_cast<double2*>(im_out + off) =
                    sw ? make_double2(i1, i0) : make_double2(i0, i1);
            }
        }
        __syncthreads();
    }
}

template <int M, bool PIPE>
__global__ void __launch_bounds__(256, 2) k_fused4(double* __restrict__ re, double* __restrict__ im,
                                                    const FusedBlock* __restrict__ blk, int n)
{
    __shared__ double s_re[16 * kPitch];
    __shared__ double s_im[16 * kPitch];
    fused_sweep<M, false, PIPE>(re, im, re, im, blk, n, blockIdx.x, gridDim.x, s_re, s_im);
}

template <int M, bool PIPE>
__global__ void __launch_bounds__(256, 2) k_fused4_perm(const double* __restrict__ re,
                                                         const double* __restrict__ im,
                                                         double* __restrict__ re_out,
                                                         double* __restrict__ im_out,
                                                         const FusedBlock* __restrict__ blk, int n)
{
    __shared__ double s_re[16 * kPitch];
    __shared__ double s_im[16 * kPitch];
    fused_sweep<M, true, PIPE>(const_cast<double*>(re), const_cast<double*>(im), re_out, im_out,
                               blk, n, blockIdx.x, gridDim.x, s_re, s_im);
}

// Batched form: blockIdx.y selects one circuit's (rho, block) pair, all of
// qubit count n.
template <int M>
__global__ void __launch_bounds__(256, 2) k_fused4_batch(const BatchEntry* __restrict__ entries,
                                                          int n)
{
    __shared__ double s_re[16 * kPitch];
    __shared__ double s_im[16 * kPitch];
    const BatchEntry e = entries[blockIdx.y];
    if (e.blk->perm_on)
        fused_sweep<M, true>(e.re, e.im, e.re_out, e.im_out, e.blk, n, blockIdx.x, gridDim.x,
                             s_re, s_im);
    else
        fused_sweep<M>(e.re, e.im, e.re, e.im, e.blk, n, blockIdx.x, gridDim.x, s_re, s_im);
}

// ---- Hermitian sweep ------------------------------------------------------
//
// rho stays Hermitian under every map the engine applies (U rho U^dagger,
// Kraus sums), so the stored A = conj(rho) satisfies T(c, r) = T(r, c)^dagger
// for the tiles of any block (r, c: the tiles' free row and column parts),
// and T'(c, r) = T'(r, c)^dagger after the block as well.  This sweep reads
// and transforms only the tiles with r <= c and writes each result twice:
// T'(r, c) in the tile's place and its conjugate transpose in the mirror
// tile's place -- half the reads and half the MFMA work of k_fused4.
// Writes go to a second pair of planes through the block's output map phi,
// which the host chooses so that every tile lands in 16 rows x one aligned
// run of 16 columns both ways (index maps: herm.hpp).  phi also carries the
// block's trailing X/CX/SWAP gates and the new bit order.
// One Hermitian block sweep over the items `next` deals this workgroup:
// next(k, w) sets w to its k-th item and returns false once they are used up
// (then for every later k as well).  NT: nontemporal loads and stores (rho
// streamed from HBM); without, cached ones (rho in the Infinity Cache).
template <bool NT, bool PIPE, class Next>
__device__ __forceinline__ void herm_sweep(const double* __restrict__ re,
                                           const double* __restrict__ im,
                                           double* __restrict__ re_out,
                                           double* __restrict__ im_out,
                                           const FusedBlock* __restrict__ blk, const HermGeom& g,
                                           Next next, double* __restrict__ s_re,
                                           double* __restrict__ s_im)
{
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int j = lane & 15;
    const int kk = lane >> 4;

    double ur[4], ui[4], upi[4], nui[4], umi[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int idx = j * 16 + 4 * c + kk;
        ur[c] = blk->ur[idx];
        ui[c] = blk->ui[idx];
        upi[c] = ur[c] + ui[c];
        nui[c] = -ui[c];
        umi[c] = ur[c] - ui[c];
    }
    uint32_t lcq_a[4];
#pragma unroll
    for (int c = 0; c < 4; c++) lcq_a[c] = (uint32_t)deposit(4 * c + kk, g.pos_q);
    const uint32_t lcq_d = (uint32_t)deposit(j, g.pos_q);

    HermThread th;
    th.init(g, tid);
    const int lc = th.lc;
    // PIPE: software pipeline, the next item's loads are issued before this
    // item's transform and stores and land in registers while they run;
    // without, each item is loaded when its turn comes (fewer registers).
    dv2 vr[8], vi[8];
    uint64_t w = 0;
    bool have = next(0, w);
    HermItem it;
    auto load = [&] {
        it = herm_item(g, w);
#pragma unroll
        for (int rep = 0; rep < 8; rep++) {
            int a;
            const uint64_t off = herm_load(g, th, it, rep, a);
            vr[rep] = ld_v2<NT>(re + off);
            vi[rep] = ld_v2<NT>(im + off);
        }
    };
    if (PIPE && have) load();
    for (uint64_t k = 0; have; k++) {
        if (!PIPE) load();
#pragma unroll
        for (int rep = 0; rep < 8; rep++) {
            const int a = th.top + 2 * rep;
            s_re[herm_lds(a, lc)] = vr[rep].x;
            s_re[herm_lds(a, lc + 1)] = vr[rep].y;
            s_im[herm_lds(a, lc)] = vi[rep].x;
            s_im[herm_lds(a, lc + 1)] = vi[rep].y;
        }
        __syncthreads();
        const HermItem cur = it;
        have = next(k + 1, w);
        if (PIPE && have) load();

        for (int tb = wave; tb < 16; tb += 4) {
            if (!herm_tile_on(cur, tb)) continue;
            const uint32_t base = (uint32_t)deposit((uint64_t)tb, g.pos_f);
            double tr[4], ti[4], ts[4];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                tr[c] = s_re[herm_lds(j, (int)(base + lcq_a[c]))];
                ti[c] = s_im[herm_lds(j, (int)(base + lcq_a[c]))];
                ts[c] = tr[c] + ti[c];
            }
            d4 p1 = {0, 0, 0, 0}, p2 = {0, 0, 0, 0}, p3 = {0, 0, 0, 0};
#pragma unroll
            for (int c = 0; c < 4; c++) {
                p1 = mfma(tr[c], ur[c], p1);
                p2 = mfma(ti[c], ui[c], p2);
                p3 = mfma(ts[c], upi[c], p3);
            }
            const d4 xr = p1 - p2;
            const d4 xi = p3 - p1 - p2;
            const d4 xs = xr + xi;
            d4 q1 = {0, 0, 0, 0}, q2 = {0, 0, 0, 0}, q3 = {0, 0, 0, 0};
#pragma unroll
            for (int c = 0; c < 4; c++) {
                q1 = mfma(ur[c], xr[c], q1);
                q2 = mfma(nui[c], xi[c], q2);
                q3 = mfma(umi[c], xs[c], q3);
            }
            const d4 yr = q1 - q2;
            const d4 yi = q3 - q1 - q2;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int a = 4 * r + kk;
                s_re[herm_lds(a, (int)(base + lcq_d))] = yr[r];
                s_im[herm_lds(a, (int)(base + lcq_d))] = yi[r];
            }
        }
        __syncthreads();

#pragma unroll
        for (int rep = 0; rep < 8; rep++) {
            uint64_t off;
            int a, c0, c1;
            if (herm_direct(g, th, cur, rep, off, a, c0, c1)) {
                st_v2<NT>(re_out + off, dv2{s_re[herm_lds(a, c0)], s_re[herm_lds(a, c1)]});
                st_v2<NT>(im_out + off, dv2{s_im[herm_lds(a, c0)], s_im[herm_lds(a, c1)]});
            }
        }
#pragma unroll
        for (int rep = 0; rep < 8; rep++) {
            uint64_t off;
            int a0, a1, col;
            if (herm_mirror(g, th, cur, rep, off, a0, a1, col)) {
                st_v2<NT>(re_out + off, dv2{s_re[herm_lds(a0, col)], s_re[herm_lds(a1, col)]});
                st_v2<NT>(im_out + off, dv2{-s_im[herm_lds(a0, col)], -s_im[herm_lds(a1, col)]});
            }
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256, 2) k_fused4_herm(const double* __restrict__ re,
                                                         const double* __restrict__ im,
                                                         double* __restrict__ re_out,
                                                         double* __restrict__ im_out,
                                                         const FusedBlock* __restrict__ blk,
                                                         int n)
{
    __shared__ double s_re[16 * kHermPitch];
    __shared__ double s_im[16 * kHermPitch];
    HermGeom g;
    g.init(*blk, n);
    herm_sweep<true, true>(re, im, re_out, im_out, blk, g,
                     [&](uint64_t k, uint64_t& w) {
                         return herm_slot(g, blockIdx.x, gridDim.x, k, w);
                     },
                     s_re, s_im);
}

// Batched form: blockIdx.y selects one circuit's (rho, block) pair, all of
// qubit count n; its items are dealt grid-stride over gridDim.x workgroups.
// NT and PIPE as in herm_sweep.
template <bool NT, bool PIPE>
__global__ void __launch_bounds__(256, 2) k_fused4_herm_batch(const BatchEntry* __restrict__ entries,
                                                               int n)
{
    __shared__ double s_re[16 * kHermPitch];
    __shared__ double s_im[16 * kHermPitch];
    const BatchEntry e = entries[blockIdx.y];
    HermGeom g;
    g.init(*e.blk, n);
    const uint64_t items = g.items();
    herm_sweep<NT, PIPE>(e.re, e.im, e.re_out, e.im_out, e.blk, g,
                      [&](uint64_t k, uint64_t& w) {
                          w = blockIdx.x + k * gridDim.x;
                          return w < items;
                      },
                      s_re, s_im);
}

// rho's index bits relabelled, entry (x, y) -> (p(x), p(y)) for a bit
// permutation p (perm_L[q] = 1 << p(q)), out of place.  A workgroup stages
// 16 consecutive rows x 256 columns whose column bits hold bits 0..3 and the
// bits p maps to 0..3, so reads and writes move runs of at least 16 doubles
// (index maps: herm.hpp).
template <bool NT>
__device__ __forceinline__ void relabel_sweep(const double* __restrict__ re,
                                              const double* __restrict__ im,
                                              double* __restrict__ re_out,
                                              double* __restrict__ im_out,
                                              const FusedBlock* __restrict__ blk, int n,
                                              double* __restrict__ s_re, double* __restrict__ s_im)
{
    const int tid = threadIdx.x;
    RelabelGeom g;
    g.init(*blk, n);
    RelabelThread th;
    th.init(g, tid);
    const int lc = th.lc;
    const uint64_t n_items = g.items();
    for (uint64_t w = blockIdx.x; w < n_items; w += gridDim.x) {
        const RelabelItem it = relabel_item(g, w);
#pragma unroll
        for (int rep = 0; rep < 8; rep++) {
            int a;
            const uint64_t off = relabel_load(g, th, it, tid, rep, a);
            const dv2 vr = ld_v2<NT>(re + off);
            const dv2 vi = ld_v2<NT>(im + off);
            s_re[a * kHermPitch + lc] = vr.x;
            s_re[a * kHermPitch + lc + 1] = vr.y;
            s_im[a * kHermPitch + lc] = vi.x;
            s_im[a * kHermPitch + lc + 1] = vi.y;
        }
        __syncthreads();
#pragma unroll
        for (int rep = 0; rep < 8; rep++) {
            int a;
            const uint64_t off = relabel_store(g, th, it, tid, rep, a);
            st_v2<NT>(re_out + off, dv2{s_re[a * kHermPitch + th.s0], s_re[a * kHermPitch + th.s1]});
            st_v2<NT>(im_out + off, dv2{s_im[a * kHermPitch + th.s0], s_im[a * kHermPitch + th.s1]});
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256, 2) k_relabel(const double* __restrict__ re,
                                                     const double* __restrict__ im,
                                                     double* __restrict__ re_out,
                                                     double* __restrict__ im_out,
                                                     const FusedBlock* __restrict__ blk, int n)
{
    __shared__ double s_re[16 * kHermPitch];
    __shared__ double s_im[16 * kHermPitch];
    relabel_sweep<true>(re, im, re_out, im_out, blk, n, s_re, s_im);
}

// Batched form: blockIdx.y selects one circuit's (rho, relabelling) pair.
// Nontemporal like the batched Hermitian sweeps it sits between (circuits of
// 10 and 11 qubits; 0.3 ms less per run of 256 circuits than cached).
__global__ void __launch_bounds__(256, 2) k_relabel_batch(const BatchEntry* __restrict__ entries,
                                                           int n)
{
    __shared__ double s_re[16 * kHermPitch];
    __shared__ double s_im[16 * kHermPitch];
    const BatchEntry e = entries[blockIdx.y];
    relabel_sweep<true>(e.re, e.im, e.re_out, e.im_out, e.blk, n, s_re, s_im);
}

namespace {

uint64_t supertiles(int n, int m) { return (1ull << (n - 4)) * (1ull << (n - m)); }

// From this many qubits rho (>= 268 MB) streams from HBM rather than staying
// in the Infinity Cache: the in-order sweeps pipeline their loads.
constexpr int kPipeQubits = 12;

// From this many qubits the batched Hermitian sweeps pipeline their loads and
// load and store nontemporally (one launch then streams >= 0.8 GB: 48 ten-qubit
// circuits).  256 hardware-efficient circuits of 6..11 qubits: 61.2 ms per run
// with cached accesses, 60.4 with nontemporal ones from 11 qubits, 59.8 from 10.
constexpr int kBatchHermStreamQubits = 10;

}  // namespace

hipError_t launch_fused4(double* re, double* im, const FusedBlock* d_blk, int n, int m,
                         int max_blocks, hipStream_t stream)
{
    if (n < 4 || m < 4 || m > 8 || m > n) return hipErrorInvalidValue;
    const uint64_t n_st = supertiles(n, m);
    const unsigned blocks = (unsigned)(n_st < (uint64_t)max_blocks ? n_st : (uint64_t)max_blocks);
    const bool pipe = n >= kPipeQubits;
    switch (m * 2 + (pipe ? 1 : 0)) {
    case 8: k_fused4<4, false><<<blocks, 256, 0, stream>>>(re, im, d_blk, n); break;
    case 9: k_fused4<4, true><<<blocks, 256, 0, stream>>>(re, im, d_blk, n); break;
    case 10: k_fused4<5, false><<<blocks, 256, 0, stream>>>(re, im, d_blk, n); break;
    case 11: k_fused4<5, true><<<blocks, 256, 0, stream>>>(re, im, d_blk, n); break;
    case 12: k_fused4<6, false><<<blocks, 256, 0, stream>>>(re, im, d_blk, n); break;
    case 13: k_fused4<6, true><<<blocks, 256, 0, stream>>>(re, im, d_blk, n); break;
    case 14: k_fused4<7, false><<<blocks, 256, 0, stream>>>(re, im, d_blk, n); break;
    case 15: k_fused4<7, true><<<blocks, 256, 0, stream>>>(re, im, d_blk, n); break;
    case 16: k_fused4<8, false><<<blocks, 256, 0, stream>>>(re, im, d_blk, n); break;
    default: k_fused4<8, true><<<blocks, 256, 0, stream>>>(re, im, d_blk, n); break;
    }
    return hipGetLastError();
}

hipError_t launch_fused4_perm(const double* re, const double* im, double* re_out, double* im_out,
                              const FusedBlock* d_blk, int n, int m, int max_blocks,
                              hipStream_t stream)
{
    if (n < 4 || n > 16 || m < 4 || m > 8 || m > n || re == re_out || im == im_out)
        return hipErrorInvalidValue;
    const uint64_t n_st = supertiles(n, m);
    const unsigned blocks = (unsigned)(n_st < (uint64_t)max_blocks ? n_st : (uint64_t)max_blocks);
    const bool pipe = n >= kPipeQubits;
#define DMX_PERM_CASE(MM)                                                                  \
    case MM:                                                                              \
        if (pipe)                                                                         \
            k_fused4_perm<MM, true><<<blocks, 256, 0, stream>>>(re, im, re_out, im_out, d_blk, n); \
        else                                                                              \
            k_fused4_perm<MM, false><<<blocks, 256, 0, stream>>>(re, im, re_out, im_out, d_blk, n); \
        break;
    switch (m) {
        DMX_PERM_CASE(4)
        DMX_PERM_CASE(5)
        DMX_PERM_CASE(6)
        DMX_PERM_CASE(7)
    default:
        if (pipe)
            k_fused4_perm<8, true><<<blocks, 256, 0, stream>>>(re, im, re_out, im_out, d_blk, n);
        else
            k_fused4_perm<8, false><<<blocks, 256, 0, stream>>>(re, im, re_out, im_out, d_blk, n);
        break;
    }
#undef DMX_PERM_CASE
    return hipGetLastError();
}

hipError_t launch_fused4_herm(const double* re, const double* im, double* re_out,
                              double* im_out, const FusedBlock* d_blk, int n, int max_blocks,
                              hipStream_t stream)
{
    if (n < 8 || n > 16 || re == re_out || im == im_out) return hipErrorInvalidValue;
    const uint64_t H = 1ull << (n - 8);
    const uint64_t items = 16 * (H * (H + 1) / 2);
    // a multiple of 128 workgroups (herm_slot): 16 per XCD group
    uint64_t blocks = ((items < (uint64_t)max_blocks ? items : (uint64_t)max_blocks) + 127) / 128 * 128;
    k_fused4_herm<<<(unsigned)blocks, 256, 0, stream>>>(re, im, re_out, im_out, d_blk, n);
    return hipGetLastError();
}

hipError_t launch_relabel(const double* re, const double* im, double* re_out, double* im_out,
                          const FusedBlock* d_blk, int n, int max_blocks, hipStream_t stream)
{
    if (n < 8 || n > 16 || re == re_out || im == im_out) return hipErrorInvalidValue;
    const uint64_t items = (1ull << (n - 4)) * (1ull << (n - 8));
    const unsigned blocks = (unsigned)(items < (uint64_t)max_blocks ? items : (uint64_t)max_blocks);
    k_relabel<<<blocks, 256, 0, stream>>>(re, im, re_out, im_out, d_blk, n);
    return hipGetLastError();
}

hipError_t launch_fused4_herm_batch(const BatchEntry* d_entries, int count, int n,
                                    int max_blocks, hipStream_t stream)
{
    if (n < 8 || n > 16 || count < 1 || count > 65535) return hipErrorInvalidValue;
    const uint64_t H = 1ull << (n - 8);
    const uint64_t items = 16 * (H * (H + 1) / 2);
    uint64_t per = ((uint64_t)max_blocks + count - 1) / count;
    if (per < 1) per = 1;
    const dim3 grid((unsigned)(items < per ? items : per), (unsigned)count);
    if (n >= kBatchHermStreamQubits)
        k_fused4_herm_batch<true, true><<<grid, 256, 0, stream>>>(d_entries, n);
    else
        k_fused4_herm_batch<false, false><<<grid, 256, 0, stream>>>(d_entries, n);
    return hipGetLastError();
}

hipError_t launch_relabel_batch(const BatchEntry* d_entries, int count, int n, int max_blocks,
                                hipStream_t stream)
{
    if (n < 8 || n > 16 || count < 1 || count > 65535) return hipErrorInvalidValue;
    const uint64_t items = (1ull << (n - 4)) * (1ull << (n - 8));
    uint64_t per = ((uint64_t)max_blocks + count - 1) / count;
    if (per < 1) per = 1;
    const dim3 grid((unsigned)(items < per ? items : per), (unsigned)count);
    k_relabel_batch<<<grid, 256, 0, stream>>>(d_entries, n);
    return hipGetLastError();
}

hipError_t launch_fused4_batch(const BatchEntry* d_entries, int count, int n, int m,
                               int max_blocks, hipStream_t stream)
{
    if (n < 4 || n > 16 || m < 4 || m > 8 || m > n || count < 1 || count > 65535)
        return hipErrorInvalidValue;
    const uint64_t n_st = supertiles(n, m);
    // about max_blocks workgroups in all, each looping over super-tiles of its
    // circuit (one workgroup per super-tile re-runs the set-up for every tile)
    uint64_t per = ((uint64_t)max_blocks + count - 1) / count;
    if (per < 1) per = 1;
    const dim3 grid((unsigned)(n_st < per ? n_st : per), (unsigned)count);
    switch (m) {
    case 4: k_fused4_batch<4><<<grid, 256, 0, stream>>>(d_entries, n); break;
    case 5: k_fused4_batch<5><<<grid, 256, 0, stream>>>(d_entries, n); break;
    case 6: k_fused4_batch<6><<<grid, 256, 0, stream>>>(d_entries, n); break;
    case 7: k_fused4_batch<7><<<grid, 256, 0, stream>>>(d_entries, n); break;
    default: k_fused4_batch<8><<<grid, 256, 0, stream>>>(d_entries, n); break;
    }
    return hipGetLastError();
}

}  // namespace dmx
