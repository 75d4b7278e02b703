// fused5.hip -- one sweep of rho applies a dense 5-qubit block on both sides,
// A <- conj(U) A U^T for the stored A = conj(rho), with FP64 MFMA
// (v_mfma_f64_16x16x4f64) on gfx950.
//
// Tile: the 32 x 32 entries whose row and column indices range over the
// block's five qubits Q.  A workgroup stages a super-tile of 32 rows x 2^m
// columns (Q plus the lowest m - 5 free column bits) in LDS; each of its
// 2^(m-5) tiles is transformed by one wave as 2 x 2 sub-tiles of 16 x 16:
//   X[I'][J] = sum_K T[I'][K] U^T[K][J]       K = eight chunks of 4
//   T'[I][J] = sum_K conj(U)[I][K] X[K][J]
// with complex products from three real MFMA products each (192 MFMAs per
// tile).  A wave computes all four X sub-tiles (registers) before it writes
// any T' entry, so it never reads an entry it has already rewritten; tiles of
// different waves occupy disjoint LDS columns.
// MFMA lane layout (as in fused.hip): A[i][k] in lane i+16k, B[k][j] in lane
// j+16k, D[4r+g][j] in lane j+16g register r.  Lane l holds U[16J + (l & 15)]
// [4K + (l >> 4)] for J = 0, 1 and K = 0..7: the B operand (U^T) of the first
// product and the A operand (conj U) of the second; register r of X[I'][J] is
// the B operand of the second product's chunk K = 4I' + r.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fused.hpp"

namespace dmx {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

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

template <int M>
__global__ void __launch_bounds__(256, 2) k_fused5(double* __restrict__ re,
                                                    double* __restrict__ im,
                                                    const FusedBlock5* __restrict__ blk, int n)
{
    constexpr int kP = (1 << M) + 1;  // LDS row pitch in doubles (odd)
    __shared__ double s_re[32 * kP];
    __shared__ double s_im[32 * kP];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int j = lane & 15;
    const int kk = lane >> 4;

    const uint64_t dim = 1ull << n;
    const uint32_t qmask = blk->qmask;
    uint32_t cmask = qmask;
    for (int q = 0; q < n && __builtin_popcount(cmask) < M; q++) cmask |= 1u << q;
    constexpr int ntiles = 1 << (M - 5);

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

    double ur[2][8], ui[2][8];
#pragma unroll
    for (int J = 0; J < 2; J++)
#pragma unroll
        for (int K = 0; K < 8; K++) {
            const int idx = (16 * J + j) * 32 + 4 * K + kk;
            ur[J][K] = blk->ur[idx];
            ui[J][K] = blk->ui[idx];
        }
    uint32_t lcq_a[8];  // super-tile column of block column 4K + kk
#pragma unroll
    for (int K = 0; K < 8; K++) lcq_a[K] = (uint32_t)deposit(4 * K + kk, pos_q);
    uint32_t lcq_d[2];  // super-tile column of block column 16J + j
#pragma unroll
    for (int J = 0; J < 2; J++) lcq_d[J] = (uint32_t)deposit(16 * J + j, pos_q);

    const uint64_t row_free = (dim - 1) & ~(uint64_t)qmask;
    const uint64_t col_free = (dim - 1) & ~(uint64_t)cmask;
    const int cg_bits = n - M;
    const uint64_t n_st = 1ull << (n - 5 + cg_bits);
    constexpr int per_plane = 32 << M;      // >= 1024
    constexpr int iters = per_plane / 512;  // double2 moves per thread and plane
    constexpr int chunk = iters < 8 ? iters : 8;

    const int lc = (2 * tid) & ((1 << M) - 1);
    const uint64_t coloff = deposit((uint64_t)lc, cmask);
    const int a0 = (2 * tid) >> M;  // first super-tile row this thread moves
    constexpr int a_step = 512 >> M;

    for (uint64_t st = blockIdx.x; st < n_st; st += gridDim.x) {
        const uint64_t rg = st >> cg_bits;
        const uint64_t cg = st & ((1ull << cg_bits) - 1);
        const uint64_t g0 = deposit(rg, row_free) * dim + deposit(cg, col_free) + coloff;

#pragma unroll
        for (int c0 = 0; c0 < iters; c0 += chunk) {
            double2 vr[chunk], vi[chunk];
#pragma unroll
            for (int t = 0; t < chunk; t++) {
                const uint64_t off = g0 + deposit((uint64_t)(a0 + (c0 + t) * a_step), qmask) * dim;
                vr[t] = *reinterpret_cast<const double2*>(re + off);
                vi[t] = *reinterpret_cast<const double2*>(im + off);
            }
#pragma unroll
            for (int t = 0; t < chunk; t++) {
                const int a = a0 + (c0 + t) * a_step;
                s_re[a * kP + lc] = vr[t].x;
                s_re[a * kP + lc + 1] = vr[t].y;
                s_im[a * kP + lc] = vi[t].x;
                s_im[a * kP + lc + 1] = vi[t].y;
            }
        }
        __syncthreads();

        for (int tb = wave; tb < ntiles; tb += 4) {
            const uint32_t base = (uint32_t)deposit((uint64_t)tb, pos_f);
            // X = T U^T, all four sub-tiles
            d4 xr[2][2], xi[2][2];
#pragma unroll
            for (int I = 0; I < 2; I++) {
                d4 p1[2], p2[2], p3[2];
#pragma unroll
                for (int J = 0; J < 2; J++) p1[J] = p2[J] = p3[J] = d4{0, 0, 0, 0};
#pragma unroll
                for (int K = 0; K < 8; K++) {
                    const int row = (16 * I + j) * kP + base + lcq_a[K];
                    const double tr = s_re[row], ti = s_im[row];
                    const double ts = tr + ti;
#pragma unroll
                    for (int J = 0; J < 2; J++) {
                        p1[J] = mfma(tr, ur[J][K], p1[J]);
                        p2[J] = mfma(ti, ui[J][K], p2[J]);
                        p3[J] = mfma(ts, ur[J][K] + ui[J][K], p3[J]);
                    }
                }
#pragma unroll
                for (int J = 0; J < 2; J++) {
                    xr[I][J] = p1[J] - p2[J];
                    xi[I][J] = p3[J] - p1[J] - p2[J];
                }
            }
            // T' = conj(U) X
#pragma unroll
            for (int I = 0; I < 2; I++)
#pragma unroll
                for (int J = 0; J < 2; J++) {
                    d4 q1 = {0, 0, 0, 0}, q2 = {0, 0, 0, 0}, q3 = {0, 0, 0, 0};
#pragma unroll
                    for (int K = 0; K < 8; K++) {
                        const double br = xr[K >> 2][J][K & 3];
                        const double bi = xi[K >> 2][J][K & 3];
                        q1 = mfma(ur[I][K], br, q1);
                        q2 = mfma(-ui[I][K], bi, q2);
                        q3 = mfma(ur[I][K] - ui[I][K], br + bi, q3);
                    }
                    const d4 yr = q1 - q2;
                    const d4 yi = q3 - q1 - q2;
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int a = 16 * I + 4 * r + kk;
                        s_re[a * kP + base + lcq_d[J]] = yr[r];
                        s_im[a * kP + base + lcq_d[J]] = yi[r];
                    }
                }
        }
        __syncthreads();

#pragma unroll
        for (int t = 0; t < iters; t++) {
            const int a = a0 + t * a_step;
            const uint64_t off = g0 + deposit((uint64_t)a, qmask) * dim;
            *reinterpret_cast<double2*>(re + off) =
                make_double2(s_re[a * kP + lc], s_re[a * kP + lc + 1]);
            *reinterpret_cast<double2*>(im + off) =
                make_double2(s_im[a * kP + lc], s_im[a * kP + lc + 1]);
        }
        __syncthreads();
    }
}

hipError_t launch_fused5(double* re, double* im, const FusedBlock5* d_blk, int n, int m,
                         int max_blocks, hipStream_t stream)
{
    if (n < 5 || m < 5 || m > 8 || m > n) return hipErrorInvalidValue;
    const uint64_t n_st = (1ull << (n - 5)) * (1ull << (n - m));
    const unsigned blocks = (unsigned)(n_st < (uint64_t)max_blocks ? n_st : (uint64_t)max_blocks);
    switch (m) {
    case 5: k_fused5<5><<<blocks, 256, 0, stream>>>(re, im, d_blk, n); break;
    case 6: k_fused5<6><<<blocks, 256, 0, stream>>>(re, im, d_blk, n); break;
    case 7: k_fused5<7><<<blocks, 256, 0, stream>>>(re, im, d_blk, n); break;
    default: k_fused5<8><<<blocks, 256, 0, stream>>>(re, im, d_blk, n); break;
    }
    return hipGetLastError();
}

}  // namespace dmx
