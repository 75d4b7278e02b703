// sweep.hip -- plain-method kernels for gfx950: one primitive sweep over rho.
//
// rho is two FP64 planes (re, im) of 4^n entries, element (r, c) at r*2^n + c.
// A primitive on qubit q updates every pair of entries whose linear index
// differs only in bit b, where b = q for the fast (column) index and b = q + n
// for the slow (row) index.  The reference runs all gates on the fast index,
// conj-transposes, and runs them on the fast index again
// (src/dmsim_amdgpu_omp.hpp:934-985).  Here the second pass sweeps the slow
// index instead, with the primitive evaluated on conjugated operands and its
// result conjugated (exact sign flips), which leaves rho conj-transposed
// bit for bit relative to the reference and needs no transpose or second
// buffer; the next run starts on the slow index and flips it back.
//
// Access patterns (64-wide wavefronts, 16-byte vector loads per lane):
//   b >= 7 : "wide"   -- a wave owns 128 consecutive pairs; pos0 and pos1 are
//                        two contiguous 1 KiB runs per plane.
//   b <= 6 : "narrow" -- a wave owns 256 consecutive entries; the partner of
//                        an entry sits in the same 16-byte lane vector (b = 0)
//                        or in lane ^ 2^(b-1), exchanged by wavefront shuffle.
//   4^n < 256          -- scalar fallback (n <= 3).
// Arithmetic follows the reference's per-element expressions and is compiled
// without FMA contraction, so results match the reference CPU path exactly.
#pragma clang fp contract(off)

#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "prim.hpp"
#include "sweep.hpp"

namespace dmx {

namespace {

constexpr double kS2I = 0.70710678118654752440;  // src/config.hpp:57

__device__ __forceinline__ uint64_t ins0(uint64_t x, int b)
{
    const uint64_t lo = x & ((1ull << b) - 1);
    return ((x ^ lo) << 1) | lo;
}

// Output for the pos0 (SEL = 0) or pos1 (SEL = 1) entry of a pair
// (a0 + i b0, a1 + i b1); expressions as in src/dmsim_amdgpu_omp.hpp:1019-1480.
template <int K>
__device__ __forceinline__ void out0(const double* c, double a0, double b0, double a1, double b1,
                                     double& yr, double& yi)
{
    if constexpr (K == PK_C1) {
        yr = (c[0] * a0) - (c[1] * b0) + (c[2] * a1) - (c[3] * b1);
        yi = (c[0] * b0) + (c[1] * a0) + (c[2] * b1) + (c[3] * a1);
    } else if constexpr (K == PK_X) {
        yr = a1; yi = b1;
    } else if constexpr (K == PK_Y) {
        yr = b1; yi = -a1;
    } else if constexpr (K == PK_H) {
        yr = kS2I * (a0 + a1); yi = kS2I * (b0 + b1);
    } else if constexpr (K == PK_SRN) {
        yr = 0.5 * (a0 + a1); yi = 0.5 * (b0 - b1);
    } else if constexpr (K == PK_RX) {
        yr = (c[0] * a0) - (c[1] * b1); yi = (c[0] * b0) + (c[1] * a1);
    } else if constexpr (K == PK_RY) {
        yr = (c[0] * a0) + (c[1] * a1); yi = (c[0] * b0) + (c[1] * b1);
    } else if constexpr (K == PK_W) {
        yr = kS2I * (a0 + b1); yi = kS2I * (b0 - a1);
    } else {  // pos1-only kinds keep pos0
        yr = a0; yi = b0;
    }
}

template <int K>
__device__ __forceinline__ void out1(const double* c, double a0, double b0, double a1, double b1,
                                     double& yr, double& yi)
{
    if constexpr (K == PK_C1) {
        yr = (c[4] * a0) - (c[5] * b0) + (c[6] * a1) - (c[7] * b1);
        yi = (c[4] * b0) + (c[5] * a0) + (c[6] * b1) + (c[7] * a1);
    } else if constexpr (K == PK_X) {
        yr = a0; yi = b0;
    } else if constexpr (K == PK_Y) {
        yr = -b0; yi = a0;
    } else if constexpr (K == PK_H) {
        yr = kS2I * (a0 - a1); yi = kS2I * (b0 - b1);
    } else if constexpr (K == PK_SRN) {
        yr = 0.5 * (a0 + a1); yi = 0.5 * (-b0 + b1);
    } else if constexpr (K == PK_RX) {
        yr = -(c[1] * b0) + (c[0] * a1); yi = (c[1] * a0) + (c[0] * b1);
    } else if constexpr (K == PK_RY) {
        yr = (c[2] * a0) + (c[3] * a1); yi = (c[2] * b0) + (c[3] * b1);
    } else if constexpr (K == PK_W) {
        yr = kS2I * (b0 + a1); yi = kS2I * (-a0 + b1);
    } else if constexpr (K == PK_Z) {
        yr = -a1; yi = -b1;
    } else if constexpr (K == PK_S) {
        yr = -b1; yi = a1;
    } else if constexpr (K == PK_SDG) {
        yr = b1; yi = -a1;
    } else if constexpr (K == PK_T) {
        yr = kS2I * (a1 - b1); yi = kS2I * (a1 + b1);
    } else if constexpr (K == PK_TDG) {
        yr = kS2I * (a1 + b1); yi = kS2I * (-a1 + b1);
    } else if constexpr (K == PK_R) {
        yr = -(b1 * c[0]); yi = a1 * c[0];
    } else if constexpr (K == PK_U1) {
        yr = (c[0] * a1) - (c[1] * b1); yi = (c[0] * b1) + (c[1] * a1);
    }
}

// CJ: evaluate on conjugated operands and conjugate the result (slow index).
template <int K, bool CJ>
__device__ __forceinline__ void eval0(const double* c, double a0, double b0, double a1, double b1,
                                      double& yr, double& yi)
{
    if constexpr (CJ) { b0 = -b0; b1 = -b1; }
    out0<K>(c, a0, b0, a1, b1, yr, yi);
    if constexpr (CJ) yi = -yi;
}

template <int K, bool CJ>
__device__ __forceinline__ void eval1(const double* c, double a0, double b0, double a1, double b1,
                                      double& yr, double& yi)
{
    if constexpr (CJ) { b0 = -b0; b1 = -b1; }
    out1<K>(c, a0, b0, a1, b1, yr, yi);
    if constexpr (CJ) yi = -yi;
}

__device__ __forceinline__ double2 ld2(const double* p)
{
    return *reinterpret_cast<const double2*>(p);
}

__device__ __forceinline__ void st2(double* p, double x, double y, bool wx, bool wy)
{
    if (wx && wy) {
        *reinterpret_cast<double2*>(p) = make_double2(x, y);
    } else if (wx) {
        p[0] = x;
    } else if (wy) {
        p[1] = y;
    }
}

struct Range {
    uint64_t first;   // this wave's (or thread's) first work item
    uint64_t stride;  // number of waves (or threads) in the grid
    int lane;
};

// Set bit c of x after inserting it (x's bits >= c move up one).
__device__ __forceinline__ uint64_t ins1(uint64_t x, int c) { return ins0(x, c) | (1ull << c); }

// b >= 7.  cbp = control bit in pair-index space (-1: none).  Only the pairs
// with the control bit set are visited: whole wave units where it lies above
// a unit (cbp >= 7), lane by lane where a lane's two pairs share it
// (1 <= cbp <= 6: runs of 2^cbp pairs); cbp = 0 splits every lane's pair of
// pairs, so those lanes load both and store one.
template <int K, bool CJ>
__device__ void sweep_wide(double* __restrict__ re, double* __restrict__ im, const double* c,
                           int b, int cbp, uint64_t N, Range w)
{
    const bool coarse = cbp >= 7;
    const bool sparse = cbp >= 1 && cbp < 7 && N >= 512;
    const uint64_t units = (coarse || sparse) ? (N >> 9) : (N >> 8);
    const uint64_t off1 = 1ull << b;
    for (uint64_t v = w.first; v < units; v += w.stride) {
        uint64_t p0;
        if (coarse)
            p0 = (ins1(v, cbp - 7) << 7) + 2 * (uint64_t)w.lane;
        else if (sparse)
            p0 = ins1((v << 7) + 2 * (uint64_t)w.lane, cbp);
        else
            p0 = (v << 7) + 2 * (uint64_t)w.lane;
        const uint64_t pos0 = ins0(p0, b);
        const uint64_t pos1 = pos0 + off1;
        bool wa = true, wb = true;
        if (cbp >= 0 && !coarse && !sparse) {
            wa = (p0 >> cbp) & 1;
            wb = ((p0 + 1) >> cbp) & 1;
        }
        if constexpr (pos1_only(K)) {
            const double2 r1 = ld2(re + pos1), i1 = ld2(im + pos1);
            double ya_r, ya_i, yb_r, yb_i;
            eval1<K, CJ>(c, 0.0, 0.0, r1.x, i1.x, ya_r, ya_i);
            eval1<K, CJ>(c, 0.0, 0.0, r1.y, i1.y, yb_r, yb_i);
            st2(re + pos1, ya_r, yb_r, wa, wb);
            st2(im + pos1, ya_i, yb_i, wa, wb);
        } else {
            const double2 r0 = ld2(re + pos0), i0 = ld2(im + pos0);
            const double2 r1 = ld2(re + pos1), i1 = ld2(im + pos1);
            double a0r, a0i, a1r, a1i, b0r, b0i, b1r, b1i;
            eval0<K, CJ>(c, r0.x, i0.x, r1.x, i1.x, a0r, a0i);
            eval1<K, CJ>(c, r0.x, i0.x, r1.x, i1.x, a1r, a1i);
            eval0<K, CJ>(c, r0.y, i0.y, r1.y, i1.y, b0r, b0i);
            eval1<K, CJ>(c, r0.y, i0.y, r1.y, i1.y, b1r, b1i);
            st2(re + pos0, a0r, b0r, wa, wb);
            st2(im + pos0, a0i, b0i, wa, wb);
            st2(re + pos1, a1r, b1r, wa, wb);
            st2(im + pos1, a1i, b1i, wa, wb);
        }
    }
}

// b <= 6.  cb = control bit in linear-index space (-1: none).  Only entries
// with the control bit set are visited where it lies above a unit (cb >= 8),
// above a half unit (cb = 7) or, lane by lane, above a lane's entry pair
// (1 <= cb <= 6: the lanes walk the set half, so a pair partner sits in lane
// ^ 2^(bx-1) with bx the target bit once the control bit is taken out);
// cb = 0 splits every lane's pair, so those lanes store one entry.
template <int K, bool CJ>
__device__ void sweep_narrow(double* __restrict__ re, double* __restrict__ im, const double* c,
                             int b, int cb, uint64_t N, Range w)
{
    const bool coarse = cb >= 8;
    const bool sparse = cb >= 1 && cb < 7 && N >= 512;
    const uint64_t units = (coarse || sparse) ? (N >> 9) : (N >> 8);
    const int bx = sparse && cb < b ? b - 1 : b;
    const int lane = w.lane;
    for (uint64_t v = w.first; v < units; v += w.stride) {
        const uint64_t u = coarse ? ins1(v, cb - 8) : v;
#pragma unroll
        for (int half = 0; half < 2; half++) {
            if (cb == 7 && half == 0) continue;  // control bit 7 clear in this half
            const uint64_t x = (u << 8) + (uint64_t)half * 128 + 2 * (uint64_t)lane;
            const uint64_t e = sparse ? ins1(x, cb) : x;
            const double2 r = ld2(re + e), i = ld2(im + e);
            double xr, xi, yr, yi;  // new values of the lane's .x and .y entries
            bool wx = true, wy = true;
            if (bx == 0) {
                eval0<K, CJ>(c, r.x, i.x, r.y, i.y, xr, xi);
                eval1<K, CJ>(c, r.x, i.x, r.y, i.y, yr, yi);
                if constexpr (pos1_only(K)) wx = false;
            } else {
                const int m = 1 << (bx - 1);
                const double prx = __shfl_xor(r.x, m), pry = __shfl_xor(r.y, m);
                const double pix = __shfl_xor(i.x, m), piy = __shfl_xor(i.y, m);
                if (((lane >> (bx - 1)) & 1) == 0) {
                    eval0<K, CJ>(c, r.x, i.x, prx, pix, xr, xi);
                    eval0<K, CJ>(c, r.y, i.y, pry, piy, yr, yi);
                    if constexpr (pos1_only(K)) wx = wy = false;
                } else {
                    eval1<K, CJ>(c, prx, pix, r.x, i.x, xr, xi);
                    eval1<K, CJ>(c, pry, piy, r.y, i.y, yr, yi);
                }
            }
            if (cb >= 0 && cb < 7 && !sparse) {
                if (cb == 0) {
                    wx = false;
                } else if (((lane >> (cb - 1)) & 1) == 0) {
                    wx = wy = false;
                }
            }
            st2(re + e, xr, yr, wx, wy);
            st2(im + e, xi, yi, wx, wy);
        }
    }
}

// 4^n < 256: one thread per pair.
template <int K, bool CJ>
__device__ void sweep_tiny(double* __restrict__ re, double* __restrict__ im, const double* c,
                           int b, int cb, uint64_t N, Range t)
{
    for (uint64_t p = t.first; p < (N >> 1); p += t.stride) {
        const uint64_t pos0 = ins0(p, b);
        const uint64_t pos1 = pos0 | (1ull << b);
        if (cb >= 0 && !((pos0 >> cb) & 1)) continue;
        const double a0 = re[pos0], b0 = im[pos0], a1 = re[pos1], b1 = im[pos1];
        double y0r, y0i, y1r, y1i;
        eval0<K, CJ>(c, a0, b0, a1, b1, y0r, y0i);
        eval1<K, CJ>(c, a0, b0, a1, b1, y1r, y1i);
        if constexpr (!pos1_only(K)) {
            re[pos0] = y0r;
            im[pos0] = y0i;
        }
        re[pos1] = y1r;
        im[pos1] = y1i;
    }
}

// General 4x4 on linear bits (b1 = high operand, b2 = low operand); quad
// order and summation order as C2_GATE (src/dmsim_amdgpu_omp.hpp:1043-1137).
template <bool CJ>
__device__ void sweep_quad(double* __restrict__ re, double* __restrict__ im,
                           const double* __restrict__ e,
                           int b1, int b2, uint64_t N, Range t)
{
    const int hi = b1 > b2 ? b1 : b2;
    const int lo = b1 > b2 ? b2 : b1;
    for (uint64_t k = t.first; k < (N >> 2); k += t.stride) {
        const uint64_t base = ins0(ins0(k, lo), hi);
        uint64_t pos[4];
        pos[0] = base;
        pos[1] = base + (1ull << b2);
        pos[2] = base + (1ull << b1);
        pos[3] = base + (1ull << b1) + (1ull << b2);
        double xr[4], xi[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            xr[j] = re[pos[j]];
            xi[j] = CJ ? -im[pos[j]] : im[pos[j]];
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const double* er = e + 8 * r;
            double s = (er[0] * xr[0]) - (er[1] * xi[0]);
            s = s + (er[2] * xr[1]);
            s = s - (er[3] * xi[1]);
            s = s + (er[4] * xr[2]);
            s = s - (er[5] * xi[2]);
            s = s + (er[6] * xr[3]);
            s = s - (er[7] * xi[3]);
            double v = (er[0] * xi[0]) + (er[1] * xr[0]);
            v = v + (er[2] * xi[1]);
            v = v + (er[3] * xr[1]);
            v = v + (er[4] * xi[2]);
            v = v + (er[5] * xr[2]);
            v = v + (er[6] * xi[3]);
            v = v + (er[7] * xr[3]);
            re[pos[r]] = s;
            im[pos[r]] = CJ ? -v : v;
        }
    }
}

// Persistent-loop dispatch of one pair primitive (4x4 sweeps are launched
// separately; see plain_segments()).
template <bool SLOW>
__device__ void run_pair(const PrimOp* __restrict__ op, double* re, double* im, int n)
{
    const uint64_t N = 1ull << (2 * n);
    const int shift = SLOW ? n : 0;
    const int kind = op->kind;
    const int b = op->q0 + shift;
    const int cb = op->q1 >= 0 ? op->q1 + shift : -1;
    const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t gthreads = (uint64_t)gridDim.x * blockDim.x;
    const Range waves{gtid >> 6, gthreads >> 6, (int)(threadIdx.x & 63)};
    const Range threads{gtid, gthreads, 0};
    double c[8];
#pragma unroll
    for (int i = 0; i < 8; i++) c[i] = op->c[i];
#define DMX_PAIR_CASE(KK)                                                             \
    case KK:                                                                          \
        if (N < 256)                                                                  \
            sweep_tiny<KK, SLOW>(re, im, c, b, cb, N, threads);                       \
        else if (b >= 7)                                                              \
            sweep_wide<KK, SLOW>(re, im, c, b, cb < 0 ? -1 : (cb < b ? cb : cb - 1), N, \
                                 waves);                                              \
        else                                                                          \
            sweep_narrow<KK, SLOW>(re, im, c, b, cb, N, waves);                       \
        break;
    switch (kind) {
        DMX_PAIR_CASE(PK_C1)
        DMX_PAIR_CASE(PK_X)
        DMX_PAIR_CASE(PK_Y)
        DMX_PAIR_CASE(PK_H)
        DMX_PAIR_CASE(PK_SRN)
        DMX_PAIR_CASE(PK_RX)
        DMX_PAIR_CASE(PK_RY)
        DMX_PAIR_CASE(PK_W)
        DMX_PAIR_CASE(PK_Z)
        DMX_PAIR_CASE(PK_S)
        DMX_PAIR_CASE(PK_SDG)
        DMX_PAIR_CASE(PK_T)
        DMX_PAIR_CASE(PK_TDG)
        DMX_PAIR_CASE(PK_R)
        DMX_PAIR_CASE(PK_U1)
    default:
        break;
    }
#undef DMX_PAIR_CASE
}

enum Path { P_TINY = 0, P_NARROW = 1, P_WIDE = 2 };

}  // namespace

// Persistent gate loop: primitives [first, last) of one pass in a single
// cooperative launch with a grid barrier after every primitive (the
// reference's BSP structure, src/dmsim_amdgpu_omp.hpp:831-837, :1004-1016).
__global__ void __launch_bounds__(256) k_plain_persistent(double* re, double* im,
                                                          const PrimOp* __restrict__ ops,
                                                          int first, int last, int n, int slow)
{
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    for (int t = first; t < last; t++) {
        if (slow)
            run_pair<true>(ops + t, re, im, n);
        else
            run_pair<false>(ops + t, re, im, n);
        if (t + 1 < last) grid.sync();
    }
}

// One pair primitive per launch, specialised by kind, index side and access
// path (each instance keeps its own small register budget).
template <int K, bool SLOW, int PATH>
__global__ void __launch_bounds__(256) k_sweep(double* re, double* im,
                                               const PrimOp* __restrict__ op, int n)
{
    const uint64_t N = 1ull << (2 * n);
    const int shift = SLOW ? n : 0;
    const int b = op->q0 + shift;
    const int cb = op->q1 >= 0 ? op->q1 + shift : -1;
    const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t gthreads = (uint64_t)gridDim.x * blockDim.x;
    double c[8];
#pragma unroll
    for (int i = 0; i < 8; i++) c[i] = op->c[i];
    if constexpr (PATH == P_TINY) {
        sweep_tiny<K, SLOW>(re, im, c, b, cb, N, Range{gtid, gthreads, 0});
    } else if constexpr (PATH == P_WIDE) {
        sweep_wide<K, SLOW>(re, im, c, b, cb < 0 ? -1 : (cb < b ? cb : cb - 1), N,
                            Range{gtid >> 6, gthreads >> 6, (int)(threadIdx.x & 63)});
    } else {
        sweep_narrow<K, SLOW>(re, im, c, b, cb, N,
                              Range{gtid >> 6, gthreads >> 6, (int)(threadIdx.x & 63)});
    }
}

// One 4x4 primitive (C2) per launch.
template <bool SLOW>
__global__ void __launch_bounds__(256) k_quad(double* re, double* im,
                                              const PrimOp* __restrict__ op, int n)
{
    const uint64_t N = 1ull << (2 * n);
    const int shift = SLOW ? n : 0;
    const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t gthreads = (uint64_t)gridDim.x * blockDim.x;
    sweep_quad<SLOW>(re, im, op->c, op->q0 + shift, op->q1 + shift, N,
                     Range{gtid, gthreads, 0});
}

// One channel superoperator (PK_SUP): the 4x4 acts on the entry quad of row
// bit q0 + n and column bit q0, in the quad order of C2.
__global__ void __launch_bounds__(256) k_super(double* re, double* im,
                                               const PrimOp* __restrict__ op, int n)
{
    const uint64_t N = 1ull << (2 * n);
    const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t gthreads = (uint64_t)gridDim.x * blockDim.x;
    sweep_quad<false>(re, im, op->c, op->q0 + n, op->q0, N, Range{gtid, gthreads, 0});
}

// In-place conjugate transpose with 32x32 LDS tiles; block (tj, ti) swaps tile
// (ti, tj) with tile (tj, ti); blocks below the diagonal exit.
__global__ void __launch_bounds__(256) k_conj_transpose(double* re, double* im, int n)
{
    constexpr int TS = 32;
    __shared__ double s[4][TS][TS + 1];
    const uint64_t dim = 1ull << n;
    const int ts = dim < TS ? (int)dim : TS;
    const int ti = blockIdx.y, tj = blockIdx.x;
    if (ti > tj) return;
    const int tx = threadIdx.x % TS, ty0 = threadIdx.x / TS;  // 8 rows per step
    for (int y = ty0; y < ts; y += 256 / TS) {
        if (tx < ts) {
            const uint64_t a = (uint64_t)(ti * ts + y) * dim + (uint64_t)tj * ts + tx;
            const uint64_t bq = (uint64_t)(tj * ts + y) * dim + (uint64_t)ti * ts + tx;
            s[0][y][tx] = re[a];
            s[1][y][tx] = im[a];
            s[2][y][tx] = re[bq];
            s[3][y][tx] = im[bq];
        }
    }
    __syncthreads();
    for (int y = ty0; y < ts; y += 256 / TS) {
        if (tx < ts) {
            const uint64_t a = (uint64_t)(ti * ts + y) * dim + (uint64_t)tj * ts + tx;
            const uint64_t bq = (uint64_t)(tj * ts + y) * dim + (uint64_t)ti * ts + tx;
            re[a] = s[2][tx][y];
            im[a] = -s[3][tx][y];
            if (ti != tj) {
                re[bq] = s[0][tx][y];
                im[bq] = -s[1][tx][y];
            }
        }
    }
}

__global__ void k_diag(const double* __restrict__ re, int n, double* __restrict__ out)
{
    const uint64_t dim = 1ull << n;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim;
         i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = re[i * dim + i];
}

// out[t] = entry idx[t] (linear index r*dim + c of the reference's layout)
// for t < count; with `transposed` the device holds conj of the transpose.
__global__ void k_gather(const double* __restrict__ re, const double* __restrict__ im, int n,
                         const uint64_t* __restrict__ idx, uint64_t count, int transposed,
                         double* __restrict__ out_re, double* __restrict__ out_im)
{
    const uint64_t dim = 1ull << n;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count;
         t += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t p = idx[t];
        if (transposed) p = (p % dim) * dim + p / dim;
        out_re[t] = re[p];
        out_im[t] = transposed ? -im[p] : im[p];
    }
}

// ---------------------------------------------------------------- host ----

namespace {

using SweepFn = void (*)(double*, double*, const PrimOp*, int);

template <int K, bool SLOW>
constexpr SweepFn pick(int path)
{
    return path == P_TINY ? k_sweep<K, SLOW, P_TINY>
                          : (path == P_WIDE ? k_sweep<K, SLOW, P_WIDE> : k_sweep<K, SLOW, P_NARROW>);
}

template <bool SLOW>
SweepFn sweep_fn(int kind, int path)
{
    switch (kind) {
    case PK_C1: return pick<PK_C1, SLOW>(path);
    case PK_X: return pick<PK_X, SLOW>(path);
    case PK_Y: return pick<PK_Y, SLOW>(path);
    case PK_H: return pick<PK_H, SLOW>(path);
    case PK_SRN: return pick<PK_SRN, SLOW>(path);
    case PK_RX: return pick<PK_RX, SLOW>(path);
    case PK_RY: return pick<PK_RY, SLOW>(path);
    case PK_W: return pick<PK_W, SLOW>(path);
    case PK_Z: return pick<PK_Z, SLOW>(path);
    case PK_S: return pick<PK_S, SLOW>(path);
    case PK_SDG: return pick<PK_SDG, SLOW>(path);
    case PK_T: return pick<PK_T, SLOW>(path);
    case PK_TDG: return pick<PK_TDG, SLOW>(path);
    case PK_R: return pick<PK_R, SLOW>(path);
    case PK_U1: return pick<PK_U1, SLOW>(path);
    default: return nullptr;
    }
}

}  // namespace

hipError_t plain_persistent_grid(int n, int* blocks)
{
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(k_plain_persistent), 256, 0);
    if (e != hipSuccess) return e;
    int dev = 0;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    // The grid barrier after every primitive costs more the more blocks take
    // part: N / 8192 blocks, at least 16, at most 256 (n <= 12) or 512, one wave
    // per 256 entries at most, measured fastest from n = 8 to n = 14 (16.2 ms
    // for 1000 gates at n = 8, 31 ms for 600 at n = 10, 52 ms for 200 at n = 12,
    // 342 ms for 120 at n = 14, against 23, 170, 100 and 365 ms with every
    // co-resident block).
    const uint64_t N = 1ull << (2 * n);
    uint64_t g = N / 8192;
    if (g < 16) g = 16;
    if (g > (n <= 12 ? 256u : 512u)) g = n <= 12 ? 256 : 512;
    const uint64_t most = (N / 1024) > 0 ? (N / 1024) : 1;  // one wave per 256 entries
    if (g > most) g = most;
    if (g > (uint64_t)per_cu * (uint64_t)cus) g = (uint64_t)per_cu * (uint64_t)cus;
    if (g < 1) g = 1;
    *blocks = (int)g;
    return hipSuccess;
}

hipError_t launch_plain_persistent(double* re, double* im, const PrimOp* d_ops, int first,
                                   int last, int n, bool slow, int blocks, hipStream_t stream)
{
    int sl = slow ? 1 : 0;
    void* args[] = {&re, &im, const_cast<PrimOp**>(&d_ops), &first, &last, &n, &sl};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_plain_persistent),
                                      dim3(blocks), dim3(256), args, 0, stream);
}

hipError_t launch_plain_one(double* re, double* im, const PrimOp* d_op, const PrimOp& host_op,
                            int n, bool slow, int max_blocks, hipStream_t stream)
{
    const uint64_t N = 1ull << (2 * n);
    const int shift = slow ? n : 0;
    if (host_op.kind == PK_SUP) {
        const uint64_t need = (N / 4 + 255) / 256;
        const unsigned blocks = (unsigned)(need < (uint64_t)max_blocks ? need : max_blocks);
        k_super<<<blocks, 256, 0, stream>>>(re, im, d_op, n);
        return hipGetLastError();
    }
    if (host_op.kind == PK_C2) {
        const uint64_t need = (N / 4 + 255) / 256;
        const unsigned blocks = (unsigned)(need < (uint64_t)max_blocks ? need : max_blocks);
        if (slow)
            k_quad<true><<<blocks, 256, 0, stream>>>(re, im, d_op, n);
        else
            k_quad<false><<<blocks, 256, 0, stream>>>(re, im, d_op, n);
        return hipGetLastError();
    }
    const int b = host_op.q0 + shift;
    int path;
    uint64_t need;
    if (N < 256) {
        path = P_TINY;
        need = 1;
    } else {
        path = b >= 7 ? P_WIDE : P_NARROW;
        need = (N / 256 + 3) / 4;  // 4 waves per block, one 256-entry unit per wave
    }
    const unsigned blocks = (unsigned)(need < (uint64_t)max_blocks ? need : max_blocks);
    SweepFn fn = slow ? sweep_fn<true>(host_op.kind, path) : sweep_fn<false>(host_op.kind, path);
    if (!fn) return hipErrorInvalidValue;
    fn<<<blocks, 256, 0, stream>>>(re, im, d_op, n);
    return hipGetLastError();
}

hipError_t launch_conj_transpose(double* re, double* im, int n, hipStream_t stream)
{
    const uint64_t dim = 1ull << n;
    const unsigned t = dim < 32 ? 1u : (unsigned)(dim / 32);
    k_conj_transpose<<<dim3(t, t), 256, 0, stream>>>(re, im, n);
    return hipGetLastError();
}

hipError_t launch_gather(const double* re, const double* im, int n, const uint64_t* d_idx,
                         uint64_t count, bool transposed, double* d_re, double* d_im,
                         hipStream_t stream)
{
    if (count == 0) return hipSuccess;
    const uint64_t need = (count + 255) / 256;
    const unsigned blocks = (unsigned)(need < 4096 ? need : 4096);
    k_gather<<<blocks, 256, 0, stream>>>(re, im, n, d_idx, count, transposed ? 1 : 0, d_re, d_im);
    return hipGetLastError();
}

hipError_t launch_diag(const double* re, int n, double* out, hipStream_t stream)
{
    const uint64_t dim = 1ull << n;
    const unsigned blocks = (unsigned)((dim + 255) / 256);
    k_diag<<<blocks, 256, 0, stream>>>(re, n, out);
    return hipGetLastError();
}

}  // namespace dmx
