// sweep.hpp -- host launchers for the plain-method kernels (sweep.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "prim.hpp"
#include "prof.hpp"

namespace dmx {

// Blocks of 256 threads for the persistent loop: 4^n / 8192 clamped to
// 16..256 (n <= 12) or 16..512, at most one wave per 256 entries of a plane
// and at most the co-resident maximum on this device.
hipError_t plain_persistent_grid(int n, int* blocks);

// Pair primitives [first, last) of one pass (slow: row index) in one
// cooperative launch.  4x4 primitives are not accepted here.
hipError_t launch_plain_persistent(double* re, double* im, const PrimOp* d_ops, int first,
                                   int last, int n, bool slow, int blocks, hipStream_t stream);

// One primitive sweep; host_op is the host copy of *d_op (selects the kernel).
hipError_t launch_plain_one(double* re, double* im, const PrimOp* d_op, const PrimOp& host_op,
                            int n, bool slow, int max_blocks, hipStream_t stream);

// Kernel class and algorithmic bytes (entries a primitive reads and writes,
// both planes) of one primitive sweep over 4^n entries.
inline int prim_class(const PrimOp& p)
{
    if (p.kind == PK_C2 || p.kind == PK_SUP) return KC_QUAD;
    if (p.kind == PK_X && p.q1 >= 0) return KC_CX;
    return pos1_only(p.kind) ? KC_PAIR_POS1 : KC_PAIR_DENSE;
}
inline double prim_bytes(const PrimOp& p, int n)
{
    const double full = 32.0 * (double)(1ull << (2 * n));
    const int c = prim_class(p);
    return (c == KC_CX || c == KC_PAIR_POS1) ? full / 2 : full;
}

// In-place rho <- conj(rho^T).
hipError_t launch_conj_transpose(double* re, double* im, int n, hipStream_t stream);

// d_re[t], d_im[t] = entry d_idx[t] (< 4^n, reference layout) for t < count.
hipError_t launch_gather(const double* re, const double* im, int n, const uint64_t* d_idx,
                         uint64_t count, bool transposed, double* d_re, double* d_im,
                         hipStream_t stream);

// out[i] = re[i*dim + i].
hipError_t launch_diag(const double* re, int n, double* out, hipStream_t stream);

}  // namespace dmx
