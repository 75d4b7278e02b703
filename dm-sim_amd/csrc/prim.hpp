// prim.hpp -- the primitive sweep record shared by host lowering and kernels.
//
// Every API gate is lowered on the host into a sequence of primitive sweeps,
// the same sequences the reference executes on the device
// (src/dmsim_amdgpu_omp.hpp:1019-1815: one OP_HEAD grid-stride loop plus a
// grid.sync() per primitive).  Gate coefficients (cos/sin of the angles) are
// computed on the host with the C library, so a sweep reads only its record.
#pragma once
#include <cstdint>

namespace dmx {

// Primitive kinds.  Pair kinds act on (pos0, pos1) = (bit b clear, bit b set);
// "pos1-only" kinds leave pos0 untouched.
enum PrimKind : int32_t {
    PK_C1 = 0,  // general 2x2, c[0..7] = e0r e0i e1r e1i e2r e2i e3r e3i
    PK_X,       // swap; with a control bit it is CX
    PK_Y,
    PK_H,
    PK_SRN,
    PK_RX,      // c[0] = cos(theta/2), c[1] = -sin(theta/2)
    PK_RY,      // c[0..3] = cos, -sin, sin, cos (of theta/2)
    PK_W,
    PK_Z,       // pos1-only from here
    PK_S,
    PK_SDG,
    PK_T,
    PK_TDG,
    PK_R,       // c[0] = phase
    PK_U1,      // c[0] = cos(lambda), c[1] = sin(lambda)
    PK_C2,      // general 4x4, c[0..31] = row-major complex (re, im) pairs
    PK_SUP,     // channel superoperator: 4x4 on (row bit q0 + n, column bit q0), applied
                // once (not per index side); pad = 0 for rho's stored orientation, 1 for
                // the conj-transposed one
    PK_NUM
};

inline constexpr bool pos1_only(int k) { return k >= PK_Z && k <= PK_U1; }

struct PrimOp {
    int32_t kind;
    int32_t q0;    // target qubit (C2: high qubit of the 4x4 index)
    int32_t q1;    // control qubit for PK_X as CX, low qubit for PK_C2, else -1
    int32_t pad;
    double c[32];
};
static_assert(sizeof(PrimOp) == 272, "PrimOp layout");

}  // namespace dmx
