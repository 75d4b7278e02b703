// lower.hpp -- API gate records and their lowering into primitive sweeps.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "prim.hpp"

namespace dmx {

// Op codes: enum OP of the reference (src/dmsim_amdgpu_omp.hpp:51-57), then
// the matrix extensions (arbitrary k-qubit matrices, k = 1..5).
enum ApiOp : int32_t {
    OP_U3, OP_U2, OP_U1, OP_CX, OP_ID, OP_X, OP_Y, OP_Z, OP_H, OP_S,
    OP_SDG, OP_T, OP_TDG, OP_RX, OP_RY, OP_RZ, OP_CZ, OP_CY, OP_SWAP, OP_CH,
    OP_CCX, OP_CSWAP, OP_CRX, OP_CRY, OP_CRZ, OP_CU1, OP_CU3, OP_RXX, OP_RZZ, OP_RCCX,
    OP_RC3X, OP_C3X, OP_C3SQRTX, OP_C4X, OP_R, OP_SRN, OP_W, OP_RYY,
    OP_MAT = 38,    // arbitrary matrix on k qubits (k in 1..5), see Gate::mat
    OP_KRAUS = 39,  // single-qubit channel sum_i K_i rho K_i^dagger, see Gate::mat
    OP_COUNT
};

extern const char* const kOpNames[OP_COUNT];

// Number of qubit operands of an op (factories, src/dmsim_amdgpu_omp.hpp:594-782).
int op_arity(int op);

// One API gate: the reference's Gate record (op, qb0..qb4, theta, phi,
// lambda; src/dmsim_amdgpu_omp.hpp:108-206) plus, for OP_MAT, a dense
// 2^k x 2^k matrix with qb[0] as the most significant bit of its index.
struct Gate {
    int32_t op = OP_ID;
    uint32_t qb[5] = {0, 0, 0, 0, 0};
    double theta = 0, phi = 0, lambda = 0;
    int32_t k = 0;              // OP_MAT only
    std::vector<double> mre;    // OP_MAT: the matrix, row-major; OP_KRAUS: the 4x4
    std::vector<double> mim;    // superoperator on (row bit, column bit) of the stored
                                // conj(rho), then its conj-transposed-orientation twin
    std::string dump() const;   // Gate::dump format (src/dmsim_amdgpu_omp.hpp:185-191)
};

// Appends the primitive sweeps of one gate.  OP_MAT with k <= 2 lowers to a
// C1/C2 sweep; k > 2 is not a primitive (the fused executor handles it) and
// makes this return false.
bool lower_gate(const Gate& g, std::vector<PrimOp>& out);

// Dense matrix M (2^k x 2^k, row-major, index bit k-1-j <-> qubits[j]) that a
// gate applies to the fast index of a row, composed from its primitive
// sweeps exactly as lowered.  `qubits` are the gate's operands in order.
void gate_matrix(const Gate& g, int& k, std::vector<uint32_t>& qubits, std::vector<double>& re,
                 std::vector<double>& im);

}  // namespace dmx
