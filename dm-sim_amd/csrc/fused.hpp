// fused.hpp -- device record and launcher of the dense 4-qubit block sweep.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dmx {

// A dense 4-qubit block U (16 x 16, row-major, local index bit t <-> the t-th
// lowest qubit of qmask), optionally followed by a permutation of the basis
// (trailing X, CX and SWAP gates): when perm_on, entry (i, j) of U rho U^dagger
// goes to (pi(i), pi(j)) with pi(x) = L x + b over GF(2), L given by its
// columns perm_L[q] = the image of bit q (q < n), and the sweep writes into a
// second pair of planes.  pi leaves bit 0 in place (perm_L[0] = 1).
struct FusedBlock {
    uint32_t qmask;  // exactly four qubits
    uint32_t m;      // super-tile column bits (the tile schedule), see tile_bits_ok
    uint32_t perm_on;
    uint32_t perm_b;
    double ur[256];
    double ui[256];
    uint32_t perm_L[16];
    // Hermitian sweeps only: the next Hermitian block's index bits in this
    // block's output order (its sweep reads only the tiles on and above its
    // diagonal, so only those entries are written), 0 = write every entry.
    uint32_t next_q;
    uint32_t pad;
};

// One circuit's rho and the block to apply to it (batched launches).
struct BatchEntry {
    double* re;  // rho before the block
    double* im;
    double* re_out;  // after it: re/im, or the other pair for a permuting block
    double* im_out;
    const FusedBlock* blk;
};

// Default super-tile width for n qubits.
inline int default_tile_bits(int n) { return n < 8 ? n : 8; }

// Super-tile column bits are the block's four qubits plus the lowest other
// qubits, m in all.  Global accesses move column pairs (bit 0 must be a
// column bit), so m = 4 needs qubit 0 in the block.
inline bool tile_bits_ok(int n, int m, uint32_t qmask)
{
    return m >= 4 && m <= (n < 8 ? n : 8) && (m > 4 || (qmask & 1u));
}

// A dense 5-qubit block (32 x 32, the same conventions).
struct FusedBlock5 {
    uint32_t qmask;  // exactly five qubits
    uint32_t m;      // super-tile column bits, see tile_bits5_ok
    uint32_t pad[2];
    double ur[1024];
    double ui[1024];
};

inline int default_tile_bits5(int n) { return n < 7 ? n : 7; }

// Column bits = the block's five qubits plus the lowest others, m in all;
// m = 5 needs qubit 0 in the block (column pairs move together).
inline bool tile_bits5_ok(int n, int m, uint32_t qmask)
{
    return m >= 5 && m <= (n < 8 ? n : 8) && (m > 5 || (qmask & 1u));
}

// rho <- U rho U^dagger for one 5-qubit block (n >= 5), staged in super-tiles
// of 32 rows x 2^m columns.
hipError_t launch_fused5(double* re, double* im, const FusedBlock5* d_blk, int n, int m,
                         int max_blocks, hipStream_t stream);

// rho <- U rho U^dagger for one block (n >= 4), staged in super-tiles of
// 16 rows x 2^m columns.
hipError_t launch_fused4(double* re, double* im, const FusedBlock* d_blk, int n, int m,
                         int max_blocks, hipStream_t stream);

// The same for a block with a trailing permutation (perm_on; n <= 16): reads
// rho from (re, im), writes the permuted result to (re_out, im_out), which
// must not overlap the input.
hipError_t launch_fused4_perm(const double* re, const double* im, double* re_out, double* im_out,
                              const FusedBlock* d_blk, int n, int m, int max_blocks,
                              hipStream_t stream);

// Hermitian form of the block sweep for n = 8..16 (rho Hermitian): reads
// the tiles on and above the diagonal of (re, im) and writes every entry of
// (re_out, im_out), entry (i, j) of U rho U^dagger to (phi(i), phi(j)) with
// phi(x) = L x + b (perm_L, perm_b).  L must map the span of the block's
// qmask bits onto bits 0..3 (not checked here).  No overlap of input and output.
hipError_t launch_fused4_herm(const double* re, const double* im, double* re_out,
                              double* im_out, const FusedBlock* d_blk, int n, int max_blocks,
                              hipStream_t stream);

// rho's index bits relabelled by a permutation p (perm_L[q] = 1 << p(q)),
// out of place, n = 8..16.
hipError_t launch_relabel(const double* re, const double* im, double* re_out, double* im_out,
                          const FusedBlock* d_blk, int n, int max_blocks, hipStream_t stream);

// The same for `count` circuits of n qubits each (one launch of about
// max_blocks workgroups in all).
// Batched Hermitian block sweeps and relabellings (n = 8..16): entry c of
// d_entries is circuit c's (rho, second planes, block); about max_blocks
// workgroups in all, each looping over items of its circuit.  Out of place.
hipError_t launch_fused4_herm_batch(const BatchEntry* d_entries, int count, int n,
                                    int max_blocks, hipStream_t stream);
hipError_t launch_relabel_batch(const BatchEntry* d_entries, int count, int n, int max_blocks,
                                hipStream_t stream);
hipError_t launch_fused4_batch(const BatchEntry* d_entries, int count, int n, int m,
                               int max_blocks, hipStream_t stream);

}  // namespace dmx
