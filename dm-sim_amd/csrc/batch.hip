// batch.hip -- kernels of the batched executor: pool reset and Pauli energies.
//
// Every circuit's rho lives in one device pool (re plane then im plane).  The
// reset writes zeros with a fill kernel and sets the |0..0><0..0| entries with
// a second one (no hipMemsetAsync, so the sequence can be captured in a graph
// and replayed).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "batch.hpp"

namespace dmx {

__global__ void k_pool_zero(double* __restrict__ pool, uint64_t n_double2)
{
    double2* p = reinterpret_cast<double2*>(pool);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_double2;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = make_double2(0.0, 0.0);
}

__global__ void k_pool_ground(double* __restrict__ pool, const uint64_t* __restrict__ re_off,
                              int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) pool[re_off[i]] = 1.0;
}

// E_c = sum_s w_s Re Tr(P_s rho_c).  With P|k> = i^nY (-1)^{|k & zy|} |k ^ x>
// (x: qubits with X or Y, zy: qubits with Z or Y) and the stored A = conj(rho),
// Tr(P rho) = sum_k i^nY (-1)^{|k & zy|} conj(A[k][k ^ x]).  One workgroup per
// circuit.
__global__ void __launch_bounds__(256) k_pauli_energy(const EnergyJob* __restrict__ jobs,
                                                      const PauliTerm* __restrict__ terms,
                                                      double* __restrict__ out)
{
    __shared__ double red[256];
    const EnergyJob job = jobs[blockIdx.x];
    const uint64_t dim = 1ull << job.n;
    double acc = 0.0;
    for (int s = job.first_term; s < job.first_term + job.n_terms; s++) {
        const PauliTerm t = terms[s];
        double part = 0.0;
        for (uint64_t k = threadIdx.x; k < dim; k += blockDim.x) {
            const uint64_t idx = k * dim + (k ^ t.xmask);
            const double ar = job.re[idx], ai = job.im[idx];
            const double sgn = (__popcll(k & t.zymask) & 1) ? -1.0 : 1.0;
            double v;
            switch (t.ny & 3) {  // Re(i^ny * (ar - i ai))
            case 0: v = ar; break;
            case 1: v = ai; break;
            case 2: v = -ar; break;
            default: v = -ai; break;
            }
            part += sgn * v;
        }
        acc += t.weight * part;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

__global__ void k_batch_diag(const EnergyJob* __restrict__ jobs,
                             const uint64_t* __restrict__ diag_off, double* __restrict__ out)
{
    const EnergyJob j = jobs[blockIdx.y];
    const uint64_t dim = 1ull << j.n;
    double* o = out + diag_off[blockIdx.y];
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim;
         i += (uint64_t)gridDim.x * blockDim.x)
        o[i] = j.re[i * dim + i];
}

hipError_t launch_batch_diag(const EnergyJob* d_jobs, const uint64_t* d_diag_off, int count,
                             double* d_out, hipStream_t stream)
{
    if (count < 1 || count > 65535) return hipErrorInvalidValue;
    k_batch_diag<<<dim3(64, (unsigned)count), 256, 0, stream>>>(d_jobs, d_diag_off, d_out);
    return hipGetLastError();
}

hipError_t launch_pool_reset(double* pool, uint64_t n_doubles, const uint64_t* d_re_off,
                             int count, hipStream_t stream)
{
    const uint64_t n2 = n_doubles / 2;
    const unsigned blocks = (unsigned)((n2 + 255) / 256 < 16384 ? (n2 + 255) / 256 : 16384);
    k_pool_zero<<<blocks, 256, 0, stream>>>(pool, n2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_pool_ground<<<(count + 255) / 256, 256, 0, stream>>>(pool, d_re_off, count);
    return hipGetLastError();
}

hipError_t launch_pauli_energy(const EnergyJob* d_jobs, const PauliTerm* d_terms, int count,
                               double* d_out, hipStream_t stream)
{
    k_pauli_energy<<<count, 256, 0, stream>>>(d_jobs, d_terms, d_out);
    return hipGetLastError();
}

}  // namespace dmx
