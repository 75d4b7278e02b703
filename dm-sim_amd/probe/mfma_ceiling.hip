// mfma_ceiling.hip -- measures the FP64 MFMA throughput this GPU sustains, the
// compute ceiling of the fused kernels (k_fused4: 24, k_fused5: 192
// v_mfma_f64_16x16x4f64 per tile).  Every wave runs `iters` x 8 independent
// MFMAs; the grid is swept from 1 to 16 workgroups of 256 threads per CU
// (1 to 16 waves per SIMD).  Each thread writes one double of `out`
// (grid x 256 entries, allocated for the largest grid).
//   dm-sim_amd/mfma_ceiling            prints one line per grid
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_mfma_ceiling(double* out, int iters, double a0)
{
    constexpr int kChains = 8;
    d4 acc[kChains];
    for (int c = 0; c < kChains; c++) acc[c] = d4{0, 0, 0, 0};
    const double a = a0 + threadIdx.x * 1e-3, b = 1.0 - threadIdx.x * 1e-4;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int c = 0; c < kChains; c++)
            acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
    }
    double s = 0;
    for (int c = 0; c < kChains; c++) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

bool ok(hipError_t e, const char* what)
{
    if (e != hipSuccess) std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    return e == hipSuccess;
}

}  // namespace

int main()
{
    constexpr int kIters = 2048;
    constexpr int kMaxGrid = 4096;
    int dev = 0, cus = 0;
    if (!ok(hipGetDevice(&dev), "hipGetDevice") ||
        !ok(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count"))
        return 1;
    double* out = nullptr;
    if (!ok(hipMalloc(&out, (size_t)kMaxGrid * 256 * sizeof(double)), "hipMalloc")) return 1;
    hipEvent_t e0, e1;
    if (!ok(hipEventCreate(&e0), "event") || !ok(hipEventCreate(&e1), "event")) return 1;
    std::printf("FP64 MFMA (v_mfma_f64_16x16x4f64) throughput, %d CUs\n", cus);
    for (int per_cu : {1, 2, 3, 4, 8, 16}) {
        const int grid = per_cu * cus;
        if (grid > kMaxGrid) break;
        float ms = 0.f;
        for (int rep = 0; rep < 2; rep++) {  // the first launch warms up
            if (!ok(hipEventRecord(e0), "record")) return 1;
            k_mfma_ceiling<<<grid, 256>>>(out, kIters, 0.5);
            if (!ok(hipGetLastError(), "launch") || !ok(hipEventRecord(e1), "record") ||
                !ok(hipEventSynchronize(e1), "sync") || !ok(hipEventElapsedTime(&ms, e0, e1), "time"))
                return 1;
        }
        const double flops = (double)grid * 4 * kIters * 8 * 2048.0;
        std::printf("%2d workgroups/CU (%2d waves/SIMD): %8.3f ms  %6.1f TFLOP/s\n", per_cu, per_cu,
                    ms, flops / (ms * 1e-3) / 1e12);
    }
    (void)hipFree(out);
    return 0;
}
