// gi_valu_ceiling.hip -- calibration of the FP32 vector issue ceiling of the
// GPU this library runs on (gfx950), for reading the all-pairs kernel's
// VALU counters against.  Two kernels, each a loop of independent chains
// (no memory traffic but one float per thread at the end):
//   fma: 32 independent v_fmac_f32 chains per lane,
//   rsq: 16 independent v_rsq_f32 chains per lane,
// at 8 wavefronts per SIMD.  Run it under rocprofv3 --pmc SQ_INSTS_VALU
// GRBM_GUI_ACTIVE: VALU instructions per SIMD per cycle is the ceiling a
// kernel of plain (non-packed) FP32 operations can issue at.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

template <bool RSQ, int NC>
__global__ __launch_bounds__(256) void k_valu_chains(float* __restrict__ out, int iters, float b, float c) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    float a[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) a[k] = tid * 1e-7f + (float)k + 1.0f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            if constexpr (RSQ)
                a[k] = __builtin_amdgcn_rsqf(a[k]);
            else
                a[k] = __builtin_fmaf(a[k], b, c);
        }
    }
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < NC; ++k) sum += a[k];
    out[tid] = sum;
}

#define CHECK(call)                                                                           \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            std::fprintf(stderr, "%s failed: %s\n", #call, hipGetErrorString(e_));             \
            std::exit(1);                                                                     \
        }                                                                                     \
    } while (0)

template <bool RSQ, int NC>
static void run(const char* name, float* d, int grid, int iters) {
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    CHECK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL((k_valu_chains<RSQ, NC>), dim3(grid), dim3(256), 0, 0, d, iters, 0.999f, 1e-3f);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    const double instr = (double)grid * 4 * iters * NC;  // wave-level VALU instructions in the loop
    std::printf("%-4s %2d chains: %8.3f ms, %.4e wave-level VALU instructions in the loop\n", name, NC, ms, instr);
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
}

int main() {
    int dev = 0, cus = 0;
    CHECK(hipGetDevice(&dev));
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = cus * 8;  // 256-thread workgroups: 4 wavefronts each, 8 per SIMD
    float* d = nullptr;
    CHECK(hipMalloc((void**)&d, (size_t)grid * 256 * sizeof(float)));
    run<false, 32>("fma", d, grid, 8);      // warm-up
    run<false, 32>("fma", d, grid, 10000);
    run<true, 16>("rsq", d, grid, 10000);
    CHECK(hipDeviceSynchronize());
    CHECK(hipFree(d));
    std::printf("%d CUs, %d workgroups of 256\n", cus, grid);
    return 0;
}
