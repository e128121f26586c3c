// gi_issue_probe.hip -- calibration of the VALU issue cost of each operation in
// the all-pairs kernels' instruction mix on the GPU it runs on (gfx950).  Not
// part of the library: `make -C giant-impact_amd calib` builds
// giant-impact_amd/lib/gi_issue_probe.
//
// For each operation (FMA, mul, add, rsq, the DPP lane rotation) a kernel runs
// 32 independent chains of it per lane at 4 and 8 wavefronts per SIMD; timed
// with events, the wall time per wave-level instruction per SIMD is that
// operation's issue cost.  A last kernel runs the pair arithmetic of the
// mutual pass (3 sub, r^2 by 3 FMA, rsq, 2 mul, 3 + 3 FMA per pair, R pairs
// per j) for comparison.  Costs are printed in cycles of an assumed 2.4 GHz
// clock; the ceiling of a mix, sum(n_op x cost_op), does not depend on that
// assumption.  A last section times where a v_rsq_f32's consumer sits: the
// same instructions (12 independent FMAs, one v_rsq_f32 and a multiply of its
// result) with the multiply right after the v_rsq_f32 (after its one wait
// state), three instructions later, or twelve later (inline asm of register
// operations only, so the order is the one written).  Each thread reads in[]
// and writes out[tid] only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>

#define CHECK(call)                                                                   \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            std::fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));           \
            std::exit(1);                                                             \
        }                                                                             \
    } while (0)

constexpr int kChains = 32;

template <int OP>
__global__ __launch_bounds__(256) void k_op(float* __restrict__ out, const float* __restrict__ in, int iters) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const float x = in[tid & 1023], y = in[(tid + 7) & 1023];
    float a[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = x + (float)k;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) {
            if constexpr (OP == 0)
                a[k] = __builtin_fmaf(a[k], x, y);
            else if constexpr (OP == 1)
                a[k] = a[k] * x;
            else if constexpr (OP == 2)
                a[k] = a[k] + x;
            else if constexpr (OP == 3)
                a[k] = __builtin_amdgcn_rsqf(a[k]);
            else  // v_mov_b32_dpp wave_ror:1, as the mutual pass rotates its j-side sums
                a[k] = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, a[k]), 0x13c,
                                                                          0xf, 0xf, false));
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += a[k];
    out[tid] = s;
}

template <int R>
__global__ __launch_bounds__(256) void k_pairmix(float* __restrict__ out, const float* __restrict__ in, int iters) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    float xi[R], yi[R], zi[R], ax[R], ay[R], az[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        xi[r] = in[(tid + r) & 1023];
        yi[r] = in[(tid + 3 * r) & 1023];
        zi[r] = in[(tid + 5 * r) & 1023];
        ax[r] = ay[r] = az[r] = 0.0f;
    }
    float xj = in[(tid + 11) & 1023], yj = in[(tid + 13) & 1023], zj = in[(tid + 17) & 1023];
    float jx = 0.0f, jy = 0.0f, jz = 0.0f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const float dx = xj - xi[r], dy = yj - yi[r], dz = zj - zi[r];
            const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, 1e-20f)));
            const float ri = __builtin_amdgcn_rsqf(r2);
            const float s = ri * ri * ri;
            ax[r] = __builtin_fmaf(s, dx, ax[r]);
            ay[r] = __builtin_fmaf(s, dy, ay[r]);
            az[r] = __builtin_fmaf(s, dz, az[r]);
            jx = __builtin_fmaf(-s, dx, jx);
            jy = __builtin_fmaf(-s, dy, jy);
            jz = __builtin_fmaf(-s, dz, jz);
        }
        xj += 0.001f;  // the j moves: 3 adds per R pairs
        yj += 0.002f;
        zj += 0.003f;
    }
    float t = jx + jy + jz;
#pragma unroll
    for (int r = 0; r < R; ++r) t += ax[r] + ay[r] + az[r];
    out[tid] = t;
}

// 12 independent FMAs per iteration on a0..a11 (asm operands 0..11); t (12)
// takes v_rsq_f32 of u (13), m (14) the product t * t; x (15), y (16).
#define GI_FMA(k) "v_fma_f32 %" #k ", %" #k ", %15, %16\n\t"
#define GI_FMA3(a, b, c) GI_FMA(a) GI_FMA(b) GI_FMA(c)
#define GI_FMA12 GI_FMA3(0, 1, 2) GI_FMA3(3, 4, 5) GI_FMA3(6, 7, 8) GI_FMA3(9, 10, 11)
#define GI_RSQ "v_rsq_f32_e64 %12, %13\n\t"
#define GI_MUL "v_mul_f32_e32 %14, %12, %12\n\t"

template <int P>
__global__ __launch_bounds__(256) void k_rsq_place(float* __restrict__ out, const float* __restrict__ in, int iters) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const float x = in[tid & 1023], y = in[(tid + 5) & 1023];
    float a0 = x, a1 = x + 1.0f, a2 = x + 2.0f, a3 = x + 3.0f, a4 = x + 4.0f, a5 = x + 5.0f, a6 = x + 6.0f,
          a7 = x + 7.0f, a8 = x + 8.0f, a9 = x + 9.0f, a10 = x + 10.0f, a11 = x + 11.0f;
    float t = 0.0f, u = y + 1.0f, m = 0.0f;
    for (int it = 0; it < iters; ++it) {
#define GI_OPS                                                                                                   \
    : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7), "+v"(a8), "+v"(a9), "+v"(a10), \
      "+v"(a11), "+v"(t), "+v"(u), "+v"(m)                                                                         \
    : "v"(x), "v"(y)
        if constexpr (P == 0)
            asm volatile(GI_FMA12 GI_OPS);
        else if constexpr (P == 1)
            asm volatile(GI_FMA12 GI_MUL GI_OPS);
        else if constexpr (P == 2)
            asm volatile(GI_RSQ "s_nop 0\n\t" GI_MUL GI_FMA12 GI_OPS);
        else if constexpr (P == 3)
            asm volatile(GI_RSQ GI_FMA3(0, 1, 2) GI_MUL GI_FMA3(3, 4, 5) GI_FMA3(6, 7, 8) GI_FMA3(9, 10, 11) GI_OPS);
        else
            asm volatile(GI_RSQ GI_FMA12 GI_MUL GI_OPS);
#undef GI_OPS
    }
    out[tid] = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7 + a8 + a9 + a10 + a11 + t + u + m;
}

using Launch = void (*)(float*, const float*, int, int);
template <int OP>
static void launch_op(float* d, const float* in, int grid, int iters) {
    hipLaunchKernelGGL((k_op<OP>), dim3(grid), dim3(256), 0, 0, d, in, iters);
}
template <int P>
static void launch_place(float* d, const float* in, int grid, int iters) {
    hipLaunchKernelGGL((k_rsq_place<P>), dim3(grid), dim3(256), 0, 0, d, in, iters);
}
template <int R>
static void launch_pairs(float* d, const float* in, int grid, int iters) {
    hipLaunchKernelGGL((k_pairmix<R>), dim3(grid), dim3(256), 0, 0, d, in, iters);
}

static double time_ms(Launch f, float* d, const float* in, int grid, int iters) {
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    f(d, in, grid, 2);  // warm-up
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0, 0));
    f(d, in, grid, iters);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
    return ms;
}

int main() {
    int dev = 0, cus = 0;
    CHECK(hipGetDevice(&dev));
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const double clk = 2.4e9;
    const int simds = cus * 4;
    float *d = nullptr, *in = nullptr;
    CHECK(hipMalloc((void**)&d, (size_t)cus * 8 * 256 * sizeof(float)));  // the largest grid's threads
    CHECK(hipMalloc((void**)&in, 1024 * sizeof(float)));
    float h[1024];
    for (int k = 0; k < 1024; ++k) h[k] = 1.0f + 0.001f * k;
    CHECK(hipMemcpy(in, h, sizeof h, hipMemcpyHostToDevice));
    const char* names[5] = {"fma", "mul", "add", "rsq", "dpp_ror1"};
    const Launch ops[5] = {launch_op<0>, launch_op<1>, launch_op<2>, launch_op<3>, launch_op<4>};
    const int iters = 20000;
    for (int wps : {4, 8}) {  // wavefronts per SIMD (256-thread workgroups: 4 wavefronts)
        const int grid = cus * wps;
        for (int o = 0; o < 5; ++o) {
            const double ms = time_ms(ops[o], d, in, grid, iters);
            const double wi = (double)grid * 4 * iters * kChains;  // wave-level instructions in the loops
            const double per_simd_per_s = wi / simds / (ms * 1e-3);
            std::printf("%-8s waves/SIMD %d: %8.3f ms  %.3f wave-instr per SIMD per 2.4 GHz cycle = %.2f cycles each\n",
                        names[o], wps, ms, per_simd_per_s / clk, clk / per_simd_per_s);
        }
        for (int R : {4, 8}) {
            const double ms = time_ms(R == 4 ? launch_pairs<4> : launch_pairs<8>, d, in, grid, iters);
            const double pairs = (double)grid * 256 * iters * R;
            std::printf("pairmix R=%d waves/SIMD %d: %8.3f ms  %.4e evaluated pairs/s\n", R, wps, ms, pairs / (ms * 1e-3));
        }
    }
    // v_rsq_f32 placement: a warm-up of ~0.3 s first, then the best of five
    // runs of each pattern; cost of the v_rsq_f32 and its multiply in each
    // placement = its time minus that of the 12 FMAs alone.
    const char* pnames[5] = {"12 fma", "12 fma + mul", "rsq, wait, mul, 12 fma", "rsq, 3 fma, mul, 9 fma",
                             "rsq, 12 fma, mul"};
    const Launch places[5] = {launch_place<0>, launch_place<1>, launch_place<2>, launch_place<3>, launch_place<4>};
    const int piters = 40000;
    for (int k = 0; k < 40; ++k) places[0](d, in, cus * 8, piters);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    for (int wps : {5, 6, 8}) {
        const int grid = cus * wps;
        double base = 0.0;
        for (int p = 0; p < 5; ++p) {
            double best = 1e30;
            for (int rep = 0; rep < 5; ++rep) best = std::min(best, time_ms(places[p], d, in, grid, piters));
            // cycles per iteration of one wavefront, with wps of them sharing a SIMD
            const double cyc = best * 1e-3 * clk / ((double)wps * piters);
            if (p == 0) base = cyc;
            std::printf("placement waves/SIMD %d  %-24s %7.3f ms  %6.2f cycles per iteration  (+%5.2f)\n", wps,
                        pnames[p], best, cyc, cyc - base);
        }
    }
    CHECK(hipFree(d));
    CHECK(hipFree(in));
    std::printf("%d CUs\n", cus);
    return 0;
}
