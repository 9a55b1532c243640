/* peak.hip - MI355X peaks the tracer's rooflines are priced against, measured on the machine:
   vector FP64 FMA rate, vector FP32 FMA rate as this code is compiled (packed FP32 instructions
   disabled), HBM copy bandwidth, and the shader clock while the FMA kernel runs (shader cycle
   counter against the constant 100 MHz clock). Then the issue cost of single instructions for ONE
   wave alone on the chip (what bounds a frame's last, longest ray): shader cycles per instruction
   for 8 independent chains (the issue interval) and for one dependent chain (the latency), for the
   FP64 FMA, the FP64 reciprocal, the FP32 exp2 and the FP32 FMA. Every loop is bounded and every
   access in range.

   usage: bhrt_peak [iterations]        (prints one line per figure) */
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                             \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                  \
            std::exit(1);                                                                 \
        }                                                                                 \
    } while (0)

constexpr int CHAINS = 8;   // independent FMA chains per lane (covers the FMA latency)

template <typename T>
__global__ __launch_bounds__(256) void k_fma(const T* __restrict__ in, T* __restrict__ out, int iters,
                                             unsigned long long* __restrict__ clk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long c0 = clock64(), t0 = wall_clock64();
    T a[CHAINS];
    const T b = in[0], c = in[1];
    for (int k = 0; k < CHAINS; ++k) a[k] = in[2 + k] + (T)i;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < CHAINS; ++k) a[k] = fma(a[k], b, c);
    }
    T s = 0;
    for (int k = 0; k < CHAINS; ++k) s += a[k];
    out[i] = s;
    if (i == 0) {
        clk[0] = clock64() - c0;
        clk[1] = wall_clock64() - t0;
    }
}

__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

template <typename T>
static void fma_rate(const char* name, int cus, int iters) {
    const int blocks = cus * 32, threads = blocks * 256;   // 32 workgroups of 4 waves per CU
    T h[2 + CHAINS] = {(T)0.999999, (T)1e-7};
    for (int k = 0; k < CHAINS; ++k) h[2 + k] = (T)(k + 1);
    T *din, *dout;
    unsigned long long* dclk;
    CK(hipMalloc(&din, sizeof h));
    CK(hipMalloc(&dout, sizeof(T) * threads));
    CK(hipMalloc(&dclk, 2 * sizeof(unsigned long long)));
    CK(hipMemcpy(din, h, sizeof h, hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_fma<T>, dim3(blocks), dim3(256), 0, 0, din, dout, iters, dclk);   // warm-up
    CK(hipGetLastError());
    CK(hipEventRecord(e0));
    hipLaunchKernelGGL(k_fma<T>, dim3(blocks), dim3(256), 0, 0, din, dout, iters, dclk);
    CK(hipGetLastError());
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    unsigned long long clk[2];
    CK(hipMemcpy(clk, dclk, sizeof clk, hipMemcpyDeviceToHost));
    const double flops = 2.0 * CHAINS * (double)iters * threads;
    std::printf("%s FMA: %.1f TFLOP/s (%d workgroups x 256 lanes x %d FMA chains x %d iterations, %.3f ms); "
                "shader clock during the run %.0f MHz\n",
                name, flops / (ms * 1e-3) * 1e-12, blocks, CHAINS, iters, ms, clk[0] / (clk[1] * 0.01));
    CK(hipFree(din));
    CK(hipFree(dout));
    CK(hipFree(dclk));
    CK(hipEventDestroy(e0));
    CK(hipEventDestroy(e1));
}

// One wave: CHAINS instructions OP per iteration spread over NCH independent chains (CHAINS: the
// issue interval; 1: the dependent latency); shader cycles of the loop in clk[0]
enum Op { OP_FMA64, OP_RCP64, OP_EXP32, OP_FMA32 };

template <Op OP>
__device__ __forceinline__ void op_step(double* a, float* f, int k) {
    if constexpr (OP == OP_FMA64) a[k] = __builtin_fma(a[k], 0.999999, 1e-7);
    if constexpr (OP == OP_RCP64) a[k] = __builtin_amdgcn_rcp(a[k]);             // x, 1/x, x, ...
    if constexpr (OP == OP_EXP32) f[k] = __builtin_amdgcn_exp2f(-f[k]);          // stays in (0, 1]
    if constexpr (OP == OP_FMA32) f[k] = __builtin_fmaf(f[k], 0.999999f, 1e-7f);
}

template <Op OP, int NCH>
__global__ __launch_bounds__(64) void k_issue(const double* __restrict__ in, double* __restrict__ out, int iters,
                                              unsigned long long* __restrict__ clk) {
    const int i = threadIdx.x;
    double a[NCH];
    float f[NCH];
    for (int k = 0; k < NCH; ++k) {
        a[k] = in[k] + 1e-3 * i;
        f[k] = (float)a[k];
    }
    const unsigned long long c0 = clock64();
    for (int it = 0; it < iters; ++it) {   // CHAINS instructions per iteration either way
#pragma unroll
        for (int k = 0; k < CHAINS; ++k) op_step<OP>(a, f, k % NCH);
    }
    const unsigned long long c1 = clock64();
    double s = 0;
    for (int k = 0; k < NCH; ++k) s += a[k] + (double)f[k];
    out[i] = s;
    if (i == 0) clk[0] = c1 - c0;
}

template <Op OP, int NCH>
static double issue_cycles(int iters) {
    double h[CHAINS];
    for (int k = 0; k < CHAINS; ++k) h[k] = 0.5 + 0.0625 * k;
    double *din, *dout;
    unsigned long long* dclk;
    CK(hipMalloc(&din, sizeof h));
    CK(hipMalloc(&dout, 64 * sizeof(double)));
    CK(hipMalloc(&dclk, sizeof(unsigned long long)));
    CK(hipMemcpy(din, h, sizeof h, hipMemcpyHostToDevice));
    for (int r = 0; r < 2; ++r) {   // the second launch is the one read
        hipLaunchKernelGGL((k_issue<OP, NCH>), dim3(1), dim3(64), 0, 0, din, dout, iters, dclk);
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
    }
    unsigned long long clk = 0;
    CK(hipMemcpy(&clk, dclk, sizeof clk, hipMemcpyDeviceToHost));
    CK(hipFree(din));
    CK(hipFree(dout));
    CK(hipFree(dclk));
    return (double)clk / ((double)iters * CHAINS);
}

template <Op OP>
static void issue_line(const char* name, int iters) {
    std::printf("one wave, %s: %.2f cycles per instruction over %d independent chains, %.2f in one dependent "
                "chain\n",
                name, issue_cycles<OP, CHAINS>(iters), CHAINS, issue_cycles<OP, 1>(iters));
}

static void copy_rate(int cus) {
    const size_t bytes = (size_t)4 << 30, n = bytes / sizeof(float4);
    float4 *src, *dst;
    CK(hipMalloc(&src, bytes));
    CK(hipMalloc(&dst, bytes));
    CK(hipMemset(src, 0, bytes));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const int blocks = cus * 16;
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, src, dst, n);
    CK(hipGetLastError());
    CK(hipEventRecord(e0));
    for (int r = 0; r < 5; ++r) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, src, dst, n);
    CK(hipGetLastError());
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("HBM copy: %.0f GB/s (4 GiB read + 4 GiB written per pass, 5 passes, %.2f ms per pass)\n",
                2.0 * bytes * 5 / (ms * 1e-3) * 1e-9, ms / 5);
    CK(hipFree(src));
    CK(hipFree(dst));
    CK(hipEventDestroy(e0));
    CK(hipEventDestroy(e1));
}

int main(int argc, char** argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 20000;
    if (iters < 1 || iters > 10000000) {
        std::fprintf(stderr, "iterations must be in [1, 1e7]\n");
        return 2;
    }
    hipDeviceProp_t p;
    CK(hipGetDeviceProperties(&p, 0));
    std::printf("device %s: %d CUs, peak engine clock %d MHz\n", p.gcnArchName, p.multiProcessorCount,
                p.clockRate / 1000);
    fma_rate<double>("FP64", p.multiProcessorCount, iters);
    fma_rate<float>("FP32 (no packed FP32)", p.multiProcessorCount, iters);
    copy_rate(p.multiProcessorCount);
    issue_line<OP_FMA64>("v_fma_f64", iters);
    issue_line<OP_RCP64>("v_rcp_f64", iters);
    issue_line<OP_EXP32>("v_exp_f32", iters);
    issue_line<OP_FMA32>("v_fma_f32", iters);
    return 0;
}
