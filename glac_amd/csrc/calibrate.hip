// calibrate.hip -- issue-rate and L1 bandwidth calibration of the GPU, the
// ceilings the kernels' counters are read against (bench.py --calibration).
//
//  * FP32 / FP64 FMA issue: a wave64 FMA chain with 16 independent
//    accumulators per lane (no packed FP32: the build disables it), enough
//    waves to fill every SIMD; reported as cycles per wave64 instruction
//    per SIMD at the device's peak clock.
//  * L1 -> register bandwidth: every wavefront streams 16-byte loads over a
//    16 KiB window that stays in the CU's L1.
//
// Prints one JSON object.  Every index stays inside its buffer: the FMA
// kernels write blocks x 256 results, the L1 kernel reads 4096 float4.
#include <hip/hip_runtime.h>

#include <cstdio>

#define CHECK(x)                                                                                        \
    do {                                                                                                \
        hipError_t e_ = (x);                                                                            \
        if (e_ != hipSuccess) {                                                                         \
            std::fprintf(stderr, "calibrate: %s at line %d\n", hipGetErrorString(e_), __LINE__);        \
            return 1;                                                                                   \
        }                                                                                               \
    } while (0)

namespace {

constexpr int kThreads = 256;
constexpr int kAcc = 16;

template <class T> __global__ __launch_bounds__(kThreads) void fma_chain(T *out, int iters, T a, T b)
{
    T acc[kAcc];
#pragma unroll
    for (int i = 0; i < kAcc; i++) acc[i] = (T)(threadIdx.x + i);
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < kAcc; i++) acc[i] = acc[i] * a + b;
    }
    T s = 0;
#pragma unroll
    for (int i = 0; i < kAcc; i++) s += acc[i];
    out[blockIdx.x * kThreads + threadIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void l1_stream(const float4 *__restrict__ in, float *out, int iters)
{
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    const int base = (blockIdx.x & 3) * 1024;  // one of four 16 KiB windows
    int off = threadIdx.x;
    for (int it = 0; it < iters; it++) {
        const float4 v = in[base + (off & 1023)];
        s.x += v.x;
        s.y += v.y;
        s.z += v.z;
        s.w += v.w;
        off += kThreads;
    }
    out[blockIdx.x * kThreads + threadIdx.x] = s.x + s.y + s.z + s.w;
}

template <class T> int fma_rate(int blocks, int iters, double clock_hz, int cus, double *cycles, double *tflops)
{
    T *out = nullptr;
    CHECK(hipMalloc(&out, sizeof(T) * blocks * kThreads));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    fma_chain<T><<<blocks, kThreads>>>(out, 16, (T)0.999, (T)0.001);
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0));
    fma_chain<T><<<blocks, kThreads>>>(out, iters, (T)0.999, (T)0.001);
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    const double wave_instr = (double)blocks * (kThreads / 64) * iters * kAcc;
    const double simds = 4.0 * cus;
    *cycles = simds * clock_hz * (ms * 1e-3) / wave_instr;
    *tflops = wave_instr * 64 * 2 / (ms * 1e-3) / 1e12;
    CHECK(hipFree(out));
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
    return 0;
}

int l1_rate(int blocks, int iters, double clock_hz, int cus, double *bytes_per_cu_cycle, double *tbs)
{
    float4 *in = nullptr;
    float *out = nullptr;
    CHECK(hipMalloc(&in, 4096 * sizeof(float4)));
    CHECK(hipMemset(in, 0, 4096 * sizeof(float4)));
    CHECK(hipMalloc(&out, sizeof(float) * blocks * kThreads));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    l1_stream<<<blocks, kThreads>>>(in, out, 16);
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0));
    l1_stream<<<blocks, kThreads>>>(in, out, iters);
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    const double bytes = (double)blocks * kThreads * iters * 16;
    *tbs = bytes / (ms * 1e-3) / 1e12;
    *bytes_per_cu_cycle = bytes / (ms * 1e-3) / cus / clock_hz;
    CHECK(hipFree(in));
    CHECK(hipFree(out));
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
    return 0;
}

}  // namespace

int main()
{
    int dev = 0;
    CHECK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    CHECK(hipGetDeviceProperties(&p, dev));
    const double clock_hz = p.clockRate * 1e3;
    const int cus = p.multiProcessorCount;
    const int blocks = cus * 32;  // 8 wavefronts per SIMD
    double c32, t32, c64, t64, l1b, l1t;
    if (fma_rate<float>(blocks, 20000, clock_hz, cus, &c32, &t32)) return 1;
    if (fma_rate<double>(blocks, 10000, clock_hz, cus, &c64, &t64)) return 1;
    if (l1_rate(blocks, 4000, clock_hz, cus, &l1b, &l1t)) return 1;
    std::printf("{\"device\": \"%s\", \"cus\": %d, \"peak_clock_mhz\": %.0f, "
                "\"fp32_fma_cycles_per_wave_instr\": %.2f, \"fp32_fma_tflops\": %.1f, "
                "\"fp64_fma_cycles_per_wave_instr\": %.2f, \"fp64_fma_tflops\": %.1f, "
                "\"l1_float4_bytes_per_cu_cycle\": %.1f, \"l1_float4_tbs\": %.1f}\n",
                p.gcnArchName, cus, clock_hz / 1e6, c32, t32, c64, t64, l1b, l1t);
    return 0;
}
