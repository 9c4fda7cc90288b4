// Issue-rate calibration of the vector ALU: cycles per wave64 instruction per
// SIMD for the instruction classes of the pair kernels, alone and interleaved
// with FP32 FMA.  Eight waves per SIMD run eight independent chains each; every
// wave reads the shader clock (clock64) and the 100 MHz wall clock at the
// start and end of its loop, and the rate is the number of wave instructions
// per SIMD over the kernel-wide window (first start to last end) in shader
// cycles.  The ceiling of a kernel that issues non-packed FP32 follows from
// the FMA figure.
//
//   valu_probe
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

namespace {

constexpr int kIter = 16384;
constexpr int kBlock = 256;      // 4 waves
constexpr int kWavesPerSimd = 8; // 8 blocks per CU

enum Kind { FMA_F32, MUL_F32, FMA_F64, EXP_F32, RSQ_F32, RSQ_F64, DPP, SUB_I32, CVT,
            FMA_DPP, FMA_EXP, FMA_SUB, FMA_CVT, NKIND };
const char* kNames[NKIND] = {"v_fma_f32", "v_mul_f32", "v_fma_f64", "v_exp_f32", "v_rsq_f32",
                             "v_rsq_f64", "v_mov_b32_dpp", "v_sub_u32", "v_cvt_f32_i32",
                             "fma_f32 + dpp", "fma_f32 + exp_f32", "fma_f32 + sub_u32",
                             "fma_f32 + cvt_f32_i32"};
// wave instructions per loop iteration (eight chains, sixteen for the mixes)
constexpr int ops(int k) { return k >= FMA_DPP ? 16 : 8; }

template <int K>
__global__ __launch_bounds__(kBlock) void probe_kernel(float* out, long long* stamps, float a, float b,
                                                       int ia)
{
    float x[8];
    int y[8];
    double d[8];
    for (int c = 0; c < 8; ++c) {
        x[c] = threadIdx.x * 1e-3f + c;
        y[c] = threadIdx.x + c;
        d[c] = threadIdx.x * 1e-3 + c;
    }
    __syncthreads();
    const long long c0 = clock64(), w0 = wall_clock64();
    for (int i = 0; i < kIter; ++i) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if constexpr (K == FMA_F32 || K >= FMA_DPP) x[c] = fmaf(x[c], a, b);
            if constexpr (K == MUL_F32) x[c] = x[c] * a;
            if constexpr (K == FMA_F64) d[c] = fma(d[c], (double)a, (double)b);
            if constexpr (K == EXP_F32) x[c] = __builtin_amdgcn_exp2f(x[c]);
            if constexpr (K == RSQ_F32) x[c] = __builtin_amdgcn_rsqf(x[c]);
            if constexpr (K == RSQ_F64) d[c] = __builtin_amdgcn_rsq(d[c]);
            if constexpr (K == DPP || K == FMA_DPP) y[c] = __builtin_amdgcn_mov_dpp(y[c], 0x13C, 0xF, 0xF, false);
            if constexpr (K == FMA_EXP) y[c] = __float_as_int(__builtin_amdgcn_exp2f(__int_as_float(y[c])));
            if constexpr (K == SUB_I32 || K == FMA_SUB) y[c] = y[c] - ia;
            if constexpr (K == CVT || K == FMA_CVT) y[c] = __float_as_int((float)y[c]);
        }
        // keep every chain in registers and opaque to the optimiser
        asm volatile("" : "+v"(x[0]), "+v"(x[1]), "+v"(x[2]), "+v"(x[3]), "+v"(x[4]), "+v"(x[5]),
                     "+v"(x[6]), "+v"(x[7]));
        asm volatile("" : "+v"(y[0]), "+v"(y[1]), "+v"(y[2]), "+v"(y[3]), "+v"(y[4]), "+v"(y[5]),
                     "+v"(y[6]), "+v"(y[7]));
        asm volatile("" : "+v"(d[0]), "+v"(d[1]), "+v"(d[2]), "+v"(d[3]), "+v"(d[4]), "+v"(d[5]),
                     "+v"(d[6]), "+v"(d[7]));
    }
    const long long c1 = clock64(), w1 = wall_clock64();
    if (threadIdx.x % 64 == 0) {
        const size_t w = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) / 64;
        stamps[4 * w] = c0;
        stamps[4 * w + 1] = w0;
        stamps[4 * w + 2] = c1;
        stamps[4 * w + 3] = w1;
    }
    float s = 0.f;
    for (int c = 0; c < 8; ++c) s += x[c] + (float)y[c] + (float)d[c];
    out[blockIdx.x * (size_t)blockDim.x + threadIdx.x] = s;
}

template <int K>
bool run(int cus, float* out, long long* stamps, std::vector<long long>& h)
{
    const int blocks = cus * kWavesPerSimd;
    for (int rep = 0; rep < 3; ++rep) {  // the last launch is read: clocks are up by then
        hipLaunchKernelGGL(probe_kernel<K>, dim3(blocks), dim3(kBlock), 0, 0, out, stamps, 0.999f, 1e-3f, 7);
        if (hipDeviceSynchronize() != hipSuccess) return false;
    }
    const size_t waves = (size_t)blocks * (kBlock / 64);
    if (hipMemcpy(h.data(), stamps, waves * 4 * sizeof(long long), hipMemcpyDeviceToHost) != hipSuccess)
        return false;
    long long wmin = h[1], wmax = h[3];
    double ratio = 0;  // shader cycles per wall tick, averaged over waves
    for (size_t w = 0; w < waves; ++w) {
        wmin = h[4 * w + 1] < wmin ? h[4 * w + 1] : wmin;
        wmax = h[4 * w + 3] > wmax ? h[4 * w + 3] : wmax;
        ratio += (double)(h[4 * w + 2] - h[4 * w]) / (double)(h[4 * w + 3] - h[4 * w + 1]);
    }
    ratio /= (double)waves;
    const double window = (double)(wmax - wmin) * ratio;
    const double per_simd = (double)waves * kIter * ops(K) / (cus * 4.0);
    std::printf("| %-22s | %5.2f | %4.0f |\n", kNames[K], window / per_simd, ratio * 100.0);
    return true;
}

template <int... K>
bool run_all(int cus, float* out, long long* stamps, std::vector<long long>& h,
             std::integer_sequence<int, K...>)
{
    return (run<K>(cus, out, stamps, h) && ...);
}

}  // namespace

int main()
{
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess || cus < 1) {
        std::fprintf(stderr, "no device\n");
        return 1;
    }
    const size_t threads = (size_t)cus * kWavesPerSimd * kBlock;
    const size_t waves = threads / 64;
    float* out = nullptr;
    long long* stamps = nullptr;
    if (hipMalloc(&out, threads * sizeof(float)) != hipSuccess ||
        hipMalloc(&stamps, waves * 4 * sizeof(long long)) != hipSuccess) {
        std::fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    std::vector<long long> h(waves * 4);
    std::printf("%d CUs, %d waves per SIMD, %d iterations of 8 independent chains per wave\n\n", cus,
                kWavesPerSimd, kIter);
    std::printf("| instructions | cycles per wave instruction per SIMD | shader clock MHz |\n|---|---|---|\n");
    const bool ok = run_all(cus, out, stamps, h, std::make_integer_sequence<int, NKIND>{});
    (void)hipFree(out);
    (void)hipFree(stamps);
    return ok ? 0 : 1;
}
