// Plain device copy of a known size, the calibration kernel for HBM traffic
// counters: FETCH_SIZE and WRITE_SIZE of one launch of copy_kernel are compared
// with its algorithmic bytes (n doubles read, n doubles written).
//
//   copy_probe [MiB] [launches]      default 512 MiB, 5 launches
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void copy_kernel(const double2* __restrict__ a,
                                                   double2* __restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

int main(int argc, char** argv)
{
    const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 512;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
    const size_t n = mib * 1024 * 1024 / sizeof(double2);
    double2 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * sizeof(double2)) != hipSuccess || hipMalloc(&b, n * sizeof(double2)) != hipSuccess) {
        std::fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    hipMemset(a, 0, n * sizeof(double2));
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    float best = 1e30f;
    for (int r = 0; r < reps; ++r) {
        hipEventRecord(e0);
        hipLaunchKernelGGL(copy_kernel, dim3(8192), dim3(256), 0, 0, a, b, n);
        hipEventRecord(e1);
        hipEventSynchronize(e1);
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) best = ms;
    }
    std::printf("copy_probe: %zu MiB read + %zu MiB written per launch, best %.3f ms, %.1f GB/s\n",
                mib, mib, best, 2.0 * mib * 1048576.0 / (best * 1e-3) / 1e9);
    hipFree(a);
    hipFree(b);
    return 0;
}
