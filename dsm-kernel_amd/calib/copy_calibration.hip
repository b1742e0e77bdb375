// copy_calibration.hip -- a plain device copy of a known size (1 GiB read,
// 1 GiB written), run under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to
// scale those counters to bytes for the kernels of this library.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void copy_calibration_kernel(const double4 *__restrict__ s, double4 *__restrict__ d, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}

int main()
{
    const size_t bytes = (size_t)1 << 30, n = bytes / sizeof(double4);
    double4 *s = nullptr, *d = nullptr;
    if (hipMalloc(&s, bytes) != hipSuccess || hipMalloc(&d, bytes) != hipSuccess) return 1;
    if (hipMemset(s, 1, bytes) != hipSuccess) return 1;
    hipLaunchKernelGGL(copy_calibration_kernel, dim3(4096), dim3(256), 0, 0, s, d, n);
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    std::printf("copied %zu bytes (read %zu, written %zu)\n", bytes, bytes, bytes);
    return 0;
}
