// Plain device-to-device copy of a known size: the yardstick against which
// rocprofv3's FETCH_SIZE / WRITE_SIZE readings of the FFT kernels are scaled
// (a copy of B bytes must fetch B and write B).
// With a third argument 1 the copy runs in place (each 16-byte word read and
// written back), the footprint of the in-place FFT passes.
//   usage: copy_calib [MiB=256] [reps=10] [inplace=0]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

__global__ void touch16(uint4 *p, long long n, unsigned k)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        uint4 v = p[i];
        v.x ^= k;
        p[i] = v;
    }
}

int main(int argc, char **argv)
{
    const long long mib = argc > 1 ? atoll(argv[1]) : 256;
    const int reps = argc > 2 ? atoi(argv[2]) : 10;
    const bool inplace = argc > 3 && atoi(argv[3]) != 0;
    const long long bytes = mib << 20, n = bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
    if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    const int grid = 256 * 32, block = 256;
    auto launch = [&] {
        if (inplace) hipLaunchKernelGGL(touch16, dim3(grid), dim3(block), 0, 0, a, n, 0u);
        else hipLaunchKernelGGL(copy16, dim3(grid), dim3(block), 0, 0, a, b, n);
    };
    launch();
    if (hipEventRecord(e0, 0) != hipSuccess) return 1;
    for (int r = 0; r < reps; ++r) launch();
    if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
    const double t = ms * 1e-3 / reps;
    std::printf("%s %lld MiB: %.2f us per copy, %.1f GB/s (read + write)\n", inplace ? "in-place copy" : "copy", mib,
                t * 1e6, 2.0 * bytes / t / 1e9);
    return (hipFree(a) == hipSuccess && hipFree(b) == hipSuccess) ? 0 : 1;
}
