// Memory yardsticks for the FFT passes.
//
// A plain device-to-device copy of a known size is the yardstick against
// which rocprofv3's FETCH_SIZE / WRITE_SIZE readings of the FFT kernels are
// scaled (a copy of B bytes must fetch B and write B).  The other modes time
// the access patterns of the axis passes with no arithmetic, so a pass can
// be set beside the copy of its own pattern:
//   copy     grid-stride 16-byte copy, out of place
//   inplace  grid-stride 16-byte read-modify-write, in place
//   memcpy   hipMemcpyAsync device to device (the runtime's copy engine path)
//   rows     in place, one block of 256 threads per 4096-element row, 16
//            elements per thread (the 2D 4096^2 X pass's pattern)
//   cols     in place, one block of 1024 threads per tile of 4 adjacent
//            columns of 4096 rows, lanes walk the 4 columns first, blocks
//            laid out per XCD as the axis passes lay them out (the Y pass's
//            pattern)
//   all      every mode above
// `work` adds that many rounds of two fused multiply-adds per element
// between the loads and the stores (a stand-in for a pass's arithmetic;
// values are unchanged).  rows and cols need MiB = 256 (a 4096^2 complex128
// array).
//   usage: copy_calib [MiB=256] [reps=10] [mode=copy] [work=0]
// (mode 0 / 1 are accepted for copy / inplace.)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

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

constexpr int SIDE = 4096;

__device__ __forceinline__ void busy(double2 (&v)[16], int work, double f, double g)
{
    for (int w = 0; w < work; ++w)
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            v[e].x = __builtin_fma(v[e].x, f, g);
            v[e].y = __builtin_fma(v[e].y, f, g);
        }
}

// one row of SIDE elements per block: thread p holds p + 256 e
__global__ void __launch_bounds__(256) rows_copy(double2 *a, int work, double f, double g)
{
    double2 *row = a + (long long)blockIdx.x * SIDE;
    const int p = threadIdx.x;
    double2 v[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = row[p + 256 * e];
    busy(v, work, f, g);
#pragma unroll
    for (int e = 0; e < 16; ++e) row[p + 256 * e] = v[e];
}

// a tile of 4 adjacent columns per block: lane (t = tid % 4, p = tid / 4)
// holds rows p + 256 e of column t; consecutive tiles on one XCD
__global__ void __launch_bounds__(1024) cols_copy(double2 *a, int work, double f, double g)
{
    const long long b = blockIdx.x, G = gridDim.x;
    const long long q = G >> 3, r = G & 7, x = b & 7, k = b >> 3;
    const long long tile = (G & 255) == 0 ? ((k >> 5) << 8) + (x << 5) + (k & 31)
                                          : (x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k);
    const int t = threadIdx.x & 3, p = threadIdx.x >> 2;
    double2 *col = a + tile * 4 + t;
    double2 v[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = col[(long long)(p + 256 * e) * SIDE];
    busy(v, work, f, g);
#pragma unroll
    for (int e = 0; e < 16; ++e) col[(long long)(p + 256 * e) * SIDE] = v[e];
}

int main(int argc, char **argv)
{
    const long long mib = argc > 1 ? atoll(argv[1]) : 256;
    const int reps = argc > 2 ? atoi(argv[2]) : 10;
    const char *mode = argc > 3 ? argv[3] : "copy";
    const int work = argc > 4 ? atoi(argv[4]) : 0;
    if (!std::strcmp(mode, "0")) mode = "copy";
    if (!std::strcmp(mode, "1")) mode = "inplace";
    const long long bytes = mib << 20, n = bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
    if (hipMemset(a, 0, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    const char *modes[] = {"copy", "inplace", "memcpy", "rows", "cols"};
    const bool all = !std::strcmp(mode, "all");
    int rc = 0;
    for (const char *m : modes) {
        if (!all && std::strcmp(m, mode)) continue;
        const bool tiled = !std::strcmp(m, "rows") || !std::strcmp(m, "cols");
        if (tiled && bytes != (long long)SIDE * SIDE * 16) {
            std::fprintf(stderr, "%s needs MiB = 256\n", m);
            rc = 2;
            continue;
        }
        auto launch = [&] {
            if (!std::strcmp(m, "copy")) hipLaunchKernelGGL(copy16, dim3(256 * 32), dim3(256), 0, 0, a, b, n);
            else if (!std::strcmp(m, "inplace")) hipLaunchKernelGGL(touch16, dim3(256 * 32), dim3(256), 0, 0, a, n, 0u);
            else if (!std::strcmp(m, "memcpy")) (void)hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, 0);
            else if (!std::strcmp(m, "rows"))
                hipLaunchKernelGGL(rows_copy, dim3(SIDE), dim3(256), 0, 0, (double2 *)a, work, 1.0, 0.0);
            else hipLaunchKernelGGL(cols_copy, dim3(SIDE / 4), dim3(1024), 0, 0, (double2 *)a, work, 1.0, 0.0);
        };
        launch();
        if (hipEventRecord(e0, 0) != hipSuccess) return 1;
        for (int r = 0; r < reps; ++r) launch();
        if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
        if (hipGetLastError() != hipSuccess) return 1;
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
        const double t = ms * 1e-3 / reps;
        std::printf("%s %lld MiB work %d: %.2f us per pass, %.1f GB/s (read + write)\n", m, mib, tiled ? work : 0,
                    t * 1e6, 2.0 * bytes / t / 1e9);
    }
    return (hipFree(a) == hipSuccess && hipFree(b) == hipSuccess) ? rc : 1;
}
