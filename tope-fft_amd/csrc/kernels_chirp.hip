// Non-power-of-two axes by Bluestein's chirp-z identity on top of the
// power-of-two passes.  The reference runs 1D non-powers of two as an O(n^2)
// DFT (src/fft1d.c:148-232, src/kernels1D.cl:549-573) and rejects them in
// 2D/3D (src/fft2d.c:355-362, src/fft3d.c:522-533).
//
// With c[k] = W_2n^(k^2) (k^2 reduced mod 2n in integers, so the phase is
// exact) and jk = (j^2 + k^2 - (j-k)^2) / 2:
//   X[j] = c[j] sum_k (x[k] c[k]) conj(c[j-k]),
// a linear convolution of a[k] = x[k] c[k] with b[m] = conj(c[m]),
// |m| < n, evaluated as a circular one of power-of-two length M >= 2n - 1:
// A = FFT_M(a zero-padded), C = IFFT_M(A * FFT_M(b wrapped)), X[j] = c[j] C[j].
// The inverse transform uses conj(c) throughout.  Every line of an axis is
// padded into its own row of M elements of a scratch array, so the two inner
// FFTs are contiguous row passes.
#include "internal.hpp"

namespace tope {

__device__ __forceinline__ cpx<double> chirp(long long k, long long n, int inverse)
{
    // W_2n^(k^2): k^2 mod 2n in 64-bit integers (k < n <= 2^31)
    const unsigned long long m2 = 2ULL * (unsigned long long)n;
    const unsigned long long e = ((unsigned long long)k * (unsigned long long)k) % m2;
    double s, c;
    sincospi((double)e / (double)n, &s, &c);  // angle 2 pi e / 2n = pi e / n
    return cpx<double>{c, inverse ? s : -s};
}

// b (length M): b[m] = conj(c[m]) for m < n, b[M - m] = conj(c[m]) for
// 0 < m < n, zero elsewhere (c conjugated for the inverse's filter).
__global__ void chirp_filter_kernel(cpx<double> *b, long long n, long long M, int inverse)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    long long m = -1;
    if (i < n) m = i;
    else if (i > M - n) m = M - i;
    cpx<double> v{0.0, 0.0};
    if (m >= 0) {
        const cpx<double> c = chirp(m, n, inverse);
        v = cpx<double>{c.x, -c.y};
    }
    b[i] = v;
}

// Element k of line l (o = l / C, c = l % C) of the axis: o os + c cs + k is.
// Threads run along k when the axis is contiguous (is = 1), else along the
// lines (adjacent lines are adjacent in memory for the strided axes).
__device__ __forceinline__ void line_of(long long idx, const ChirpGeo &g, long long span, long long &l, long long &k)
{
    if (g.is == 1) {
        l = idx / span;
        k = idx % span;
    } else {
        const long long lines = g.O * g.C;
        k = idx / lines;
        l = idx % lines;
    }
}

__device__ __forceinline__ long long elem_of(long long l, long long k, const ChirpGeo &g)
{
    return (l / g.C) * g.os + (l % g.C) * g.cs + k * g.is;
}

// S[l M + k] = x_l[k] c[k] (k < n), 0 (n <= k < M)
__global__ void chirp_pre_kernel(const cpx<double> *__restrict__ x, cpx<double> *__restrict__ S, const ChirpGeo g,
                                 long long n, long long M, int inverse)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.O * g.C * M) return;
    long long l, k;
    line_of(idx, g, M, l, k);
    cpx<double> v{0.0, 0.0};
    if (k < n) v = cmul(x[elem_of(l, k, g)], chirp(k, n, inverse));
    S[l * M + k] = v;
}

// S[l M + k] *= B[k] (the transformed filter of the direction)
__global__ void chirp_mul_kernel(cpx<double> *__restrict__ S, const cpx<double> *__restrict__ B, long long M,
                                 long long total)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) S[i] = cmul(S[i], B[i % M]);
}

// x_l[j] = c[j] S[l M + j] s (j < n)
__global__ void chirp_post_kernel(const cpx<double> *__restrict__ S, cpx<double> *__restrict__ x, const ChirpGeo g,
                                  long long n, long long M, double s, int inverse)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.O * g.C * n) return;
    long long l, j;
    line_of(idx, g, n, l, j);
    const cpx<double> v = cmul(S[l * M + j], chirp(j, n, inverse));
    x[elem_of(l, j, g)] = cpx<double>{v.x * s, v.y * s};
}

static unsigned grid_of(long long count) { return (unsigned)((count + 255) / 256); }

hipError_t launch_chirp(int op, const void *in, void *out, const void *table, const ChirpGeo &g, long long n,
                        long long M, double s, int inverse, hipStream_t st)
{
    const cpx<double> *i = static_cast<const cpx<double> *>(in);
    cpx<double> *o = static_cast<cpx<double> *>(out);
    const long long lines = g.O * g.C;
    if ((lines * M + 255) / 256 > 0x7fffffffLL) return hipErrorInvalidValue;
    switch (op) {
    case CHIRP_FILTER:
        hipLaunchKernelGGL(chirp_filter_kernel, dim3(grid_of(M)), dim3(256), 0, st, o, n, M, inverse);
        break;
    case CHIRP_PRE:
        hipLaunchKernelGGL(chirp_pre_kernel, dim3(grid_of(lines * M)), dim3(256), 0, st, i, o, g, n, M, inverse);
        break;
    case CHIRP_MUL:
        hipLaunchKernelGGL(chirp_mul_kernel, dim3(grid_of(lines * M)), dim3(256), 0, st, o,
                           static_cast<const cpx<double> *>(table), M, lines * M);
        break;
    case CHIRP_POST:
        hipLaunchKernelGGL(chirp_post_kernel, dim3(grid_of(lines * n)), dim3(256), 0, st, i, o, g, n, M, s, inverse);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace tope
