// Non-power-of-two 1D transforms by Bluestein's chirp-z identity on top of
// the power-of-two passes.  The reference runs these lengths as an O(n^2)
// DFT (src/fft1d.c:148-232, src/kernels1D.cl:549-573); this replaces it
// above a small length, with the same result to FP64 rounding.
//
// With c[k] = W_2n^(k^2) (k^2 reduced mod 2n in integers, so the phase is
// exact) and jk = (j^2 + k^2 - (j-k)^2) / 2:
//   X[j] = c[j] sum_k (x[k] c[k]) conj(c[j-k]),
// a linear convolution of a[k] = x[k] c[k] with b[m] = conj(c[m]),
// |m| < n, evaluated as a circular one of power-of-two length M >= 2n - 1:
// A = FFT_M(a zero-padded), C = IFFT_M(A * FFT_M(b wrapped)), X[j] = c[j] C[j].
// The inverse transform uses conj(c) throughout and scales by 1/n.
#include "internal.hpp"

namespace tope {

__device__ __forceinline__ cpx<double> chirp(long long k, long long n)
{
    // W_2n^(k^2): k^2 mod 2n in 64-bit integers (k < n <= 2^31)
    const unsigned long long m2 = 2ULL * (unsigned long long)n;
    const unsigned long long e = ((unsigned long long)k * (unsigned long long)k) % m2;
    double s, c;
    sincospi((double)e / (double)n, &s, &c);  // angle 2 pi e / 2n = pi e / n
    return cpx<double>{c, -s};
}

// b (length M, forward transform's filter): b[m] = conj(c[m]) for m < n,
// b[M - m] = conj(c[m]) for 0 < m < n, zero elsewhere; conj_out selects the
// inverse transform's filter (c itself).
__global__ void chirp_filter_kernel(cpx<double> *b, long long n, long long M, int inverse)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    long long m = -1;
    if (i < n) m = i;
    else if (i > M - n) m = M - i;
    cpx<double> v{0.0, 0.0};
    if (m >= 0) {
        const cpx<double> c = chirp(m, n);
        v = inverse ? c : cpx<double>{c.x, -c.y};
    }
    b[i] = v;
}

// a[k] = x[k] c[k] (k < n), 0 (n <= k < M); c conjugated for the inverse.
__global__ void chirp_pre_kernel(const cpx<double> *__restrict__ x, cpx<double> *__restrict__ a, long long n,
                                 long long M, int inverse)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= M) return;
    cpx<double> v{0.0, 0.0};
    if (k < n) {
        cpx<double> c = chirp(k, n);
        if (inverse) c.y = -c.y;
        v = cmul(x[k], c);
    }
    a[k] = v;
}

// a[i] *= B[i] (the transformed filter of the direction)
__global__ void chirp_mul_kernel(cpx<double> *__restrict__ a, const cpx<double> *__restrict__ B, long long M)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) a[i] = cmul(a[i], B[i]);
}

// X[j] = c[j] C[j] s (j < n), s = 1/M (unnormalised inverse FFT) times 1/n
// for the inverse transform.
__global__ void chirp_post_kernel(const cpx<double> *__restrict__ C, cpx<double> *__restrict__ X, long long n,
                                  double s, int inverse)
{
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    cpx<double> c = chirp(j, n);
    if (inverse) c.y = -c.y;
    const cpx<double> v = cmul(C[j], c);
    X[j] = cpx<double>{v.x * s, v.y * s};
}

static unsigned grid_of(long long count) { return (unsigned)((count + 255) / 256); }

hipError_t launch_chirp(int op, const void *in, void *out, const void *table, long long n, long long M, double s,
                        int inverse, hipStream_t st)
{
    const cpx<double> *i = static_cast<const cpx<double> *>(in);
    cpx<double> *o = static_cast<cpx<double> *>(out);
    switch (op) {
    case CHIRP_FILTER:
        hipLaunchKernelGGL(chirp_filter_kernel, dim3(grid_of(M)), dim3(256), 0, st, o, n, M, inverse);
        break;
    case CHIRP_PRE: hipLaunchKernelGGL(chirp_pre_kernel, dim3(grid_of(M)), dim3(256), 0, st, i, o, n, M, inverse); break;
    case CHIRP_MUL:
        hipLaunchKernelGGL(chirp_mul_kernel, dim3(grid_of(M)), dim3(256), 0, st, o,
                           static_cast<const cpx<double> *>(table), M);
        break;
    case CHIRP_POST:
        hipLaunchKernelGGL(chirp_post_kernel, dim3(grid_of(n)), dim3(256), 0, st, i, o, n, s, inverse);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace tope
