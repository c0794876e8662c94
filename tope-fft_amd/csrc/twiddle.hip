// FP64 twiddle generation on the GPU and the naive DFT kernel.
//
// Replaces the reference's plan-time `twid1D` / `twiddles` kernels
// (src/kernels1D.cl:7-13, src/kernels2D.cl:7-13, src/kernels3D.cl:7-13),
// which evaluate cos/sin(2*pi*b/n) into a quarter-wave table, and the
// non-power-of-two `DFT` kernel (src/kernels1D.cl:549-573).
//
// Every root of unity W_M^e is computed from the integer pair (e mod M, M):
// sincospi(2*e/M) in FP64, whose argument is exact for power-of-two M, so
// no unreduced large argument ever reaches the sin/cos (the reference's DFT
// kernel passes up to 2*pi*N^2/N unreduced).
#include "internal.hpp"

namespace tope {

__device__ __forceinline__ void root64(long long e, long long M, double &c, double &s)
{
    e %= M;
    if (e < 0) e += M;
    double sv, cv;
    sincospi(2.0 * (double)e / (double)M, &sv, &cv);
    c = cv;
    s = -sv;  // forward sign: e^{-2 pi i e / M}
}

struct StageList {
    int nst;
    int radix[16];
    int span[16];
    int off[16];
};

template <typename R>
__global__ void gen_stage_tw_kernel(cpx<R> *out, StageList L, int total)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int s = 1;
    while (s + 1 < L.nst && i >= L.off[s + 1]) ++s;
    const int local = i - L.off[s];
    const int ns = L.span[s];
    const int m = local / ns + 1;
    const int k = local - (m - 1) * ns;
    double c, sn;
    root64((long long)k * m, (long long)ns * L.radix[s], c, sn);
    out[i] = cpx<R>{(R)c, (R)sn};
}

template <typename R>
__global__ void gen_power_kernel(cpx<R> *out, long long M, long long mult, long long count)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    double c, sn;
    root64((i % M) * (mult % M), M, c, sn);
    out[i] = cpx<R>{(R)c, (R)sn};
}

HostSched host_sched(int n, int e)
{
    HostSched h{};
    const int rx = e < 16 ? e : 16;
    int ln = 0, lr = 0;
    while ((1 << ln) < n) ++ln;
    while ((1 << lr) < rx) ++lr;
    const int nfull = lr ? ln / lr : 0;
    const int rem = 1 << (lr ? ln % lr : ln);
    h.nst = nfull + (rem > 1 ? 1 : 0);
    int span = 1, off = 0;
    for (int s = 0; s < h.nst; ++s) {
        h.radix[s] = (rem > 1 && s == 0) ? rem : rx;
        h.span[s] = span;
        h.off[s] = off;
        if (s > 0) off += (h.radix[s] - 1) * span;
        span *= h.radix[s];
    }
    h.twsize = off;
    return h;
}

hipError_t gen_stage_twiddles(void *out, int prec, int n, int e, hipStream_t s)
{
    HostSched h = host_sched(n, e);
    if (h.twsize == 0) return hipSuccess;
    StageList L{};
    L.nst = h.nst;
    for (int i = 0; i < h.nst; ++i) {
        L.radix[i] = h.radix[i];
        L.span[i] = h.span[i];
        L.off[i] = h.off[i];
    }
    const int bs = 256;
    const int grid = (h.twsize + bs - 1) / bs;
    if (prec == PREC_F64)
        hipLaunchKernelGGL(gen_stage_tw_kernel<double>, dim3(grid), dim3(bs), 0, s, (cpx<double> *)out, L,
                           h.twsize);
    else
        hipLaunchKernelGGL(gen_stage_tw_kernel<float>, dim3(grid), dim3(bs), 0, s, (cpx<float> *)out, L,
                           h.twsize);
    return hipGetLastError();
}

hipError_t gen_power_table(void *out, int prec, long long M, long long mult, long long count, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    const int bs = 256;
    const unsigned grid = (unsigned)((count + bs - 1) / bs);
    if (prec == PREC_F64)
        hipLaunchKernelGGL(gen_power_kernel<double>, dim3(grid), dim3(bs), 0, s, (cpx<double> *)out, M, mult,
                           count);
    else
        hipLaunchKernelGGL(gen_power_kernel<float>, dim3(grid), dim3(bs), 0, s, (cpx<float> *)out, M, mult,
                           count);
    return hipGetLastError();
}

// One thread per output bin; the phase index (j*k) mod n is stepped in
// integers, so every twiddle comes exactly from the table.
template <typename R, int SIGN>
__global__ void naive_dft_kernel(const cpx<R> *__restrict__ in, cpx<R> *__restrict__ out,
                                 const cpx<R> *__restrict__ roots, long long n, R scale_)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    R ar = 0, ai = 0;
    long long idx = 0;
    for (long long j = 0; j < n; ++j) {
        const cpx<R> x = in[j];
        const cpx<R> w = roots[idx];
        const cpx<R> p = twmul<SIGN>(x, w);
        ar += p.x;
        ai += p.y;
        idx += k;
        if (idx >= n) idx -= n;
    }
    out[k] = cpx<R>{ar * scale_, ai * scale_};
}

hipError_t launch_naive_dft(int prec, const void *in, void *out, const void *roots, long long n, int sign,
                            double scale_, hipStream_t s)
{
    const int bs = 64;
    const unsigned grid = (unsigned)((n + bs - 1) / bs);
    if (prec == PREC_F64) {
        auto i = (const cpx<double> *)in;
        auto o = (cpx<double> *)out;
        auto r = (const cpx<double> *)roots;
        if (sign < 0)
            hipLaunchKernelGGL((naive_dft_kernel<double, -1>), dim3(grid), dim3(bs), 0, s, i, o, r, n, scale_);
        else
            hipLaunchKernelGGL((naive_dft_kernel<double, 1>), dim3(grid), dim3(bs), 0, s, i, o, r, n, scale_);
    } else {
        auto i = (const cpx<float> *)in;
        auto o = (cpx<float> *)out;
        auto r = (const cpx<float> *)roots;
        if (sign < 0)
            hipLaunchKernelGGL((naive_dft_kernel<float, -1>), dim3(grid), dim3(bs), 0, s, i, o, r, n,
                               (float)scale_);
        else
            hipLaunchKernelGGL((naive_dft_kernel<float, 1>), dim3(grid), dim3(bs), 0, s, i, o, r, n,
                               (float)scale_);
    }
    return hipGetLastError();
}

}  // namespace tope
