// Fused axis-pass kernel: every radix stage of a length-N transform runs in
// one launch, data held in registers and exchanged through LDS, one HBM read
// and one HBM write per element.
//
// It replaces, per axis, the reference's sequence "swap pass + log_r(N)
// single-stage launches + divide pass" (src/fft2d.c:12-72 / :74-134 /
// :178-189, src/fft3d.c:12-214, src/fft1d.c:30-137), each of which was a
// full-array HBM round trip.
//
// Algorithm: self-sorting Stockham, decimation in time.  A stage of radix r
// at span Ns treats "virtual thread" j in [0, N/r): it reads elements
// j + m*N/r (m < r), multiplies element m by W_{Ns*r}^{(j mod Ns)*m}, does an
// r-point DFT and writes output m to (j/Ns)*Ns*r + (j mod Ns) + m*Ns.
// Starting at Ns = 1 the output is in natural order, so no bit-reversal pass
// (reference swap kernels, src/kernels2D.cl:546-592) is needed.
//
// Each physical thread p owns E elements p + P*e (P = N/E threads per
// transform) and runs E/r virtual threads per stage.  A block handles TB
// transforms.  Element i of transform c (in a block tile) lives at
//     base + o*os + c*cs + i*is        (complex units, 64-bit offsets)
// which covers rows (is = 1), strided columns (cs = 1, is = row pitch),
// the 3D Z axis, batches, and both steps of a four-step decomposition.
#pragma once
#include "cplx.hpp"

namespace tope {

constexpr int ilog2c(long long n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }

// Radix schedule for an N-point transform with E elements per thread:
// one leading stage of radix N / E^k (if N is not a power of E), then radix-E
// stages.  Shared by the host (twiddle table layout) and the device.
template <int N, int E>
struct Sched {
    static constexpr int LN = ilog2c(N), LE = ilog2c(E);
    static constexpr int NFULL = LE ? LN / LE : 0;
    static constexpr int REM = 1 << (LE ? LN % LE : LN);
    static constexpr int NST = NFULL + (REM > 1 ? 1 : 0);
    static constexpr int P = N / E;
    __host__ __device__ static constexpr int radix(int s) { return (REM > 1 && s == 0) ? REM : E; }
    __host__ __device__ static constexpr int span(int s)
    {
        int v = 1;
        for (int i = 0; i < s; ++i) v *= radix(i);
        return v;
    }
    // offset (in entries) of stage s's twiddle block; layout [(m-1)*Ns + k]
    __host__ __device__ static constexpr int twoff(int s)
    {
        int o = 0;
        for (int i = 1; i < s; ++i) o += (radix(i) - 1) * span(i);
        return o;
    }
    static constexpr int TWSIZE = twoff(NST);
};

// Kernel arguments (all strides in complex elements).
template <typename R>
struct AxisArgs {
    long long in_os, in_cs, in_is;
    long long out_os, out_cs, out_is;
    int C;      // transforms per outer index
    int tiles;  // ceil(C / TB)
    const cpx<R> *tw;  // stage twiddles for this N (forward sign)
    // optional four-step post-twiddle: out element k of transform (o, c) is
    // multiplied by W_{pt_n}^{(o*pt_o + c*pt_c) * k}; W_{pt_n}^a is
    // pt_lo[a & pt_mask] * pt_hi[a >> pt_shift]
    const cpx<R> *pt_lo, *pt_hi;
    long long pt_o, pt_c;
    long long pt_n;
    int pt_shift;
    R scale;    // multiplies the output (1/N on the last inverse pass, else 1)
};

// LDS index of element i of block-local transform t.  One pad slot per 16
// elements breaks the power-of-two strides of the Stockham scatter.
template <int N, int TB, bool CFAST>
__device__ __forceinline__ int lds_index(int t, int i)
{
    const int pi = i + (i >> 4);
    if constexpr (CFAST) return pi * TB + t;
    else return t * (N + N / 16) + pi;
}

template <int N, int TB>
constexpr int lds_slots() { return TB * (N + N / 16); }

// LDS exchange of the E register elements of every thread: write element
// e to slot dst(e), then read slots p + P*e back.  SPLIT exchanges the real
// and imaginary parts in two rounds, halving the LDS footprint (this is what
// lets a block hold four 4096-point FP64 columns).
template <typename R, int N, int E, int TB, bool CFAST, bool SPLIT>
struct Exchange {
    static constexpr int BYTES = lds_slots<N, TB>() * (SPLIT ? sizeof(R) : 2 * sizeof(R));
};
template <typename R, int N, int E, int SIGN, int S>
__device__ __forceinline__ void run_stage(cpx<R> (&v)[E], int p, const cpx<R> *__restrict__ tw)
{
    using SC = Sched<N, E>;
    constexpr int r = SC::radix(S);
    constexpr int Ns = SC::span(S);
    constexpr int P = SC::P;
    constexpr int Q = E / r;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        cpx<R> x[r];
#pragma unroll
        for (int m = 0; m < r; ++m) x[m] = v[q + m * Q];
        if constexpr (S > 0) {
            const int j = p + q * P;
            const int k = j & (Ns - 1);
            const cpx<R> *t = tw + SC::twoff(S) + k;
#pragma unroll
            for (int m = 1; m < r; ++m) x[m] = twmul<SIGN>(x[m], t[(m - 1) * Ns]);
        }
        dft<r, SIGN>(x);
#pragma unroll
        for (int m = 0; m < r; ++m) v[q + m * Q] = x[m];
    }
}

// destination of output m of virtual thread j in stage S
template <int N, int E, int S>
__device__ __forceinline__ int stage_dst(int j, int m)
{
    using SC = Sched<N, E>;
    constexpr int r = SC::radix(S);
    constexpr int Ns = SC::span(S);
    return (j / Ns) * Ns * r + (j & (Ns - 1)) + m * Ns;
}

template <typename R, int N, int E, int TB, bool CFAST, bool SPLIT, int SIGN, int S>
__device__ __forceinline__ void run_stages(cpx<R> (&v)[E], int p, int t, const cpx<R> *__restrict__ tw,
                                           void *lds_raw)
{
    using SC = Sched<N, E>;
    run_stage<R, N, E, SIGN, S>(v, p, tw);
    if constexpr (S + 1 < SC::NST) {
        constexpr int r = SC::radix(S);
        constexpr int Q = E / r;
        constexpr int P = SC::P;
        if constexpr (P == 1) {
            // whole transform in one thread: permute registers
            cpx<R> w[E];
#pragma unroll
            for (int q = 0; q < Q; ++q)
#pragma unroll
                for (int m = 0; m < r; ++m) w[stage_dst<N, E, S>(q, m)] = v[q + m * Q];
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = w[e];
        } else {
            if constexpr (!SPLIT) {
                cpx<R> *lds = (cpx<R> *)lds_raw;
                if constexpr (S > 0) __syncthreads();
#pragma unroll
                for (int q = 0; q < Q; ++q)
#pragma unroll
                    for (int m = 0; m < r; ++m)
                        lds[lds_index<N, TB, CFAST>(t, stage_dst<N, E, S>(p + q * P, m))] = v[q + m * Q];
                __syncthreads();
#pragma unroll
                for (int e = 0; e < E; ++e) v[e] = lds[lds_index<N, TB, CFAST>(t, p + P * e)];
            } else {
                R *lds = (R *)lds_raw;
                if constexpr (S > 0) __syncthreads();
#pragma unroll
                for (int q = 0; q < Q; ++q)
#pragma unroll
                    for (int m = 0; m < r; ++m)
                        lds[lds_index<N, TB, CFAST>(t, stage_dst<N, E, S>(p + q * P, m))] = v[q + m * Q].x;
                __syncthreads();
#pragma unroll
                for (int e = 0; e < E; ++e) v[e].x = lds[lds_index<N, TB, CFAST>(t, p + P * e)];
                __syncthreads();
#pragma unroll
                for (int q = 0; q < Q; ++q)
#pragma unroll
                    for (int m = 0; m < r; ++m)
                        lds[lds_index<N, TB, CFAST>(t, stage_dst<N, E, S>(p + q * P, m))] = v[q + m * Q].y;
                __syncthreads();
#pragma unroll
                for (int e = 0; e < E; ++e) v[e].y = lds[lds_index<N, TB, CFAST>(t, p + P * e)];
            }
        }
        run_stages<R, N, E, TB, CFAST, SPLIT, SIGN, S + 1>(v, p, t, tw, lds_raw);
    }
}

// One launch = one axis pass over TB transforms per block.  The four-step
// post-twiddle (a.pt_n != 0) and the output scale are runtime-uniform.
template <typename R, int N, int E, int TB, bool CFAST, bool SPLIT, int SIGN>
__global__ void __launch_bounds__(TB *(N / E))
axis_kernel(const cpx<R> *in, cpx<R> *out, const AxisArgs<R> a)
{
    using SC = Sched<N, E>;
    constexpr int P = SC::P;
    constexpr bool USE_LDS = (P > 1) && (SC::NST > 1);
    constexpr int LDS_BYTES = USE_LDS ? Exchange<R, N, E, TB, CFAST, SPLIT>::BYTES : 16;
    __shared__ __align__(16) unsigned char lds[LDS_BYTES];

    const int tid = threadIdx.x;
    const int t = CFAST ? (tid % TB) : (tid / P);
    const int p = CFAST ? (tid / TB) : (tid % P);
    const long long o = blockIdx.x / a.tiles;
    const int tile = blockIdx.x - (int)(o * a.tiles);
    const int c = tile * TB + t;
    const bool valid = c < a.C;

    cpx<R> v[E];
    if (valid) {
        const cpx<R> *src = in + o * a.in_os + (long long)c * a.in_cs;
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = src[(long long)(p + P * e) * a.in_is];
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = cpx<R>{0, 0};
    }

    run_stages<R, N, E, TB, CFAST, SPLIT, SIGN, 0>(v, p, t, a.tw, lds);

    if (valid) {
        if (a.pt_n) {
            const long long base = o * a.pt_o + (long long)c * a.pt_c;
            const long long mask = (1LL << a.pt_shift) - 1;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const long long k = p + P * e;
                const long long ex = (base * k) & (a.pt_n - 1);
                const cpx<R> w = cmul(a.pt_lo[ex & mask], a.pt_hi[ex >> a.pt_shift]);
                v[e] = twmul<SIGN>(v[e], w);
            }
        }
        cpx<R> *dst = out + o * a.out_os + (long long)c * a.out_cs;
#pragma unroll
        for (int e = 0; e < E; ++e) dst[(long long)(p + P * e) * a.out_is] = scale(v[e], a.scale);
    }
}

}  // namespace tope
