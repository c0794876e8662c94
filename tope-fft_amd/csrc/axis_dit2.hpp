// Strided-axis pass that overlaps its loads with its arithmetic.
//
// A column tile of the long strided axis (4 columns of 4096 FP64 points,
// 256 KiB) fills half of a CU's register file, so only one block runs per
// CU and, in the one-shot axis kernel, its load, compute and store phases
// follow each other with the memory system idle during the arithmetic.
// This kernel splits each column's N-point transform by decimation in time,
//     X[k]       = E[k] + W_N^k O[k]
//     X[k + N/2] = E[k] - W_N^k O[k],   k < N/2,
// where E and O are the N/2-point transforms of the even and odd rows.
// Every thread issues the loads of both halves up front; the even half's
// transform (the fused Stockham stages of the axis kernel, on N/2 points)
// then runs while the odd half's loads are still in flight, and the odd
// half follows.  Registers per thread are the same as for the one-shot
// tile; the LDS exchange area is half as large, which leaves room for the
// stage twiddles and the W_N^k table in LDS.
//
// Replaces, like the axis kernel, the reference's per-stage strided passes
// (src/fft2d.c:74-134 with src/kernels2D.cl:15-288; src/fft3d.c:80-214 with
// src/kernels3D.cl:15-298): one HBM read and one write per element.
#pragma once
#include "axis_kernel.hpp"

namespace tope {

template <typename R, int N, int E, int TB, bool SPLIT>
struct Dit2Lds {
    static constexpr int H = N / 2;
    static constexpr int XB = Exchange<R, H, E, TB, true, SPLIT>::BYTES;
    static constexpr int TWN = TwLds<H, E>::TOTAL;
    static constexpr int CB = (int)sizeof(cpx<R>);
    // stage twiddles in LDS when they fit beside the exchange area, and the
    // W_N^k table (H entries) too when that still fits
    static constexpr bool TWL = XB + TWN * CB <= 163840;
    static constexpr bool WKL = TWL && XB + (TWN + H) * CB <= 163840;
    static constexpr int BYTES = XB + (TWL ? TWN * CB : 0) + (WKL ? H * CB : 0);
};

// a.tw: stage table of Sched<N/2, E>; a.dif_tw: W_N^k, k < N/2 (forward
// sign).  Column tiles only (lanes walk the TB columns first).
template <typename R, int N, int E, int TB, bool SPLIT, int SIGN, bool POST, bool BIG>
__device__ __forceinline__ void dit2_block(const cpx<R> *in, cpx<R> *out, const AxisArgs<R> &a, long long blk,
                                           unsigned char *lds)
{
    constexpr int H = N / 2;
    using SC = Sched<H, E>;
    using L = Dit2Lds<R, N, E, TB, SPLIT>;
    constexpr int P = SC::P;
    constexpr int CB = L::CB;
    constexpr int NT = TB * P;

    const int tid = threadIdx.x;
    const int t = tid % TB;
    const int p = tid / TB;
    cpx<R> *twl = reinterpret_cast<cpx<R> *>(lds + L::XB);
    cpx<R> *wkl = twl + L::TWN;
    if constexpr (L::TWL) {
        for (int i = tid; i < L::TWN; i += NT) twl[i] = a.tw[TwLds<H, E>::src(i)];
    }
    if constexpr (L::WKL) {
        for (int i = tid; i < H; i += NT) wkl[i] = a.dif_tw[i];
    }

    const long long o = blk / a.tiles;
    int tile = (int)(blk - o * a.tiles) + a.rot;
    if (tile >= a.tiles) tile -= a.tiles;
    const int c0 = tile * TB;
    const int c = c0 + t;
    const bool valid = c < a.C;

    // even rows 2j and odd rows 2j + 1, j = p + P e
    cpx<R> ve[E], vo[E];
    if constexpr (!BIG) {
        const long long ibase = o * a.in_os + (long long)c0 * a.in_cs;
        const __amdgpu_buffer_rsrc_t rin = make_rsrc(in + ibase, (a.in_elems - ibase) * CB);
        const int voff = (int)((t * a.in_cs + 2LL * p * a.in_is) * CB);
        const int sstep = (int)(2LL * P * a.in_is * CB);
        const int odd = (int)(a.in_is * CB);
        if (valid) {
#pragma unroll
            for (int e = 0; e < E; ++e) ve[e] = buf_load<R>(rin, voff, e * sstep);
#pragma unroll
            for (int e = 0; e < E; ++e) vo[e] = buf_load<R>(rin, voff, e * sstep + odd);
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) ve[e] = vo[e] = cpx<R>{0, 0};
        }
    } else {
        if (valid) {
            const cpx<R> *src = in + o * a.in_os + (long long)c * a.in_cs;
#pragma unroll
            for (int e = 0; e < E; ++e) ve[e] = src[(2LL * (p + P * e)) * a.in_is];
#pragma unroll
            for (int e = 0; e < E; ++e) vo[e] = src[(2LL * (p + P * e) + 1) * a.in_is];
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) ve[e] = vo[e] = cpx<R>{0, 0};
        }
    }

    run_stages<R, H, E, TB, true, SPLIT, SIGN, 0, L::TWL>(ve, p, t, a.tw, lds, twl);
    __syncthreads();  // the odd half reuses the exchange area
    run_stages<R, H, E, TB, true, SPLIT, SIGN, 0, L::TWL>(vo, p, t, a.tw, lds, twl);

    if (!valid) return;
    // radix-2 combine: thread p holds E[k], O[k] for k = p + P e
    cpx<R> lo[E], hi[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = p + P * e;
        const cpx<R> w = L::WKL ? wkl[k] : a.dif_tw[k];
        const cpx<R> ow = twmul<SIGN>(vo[e], w);
        lo[e] = ve[e] + ow;
        hi[e] = ve[e] - ow;
    }
    if constexpr (POST) {
        const long long base = o * a.pt_o + (long long)c * a.pt_c;
        const long long mask = (1LL << a.pt_shift) - 1;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const long long k = p + P * e;
            const long long e0 = (base * k) & (a.pt_n - 1), e1 = (base * (k + H)) & (a.pt_n - 1);
            lo[e] = twmul<SIGN>(lo[e], cmul(a.pt_lo[e0 & mask], a.pt_hi[e0 >> a.pt_shift]));
            hi[e] = twmul<SIGN>(hi[e], cmul(a.pt_lo[e1 & mask], a.pt_hi[e1 >> a.pt_shift]));
        }
    }
    if (a.scale != R(1)) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            lo[e] = scale(lo[e], a.scale);
            hi[e] = scale(hi[e], a.scale);
        }
    }
    if constexpr (!BIG) {
        const long long obase = o * a.out_os + (long long)c0 * a.out_cs;
        const __amdgpu_buffer_rsrc_t rout = make_rsrc(out + obase, (a.out_elems - obase) * CB);
        const int voff = (int)((t * a.out_cs + (long long)p * a.out_is) * CB);
        const int sstep = (int)((long long)P * a.out_is * CB);
        const int half = (int)((long long)H * a.out_is * CB);
#pragma unroll
        for (int e = 0; e < E; ++e) buf_store<R>(lo[e], rout, voff, e * sstep);
#pragma unroll
        for (int e = 0; e < E; ++e) buf_store<R>(hi[e], rout, voff, e * sstep + half);
    } else {
        cpx<R> *dst = out + o * a.out_os + (long long)c * a.out_cs;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            dst[(long long)(p + P * e) * a.out_is] = lo[e];
            dst[(long long)(p + P * e + H) * a.out_is] = hi[e];
        }
    }
}

template <typename R, int N, int E, int TB, bool SPLIT, int SIGN, bool POST, bool BIG>
__global__ void __launch_bounds__(TB *(N / 2 / E))
dit2_kernel(const cpx<R> *in, cpx<R> *out, const AxisArgs<R> a)
{
    __shared__ __align__(16) unsigned char lds[Dit2Lds<R, N, E, TB, SPLIT>::BYTES];
    dit2_block<R, N, E, TB, SPLIT, SIGN, POST, BIG>(in, out, a, xcd_block(a.xcd), lds);
}

}  // namespace tope
