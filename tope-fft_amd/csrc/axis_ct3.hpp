// Strided 4096-point pass with one of its two exchanges inside each wave.
//
// The one-shot column tile (axis_kernel.hpp) runs the 4096-point transform
// as three radix-16 Stockham stages with two block-wide exchanges, each
// behind workgroup barriers, so every wave waits for the slowest wave's
// loads before the second stage.  Here the same tile (TB = 4 columns, 1024
// threads, 16 elements per thread) runs a three-step Cooley-Tukey
// decomposition, n = 256 n2 + 16 n1 + n0, k = k0 + 16 k1 + 256 k2:
//   A: for each (n1, n0): DFT-16 over n2                 -> A[k0; n1, n0]
//   B: times W_256^{n1 k0}, DFT-16 over n1               -> B[k0, k1; n0]
//   C: times W_4096^{n0 (k0 + 16 k1)}, DFT-16 over n0    -> X[k0 + 16 k1 + 256 k2]
// with thread (wave w, lane group g, column t) holding n0 = w, n1 = g in
// step A.  The exchange between A and B swaps the lane-group index with the
// register index inside one wave (the wave's own LDS rows, no workgroup
// barrier); only the exchange between B and C crosses waves.  A wave whose
// loads arrive early runs steps A and B while others still wait for theirs.
// Both exchanges go through one LDS region per wave (16 rows of 68 values
// of one component): the in-wave transpose reads and writes the wave's own
// region, and the cross-wave exchange has each wave write its values into
// its own region (rows of 64) and read the other waves' regions.  So no
// workgroup barrier precedes the cross-wave writes, and the steps A and B of
// each wave run as soon as its own loads arrive.  All accesses are
// bank-conflict free (rows of 68 doubles for the transpose; contiguous
// 64-value rows for the cross-wave exchange).
//
// Replaces, like the axis kernel, the reference's per-stage strided passes
// (src/fft2d.c:74-134 with src/kernels2D.cl:15-288; src/fft3d.c:80-214 with
// src/kernels3D.cl:15-298).
#pragma once
#include "axis_kernel.hpp"

namespace tope {

// Twiddle table of the pass (a.tw): W_256^j, j < 128, then W_4096^i, i < 128.
// Step B uses W_256^{m g} and step C W_4096^{m b} = W_4096^{m g} W_256^{m w}
// for b = g + 16 w (m = 1, 2, 4, 8: every exponent below 128).  A table of
// the products W_4096^{m b} in LDS (16 KiB more) measured slower: 105.8
// against 100.4 us for the 2D 4096^2 pass.
constexpr int CT3_TW = 128 + 128;

// Four-step post-twiddle factors of a tile (POST): for each column t of
// base b_t, W^{b_t g} (g < 16), W^{16 b_t w} (w < 16) and W^{256 b_t}.
constexpr int CT3_PT = 4 * 16 + 4 * 16 + 4;

template <typename R, bool POST = false>
struct Ct3Lds {
    static constexpr int WR = 16 * 68;                   // values of one component per wave region
    static constexpr int XB = 16 * WR * (int)sizeof(R);  // 16 wave regions
    static constexpr int BYTES = XB + (CT3_TW + (POST ? CT3_PT : 0)) * (int)sizeof(cpx<R>);
};

// Orders a wave's LDS writes before its reads of other lanes' values (and
// those reads before the wave's next writes): LDS operations of one wave
// execute in order; this keeps the compiler from moving them across.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// W_{pt_n}^x as the two table values whose product it is (pt_lookup's
// operands): the reads are issued here, the product formed later.
template <typename R>
__device__ __forceinline__ void pt_fetch(long long x, const AxisArgs<R> &a, typename RawCpx<R>::type &lo,
                                         typename RawCpx<R>::type &hi)
{
    x &= a.pt_n - 1;
    lo = raw_ld(a.pt_lo + (x & ((1LL << a.pt_shift) - 1)));
    hi = raw_ld(a.pt_hi + (x >> a.pt_shift));
}

template <typename R, int SIGN, bool POST, bool BIG>
__global__ void __launch_bounds__(1024) ct3_kernel(const cpx<R> *in, cpx<R> *out, const AxisArgs<R> a)
{
    constexpr int TB = 4;
    constexpr int CB = (int)sizeof(cpx<R>);
    using L = Ct3Lds<R, POST>;
    __shared__ __align__(16) unsigned char lds[L::BYTES];
    R *xl = reinterpret_cast<R *>(lds);
    cpx<R> *ltw = reinterpret_cast<cpx<R> *>(lds + L::XB);

    const int tid = threadIdx.x;
    const int t = tid & 3, g = (tid >> 2) & 15, w = tid >> 6;
    // step C's twiddle rows (read after the first workgroup barrier).
    // Without the post-twiddle: read into a register now (every thread, at a
    // clamped index), stored to LDS once the data loads are issued -- a
    // store right after the read waited for it before any data load went
    // out (2D 4096^2 Y pass 105.0-106.3 -> 103.6-105.4 us).  With it the
    // tile is at its register budget, and the held value made the inverse
    // instantiation spill (1D 2^24 307 -> 314 us): stored at once there.
    typename RawCpx<R>::type tabv{};
    if constexpr (POST) {
        if (tid < CT3_TW) ltw[tid] = a.tw[tid];
    } else {
        tabv = raw_ld(a.tw + (tid < CT3_TW ? tid : 0));
    }

    const long long blk = xcd_block(a.xcd);
    const long long o = blk / a.tiles;
    int tile = (int)(blk - o * a.tiles) + a.rot;
    if (tile >= a.tiles) tile -= a.tiles;
    const int c0 = tile * TB;
    const int c = c0 + t;
    const bool valid = c < a.C;
    // the post-twiddle factors of the tile's columns, read from the tables
    // before the data loads (at the end of the tile, after the last barrier,
    // table reads from L2 wait behind the streaming traffic)
    cpx<R> *ptl = ltw + CT3_TW;
    if constexpr (POST) {
        if (tid < CT3_PT) {
            const int ct = tid < 128 ? (tid >> 4) & 3 : tid - 128;
            const long long bt = o * a.pt_o + (long long)(c0 + ct) * a.pt_c;
            const long long x = tid < 64 ? bt * (tid & 15) : tid < 128 ? bt * 16 * (tid & 15) : bt * 256;
            ptl[tid] = pt_lookup(x, a.pt_lo, a.pt_hi, a.pt_n, a.pt_shift);
        }
    }

    // step B's twiddles W_256^{m g}, issued ahead of the data loads (read
    // after the first exchange, they waited behind the streaming traffic)
    cpx<R> twb[16];
#pragma unroll
    for (int m = 1; m < 16; m <<= 1) twb[m] = a.tw[m * g];

    // step A inputs: rows 256 e + 16 g + w of column c
    cpx<R> v[16];
    if constexpr (!BIG) {
        const long long ibase = o * a.in_os + (long long)c0 * a.in_cs;
        const __amdgpu_buffer_rsrc_t rin = make_rsrc(in + ibase, (a.in_elems - ibase) * CB);
        const int voff = (int)((t * a.in_cs + (long long)(16 * g + w) * a.in_is) * CB);
        const int sstep = (int)(256LL * a.in_is * CB);
        if (valid) {
#pragma unroll
            for (int e = 0; e < 16; ++e) v[e] = buf_load<R>(rin, voff, e * sstep);
        } else {
#pragma unroll
            for (int e = 0; e < 16; ++e) v[e] = cpx<R>{0, 0};
        }
    } else {
        if (valid) {
            const cpx<R> *src = in + o * a.in_os + (long long)c * a.in_cs;
#pragma unroll
            for (int e = 0; e < 16; ++e) v[e] = src[(long long)(256 * e + 16 * g + w) * a.in_is];
        } else {
#pragma unroll
            for (int e = 0; e < 16; ++e) v[e] = cpx<R>{0, 0};
        }
    }
    if constexpr (!POST) {
        // keeps the compiler from joining this store to the read above
        // (which would put the wait for the read before the data loads)
        asm volatile("" ::: "memory");
        if (tid < CT3_TW) ltw[tid] = raw_cpx<R>(tabv);
    }

    dft16<SIGN>(v);  // register k0

    R *mine = xl + w * L::WR;  // this wave's region
    // in-wave exchange: register k0 -> lane group k0, lane group n1 -> register n1
    {
        const int wb = 4 * g + t, rb = g * 68 + t;
#pragma unroll
        for (int comp = 0; comp < 2; ++comp) {
#pragma unroll
            for (int k = 0; k < 16; ++k) mine[k * 68 + wb] = comp ? v[k].y : v[k].x;
            wave_lds_sync();
#pragma unroll
            for (int r = 0; r < 16; ++r) (comp ? v[r].y : v[r].x) = mine[rb + 4 * r];
            wave_lds_sync();
        }
    }

    // step B: register n1 times W_256^{n1 k0} (k0 = g), DFT over n1 -> register k1
    // (the W_256 rows come from the global table: no barrier before this step)
    apply_twiddles<SIGN>(v, twb);
    dft16<SIGN>(v);

    // cross-wave exchange: register k1 -> wave k1, wave n0 -> register n0;
    // wave w writes row k of its own region, wave k reads row k of every region
    {
        const int lane = 4 * g + t;
#pragma unroll
        for (int comp = 0; comp < 2; ++comp) {
            if (comp) __syncthreads();  // every wave has read the real parts
#pragma unroll
            for (int k = 0; k < 16; ++k) mine[k * 64 + lane] = comp ? v[k].y : v[k].x;
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 16; ++r) (comp ? v[r].y : v[r].x) = xl[r * L::WR + w * 64 + lane];
        }
    }

    // step C: register n0 times W_4096^{n0 b}, b = g + 16 w; DFT over n0 -> register k2
    {
        cpx<R> tw[16];
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) tw[m] = cmul(ltw[128 + m * g], ltw[m * w]);
        apply_twiddles<SIGN>(v, tw);
    }
    dft16<SIGN>(v);

    if (!valid) return;
    // output rows k = g + 16 w + 256 e
    if constexpr (POST)  // W^{b k}, k = g + 16 w + 256 e: W^{b g} W^{16 b w} (W^{256 b})^e
        pt_chain<SIGN>(v, cmul(ptl[16 * t + g], ptl[64 + 16 * t + w]), ptl[128 + t]);
    if (a.scale != R(1)) {
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = scale(v[e], a.scale);
    }
    if constexpr (!BIG) {
        const long long obase = o * a.out_os + (long long)c0 * a.out_cs;
        const __amdgpu_buffer_rsrc_t rout = make_rsrc(out + obase, (a.out_elems - obase) * CB);
        const int voff = (int)((t * a.out_cs + (long long)(g + 16 * w) * a.out_is) * CB);
        const int sstep = (int)(256LL * a.out_is * CB);
#pragma unroll
        for (int e = 0; e < 16; ++e) buf_store<R>(v[e], rout, voff, e * sstep);
    } else {
        cpx<R> *dst = out + o * a.out_os + (long long)c * a.out_cs;
#pragma unroll
        for (int e = 0; e < 16; ++e) dst[(long long)(g + 16 * w + 256 * e) * a.out_is] = v[e];
    }
}

// Complex64 three-step tile: 8 columns (64-byte row segments, as the FP64
// tile's 4), each thread holding the two adjacent columns 2t, 2t + 1 of its
// lane position as one 16-byte element per row.  The steps, twiddles and
// exchanges are those of ct3_kernel, run once per column of the pair, with
// the twiddles shared; each exchange moves whole complex values (8-byte LDS
// accesses, the FP64 tile's bank pattern) in one round per column.
// Requires unit column stride, an even column count (whole pairs) and even
// row and plane strides (16-byte aligned pairs).
//
// Four-step post-twiddle (POST) of column c (base b_c): output k = g + 16 w
// + 256 e takes W^{b_c g} W^{16 b_c w} W^{256 b_c e}, three table values
// per output (FP32: a fresh product instead of a chain of 16), read into LDS
// at the tile's start: rows of 16 per column for g, w and e.
constexpr int CT3P_PT = 3 * 8 * 16;

template <bool POST>
struct Ct3PairLds {
    static constexpr int WR = 16 * 68;  // complex slots per wave region
    static constexpr int XB = 16 * WR * (int)sizeof(cpx<float>);
    static constexpr int BYTES = XB + (CT3_TW + (POST ? CT3P_PT : 0)) * (int)sizeof(cpx<float>);
};

// One row of a column pair as four dwords (re, im of column 2t, then 2t + 1).
// The elements go through scalar values: __builtin_bit_cast applied to an
// ext_vector element (u.y) reads the vector's first element with this
// compiler (the IR loads offset 0 for every element).
__device__ __forceinline__ void pair_split(tope_u32x4 u, cpx<float> &a, cpx<float> &b)
{
    const unsigned u0 = u[0], u1 = u[1], u2 = u[2], u3 = u[3];
    a.x = __uint_as_float(u0);
    a.y = __uint_as_float(u1);
    b.x = __uint_as_float(u2);
    b.y = __uint_as_float(u3);
}

__device__ __forceinline__ tope_u32x4 pair_join(cpx<float> a, cpx<float> b)
{
    return tope_u32x4{__float_as_uint(a.x), __float_as_uint(a.y), __float_as_uint(b.x), __float_as_uint(b.y)};
}

// In-wave exchange of one column (register k0 -> lane group k0, lane group
// n1 -> register n1) through the wave's region `mine`.  Force-inlined: the
// register arrays must stay in registers (an outlined call passes them
// through the stack).
__device__ __forceinline__ void ct3_pair_in_wave(cpx<float> (&v)[16], cpx<float> *mine, int g, int t)
{
    const int wb = 4 * g + t, rb = g * 68 + t;
#pragma unroll
    for (int k = 0; k < 16; ++k) mine[k * 68 + wb] = v[k];
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = mine[rb + 4 * r];
    wave_lds_sync();
}

// Cross-wave exchange of one column (register k1 -> wave k1, wave n0 ->
// register n0): wave w writes row k of its own region, wave k reads row k
// of every region.
__device__ __forceinline__ void ct3_pair_cross(cpx<float> (&v)[16], cpx<float> *mine, const cpx<float> *xl, int w,
                                               int lane)
{
#pragma unroll
    for (int k = 0; k < 16; ++k) mine[k * 64 + lane] = v[k];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = xl[r * Ct3PairLds<false>::WR + w * 64 + lane];
}

// Column cc's post-twiddle applied to its 16 outputs k = g + 16 w + 256 e.
template <int SIGN>
__device__ __forceinline__ void ct3_pair_post(cpx<float> (&v)[16], const cpx<float> *ptl, int cc, int g, int w)
{
    const cpx<float> s = cmul(ptl[16 * cc + g], ptl[128 + 16 * cc + w]);
    v[0] = twmul<SIGN>(v[0], s);
#pragma unroll
    for (int e = 1; e < 16; ++e) v[e] = twmul<SIGN>(v[e], cmul(s, ptl[256 + 16 * cc + e]));
}

template <int SIGN, bool POST, bool BIG>
__global__ void __launch_bounds__(1024) ct3_pair_kernel(const cpx<float> *in, cpx<float> *out,
                                                        const AxisArgs<float> a)
{
    using R = float;
    using L = Ct3PairLds<POST>;
    constexpr int TB = 8;
    constexpr int CB = (int)sizeof(cpx<R>);
    __shared__ __align__(16) unsigned char lds[L::BYTES];
    cpx<R> *xl = reinterpret_cast<cpx<R> *>(lds);
    cpx<R> *ltw = reinterpret_cast<cpx<R> *>(lds + L::XB);

    const int tid = threadIdx.x;
    const int t = tid & 3, g = (tid >> 2) & 15, w = tid >> 6;
    // step C's twiddle rows: read now, stored to LDS once the data loads are issued
    const typename RawCpx<R>::type tabv = raw_ld(a.tw + (tid < CT3_TW ? tid : 0));  // as in ct3_kernel

    const long long blk = xcd_block(a.xcd);
    const long long o = blk / a.tiles;
    int tile = (int)(blk - o * a.tiles) + a.rot;
    if (tile >= a.tiles) tile -= a.tiles;
    const int c0 = tile * TB;
    const int c = c0 + 2 * t;
    const bool valid = c < a.C;
    // the post-twiddle factors of the tile's columns: read now, formed and
    // stored to LDS once the data loads are issued
    cpx<R> *ptl = ltw + CT3_TW;
    typename RawCpx<R>::type plo{}, phi{};
    if constexpr (POST) {
        const int kind = tid >> 7, cc = (tid >> 4) & 7, j = tid & 15;
        const long long bt = o * a.pt_o + (long long)(c0 + cc) * a.pt_c;
        const long long x = tid >= CT3P_PT ? 0 : bt * (kind == 0 ? j : kind == 1 ? 16 * j : 256 * j);
        pt_fetch(x, a, plo, phi);
    }

    // step B's twiddles, issued ahead of the data loads
    cpx<R> twb[16];
#pragma unroll
    for (int m = 1; m < 16; m <<= 1) twb[m] = a.tw[m * g];

    // step A inputs: rows 256 e + 16 g + w of columns c, c + 1
    cpx<R> v0[16], v1[16];
    if (valid) {
        if constexpr (!BIG) {
            const long long ibase = o * a.in_os + (long long)c0;
            const __amdgpu_buffer_rsrc_t rin = make_rsrc(in + ibase, (a.in_elems - ibase) * CB);
            const int voff = (int)((2 * t + (long long)(16 * g + w) * a.in_is) * CB);
            const int sstep = (int)(256LL * a.in_is * CB);
#pragma unroll
            for (int e = 0; e < 16; ++e)
                pair_split(__builtin_amdgcn_raw_buffer_load_b128(rin, voff, e * sstep, 0), v0[e], v1[e]);
        } else {
            const tope_u32x4 *src = reinterpret_cast<const tope_u32x4 *>(in + o * a.in_os + (long long)c);
            const long long st = a.in_is / 2;  // rows in pairs (even stride)
#pragma unroll
            for (int e = 0; e < 16; ++e) pair_split(src[(long long)(256 * e + 16 * g + w) * st], v0[e], v1[e]);
        }
    } else {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            v0[e] = cpx<R>{0, 0};
            v1[e] = cpx<R>{0, 0};
        }
    }
    asm volatile("" ::: "memory");  // as in ct3_kernel: the stores stay after the data loads
    if (tid < CT3_TW) ltw[tid] = raw_cpx<R>(tabv);
    if constexpr (POST) {
        if (tid < CT3P_PT) ptl[tid] = cmul(raw_cpx<R>(plo), raw_cpx<R>(phi));
    }

    dft16<SIGN>(v0);  // register k0
    dft16<SIGN>(v1);

    cpx<R> *mine = xl + w * L::WR;  // this wave's region
    // in-wave exchange, one column at a time
    ct3_pair_in_wave(v0, mine, g, t);
    ct3_pair_in_wave(v1, mine, g, t);

    // step B: register n1 times W_256^{n1 k0} (k0 = g), DFT over n1 -> register k1
    apply_twiddles<SIGN>(v0, twb);  // forms every twb[m]
#pragma unroll
    for (int m = 1; m < 16; ++m) v1[m] = twmul<SIGN>(v1[m], twb[m]);
    dft16<SIGN>(v0);
    dft16<SIGN>(v1);

    // cross-wave exchange, one column at a time
    ct3_pair_cross(v0, mine, xl, w, 4 * g + t);
    __syncthreads();  // every wave has read the first column
    ct3_pair_cross(v1, mine, xl, w, 4 * g + t);

    // step C: register n0 times W_4096^{n0 b}, b = g + 16 w; DFT over n0 -> register k2
    {
        cpx<R> tw[16];
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) tw[m] = cmul(ltw[128 + m * g], ltw[m * w]);
        apply_twiddles<SIGN>(v0, tw);
#pragma unroll
        for (int m = 1; m < 16; ++m) v1[m] = twmul<SIGN>(v1[m], tw[m]);
    }
    dft16<SIGN>(v0);
    dft16<SIGN>(v1);

    if (!valid) return;
    if constexpr (POST) {
        ct3_pair_post<SIGN>(v0, ptl, 2 * t, g, w);
        ct3_pair_post<SIGN>(v1, ptl, 2 * t + 1, g, w);
    }
    if (a.scale != R(1)) {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            v0[e] = scale(v0[e], a.scale);
            v1[e] = scale(v1[e], a.scale);
        }
    }
    // output rows k = g + 16 w + 256 e
    if constexpr (!BIG) {
        const long long obase = o * a.out_os + (long long)c0;
        const __amdgpu_buffer_rsrc_t rout = make_rsrc(out + obase, (a.out_elems - obase) * CB);
        const int voff = (int)((2 * t + (long long)(g + 16 * w) * a.out_is) * CB);
        const int sstep = (int)(256LL * a.out_is * CB);
        // whole offset in the VGPR, soffset 0 (the store-data hazard note at buf_store)
#pragma unroll
        for (int e = 0; e < 16; ++e)
            __builtin_amdgcn_raw_buffer_store_b128(pair_join(v0[e], v1[e]), rout, voff + e * sstep, 0, 0);
    } else {
        tope_u32x4 *dst = reinterpret_cast<tope_u32x4 *>(out + o * a.out_os + (long long)c);
        const long long st = a.out_is / 2;
#pragma unroll
        for (int e = 0; e < 16; ++e) dst[(long long)(g + 16 * w + 256 * e) * st] = pair_join(v0[e], v1[e]);
    }
}

}  // namespace tope
