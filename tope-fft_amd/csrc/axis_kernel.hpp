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
// stages of radix RX = min(E, 16), preceded by one smaller stage when N is
// not a power of RX.  With E = 32 a thread runs two radix-16 butterflies per
// stage.  Shared by the host (twiddle table layout) and the device.
template <int N, int E>
struct Sched {
    static constexpr int RX = E < 16 ? E : 16;
    static constexpr int LN = ilog2c(N), LR = ilog2c(RX);
    static constexpr int NFULL = LR ? LN / LR : 0;
    static constexpr int REM = 1 << (LR ? LN % LR : LN);
    static constexpr int NST = NFULL + (REM > 1 ? 1 : 0);
    static constexpr int P = N / E;
    __host__ __device__ static constexpr int radix(int s) { return (REM > 1 && s == 0) ? REM : RX; }
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
    int xcd;    // 1: remap blocks so each XCD walks a contiguous range of tiles
    int rot;    // 0 <= rot < tiles: logical tile t of an outer index runs tile
                // (t + rot) mod tiles (a partial last tile can go first)
    long long in_elems, out_elems;  // buffer sizes (bounds of the buffer resources)
    // DIF2 row passes: block-local transform t = 2q + h is row (tile*TB/2 + q)
    // shifted by h*pair elements; afterwards the two rows of each pair are
    // combined by one radix-2 decimation-in-frequency step of the NEXT axis
    // (length 2*half_rows): row h=0 <- r0 + r1, row h=1 <- (r0 - r1) W^{row},
    // W = W_{2*half_rows} read from dif_tw (forward sign).
    long long pair;
    const cpx<R> *dif_tw;
};

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Strided tiles that are neighbours in memory share 128-byte lines, so
// with remapping the blocks an XCD receives back to back are adjacent
// tiles.  When G is a multiple of 8 x 32 blocks, XCD x takes tiles
// 256 j + 32 x + (0..31) for its j-th run of 32 blocks: the blocks in flight
// at a time cover one contiguous range of tiles across the XCDs (3D 512^3
// column passes 789 -> 772 and 886 -> 851 us, 2D 8192 x 4096 Y pass
// 356 -> 297 us, 4096^2 unchanged; 16384 x 4096 569 -> 589 us).  Otherwise,
// with G = 8q + r blocks, XCD x < r takes q + 1 blocks and the others q,
// each a contiguous range of logical tiles (block b = x + 8k runs tile k
// plus the blocks of the XCDs before x).  Both are bijections on [0, G).
__device__ __forceinline__ long long xcd_block(int remap)
{
    const long long b = blockIdx.x, G = gridDim.x;
    if (!remap) return b;
    const long long x = b & 7, k = b >> 3;
    if ((G & 255) == 0) return ((k >> 5) << 8) + (x << 5) + (k & 31);
    const long long q = G >> 3, r = G & 7;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// LDS pad shift: element i sits at slot i + (i >> K).  Column tiles (TB
// transforms interleaved) use K = 4.  Rows use K = 5 once a transform spans
// at least 32 threads (P = N/E >= 32) and K = 4 below: with 64 four-byte
// banks and eight-byte accesses served 32 lanes a cycle, a bank model of
// every exchange puts the conflict cost of K = 4 at 1.50-2.00x the
// conflict-free cycles for those rows and of K = 5 at 1.17-1.75x, while
// K = 4 is the better one (1.00-1.50x) for the shorter rows.
template <int N, int E, bool CFAST>
constexpr int lds_pad_shift() { return (!CFAST && N / E >= 32) ? 5 : 4; }

// LDS index of element i of block-local transform t: one pad slot per 2^K
// elements breaks the power-of-two strides of the Stockham scatter.
template <int N, int E, int TB, bool CFAST>
__device__ __forceinline__ int lds_index(int t, int i)
{
    constexpr int K = lds_pad_shift<N, E, CFAST>();
    const int pi = i + (i >> K);
    if constexpr (CFAST) return pi * TB + t;
    else return t * (N + N / 16) + pi;
}

template <int N, int TB>
constexpr int lds_slots() { return TB * (N + N / 16); }

// The pad map i -> i + (i >> K) is linear over every access group used
// here: element p + P*e (P a power of two) and the radix-r scatter
// base + m*Ns (Ns, r powers of two) never carry across a 2^K boundary
// differently from their offset, so padded(base + d) = padded(base) +
// d + (d >> K) with d = P*e or m*Ns (checked for every compiled shape by
// a CPU model of the exchanges).  This lets every LDS access be one base
// register plus a compile-time offset.
template <int N, int E, int TB, bool CFAST>
constexpr int lds_off(int d)
{
    constexpr int K = lds_pad_shift<N, E, CFAST>();
    return (d + (d >> K)) * (CFAST ? TB : 1);
}

// LDS exchange of the E register elements of every thread: write element
// e to slot dst(e), then read slots p + P*e back.  SPLIT exchanges the real
// and imaginary parts in two rounds, halving the LDS footprint (this is what
// lets a block hold four 4096-point FP64 columns).
template <typename R, int N, int E, int TB, bool CFAST, bool SPLIT>
struct Exchange {
    static constexpr int BYTES = lds_slots<N, TB>() * (SPLIT ? sizeof(R) : 2 * sizeof(R));
};
// Buffer resource over [base, base + bytes) (gfx9 raw buffer, no swizzle);
// accesses past num_records read zero and drop stores.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, long long bytes)
{
    const long long lim = 0x7fffffffLL;
    const int nrec = (int)(bytes < lim ? (bytes > 0 ? bytes : 0) : lim);
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, nrec, 0x00020000);
}

typedef unsigned int tope_u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int tope_u32x2 __attribute__((ext_vector_type(2)));

template <typename R>
__device__ __forceinline__ cpx<R> buf_load(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    if constexpr (sizeof(R) == 8) {
        tope_u32x4 u = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
        return __builtin_bit_cast(cpx<R>, u);
    } else {
        tope_u32x2 u = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
        return __builtin_bit_cast(cpx<R>, u);
    }
}

// 128-bit stores take their whole offset in the VGPR and soffset = 0.  A
// buffer store of more than 64 bits of data whose soffset is an SGPR is
// exempt from the compiler's store-data hazard check (no wait state before
// a VALU instruction overwrites the data registers), and on gfx950 such a
// VALU write issued right after the store has replaced the upper dwords of
// the stored value in some lanes.  With an inline-constant soffset the
// compiler inserts the wait state.
template <typename R>
__device__ __forceinline__ void buf_store(cpx<R> v, __amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    if constexpr (sizeof(R) == 8)
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(tope_u32x4, v), r, voff + soff, 0, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(tope_u32x2, v), r, voff, soff, 0);
}

// Table values read at a block's start and used after its data loads are
// issued travel as raw vectors: a complex local held across the loads was
// kept by the compiler in an LDS slot of its own (16 KiB more LDS and a wait
// for the read before the first data load).
template <typename R>
struct RawCpx {
    typedef tope_u32x4 type;
};
template <>
struct RawCpx<float> {
    typedef tope_u32x2 type;
};
template <typename R>
__device__ __forceinline__ typename RawCpx<R>::type raw_ld(const cpx<R> *p)
{
    return *reinterpret_cast<const typename RawCpx<R>::type *>(p);
}
template <typename R>
__device__ __forceinline__ cpx<R> raw_cpx(typename RawCpx<R>::type u)
{
    return __builtin_bit_cast(cpx<R>, u);
}

// W_{pt_n}^x = pt_lo[x & mask] * pt_hi[x >> pt_shift] (x reduced mod pt_n).
template <typename R>
__device__ __forceinline__ cpx<R> pt_lookup(long long x, const cpx<R> *pt_lo, const cpx<R> *pt_hi, long long pt_n,
                                            int pt_shift)
{
    x &= pt_n - 1;
    return cmul(pt_lo[x & ((1LL << pt_shift) - 1)], pt_hi[x >> pt_shift]);
}

// v[e] *= w s^e: a chain of products of the step s (relative error ~E ulp
// in FP64).
template <int SIGN, typename R, int E>
__device__ __forceinline__ void pt_chain(cpx<R> (&v)[E], cpx<R> w, cpx<R> s)
{
#pragma unroll
    for (int e = 0; e < E; ++e) {
        if (e) w = cmul(w, s);
        v[e] = twmul<SIGN>(v[e], w);
    }
}

// Four-step post-twiddle of one thread's outputs k = p + P e (e < E) of a
// line with twiddle base b: v[e] *= W_{pt_n}^{b k}.  W^{b (p + P e)} is
// W^{b p} (W^{b P})^e: a fresh table value every RS outputs, products of
// the step W^{b P} in between (FP64: one chain of E - 1 products, relative
// error ~E ulp; FP32: fresh every 4 outputs); b k < 2^63.
template <int SIGN, typename R, int E>
__device__ __forceinline__ void post_twiddle(cpx<R> (&v)[E], long long b, long long p, long long P,
                                             const cpx<R> *pt_lo, const cpx<R> *pt_hi, long long pt_n, int pt_shift)
{
    constexpr int RS = sizeof(R) == 8 ? E : 4;
    if constexpr (RS == E) {
        pt_chain<SIGN>(v, pt_lookup(b * p, pt_lo, pt_hi, pt_n, pt_shift), pt_lookup(b * P, pt_lo, pt_hi, pt_n, pt_shift));
    } else {
        const cpx<R> s = pt_lookup(b * P, pt_lo, pt_hi, pt_n, pt_shift);
        cpx<R> w = pt_lookup(b * p, pt_lo, pt_hi, pt_n, pt_shift);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (e) w = (e % RS == 0) ? pt_lookup(b * (p + P * e), pt_lo, pt_hi, pt_n, pt_shift) : cmul(w, s);
            v[e] = twmul<SIGN>(v[e], w);
        }
    }
}

// Twiddles W^{k*m}, m = 1..r-1, of one butterfly: only the power-of-two
// exponents m = 1, 2, 4, 8 are read; every other one is the product of two
// earlier entries (m = 2^a + b, b < 2^a), at most three products deep, and
// is applied as soon as it is formed (apply_twiddles) so that at most eight
// twiddles are live beside the data.
//
// The global stage table has layout [(m-1)*Ns + k] from entry `base`; it is
// read through a buffer resource so the lane part of the address is one
// 32-bit offset (k) shared by every exponent and stage, the rest uniform.
template <typename R, int r>
__device__ __forceinline__ void load_twiddles(cpx<R> (&w)[r], __amdgpu_buffer_rsrc_t t, int kb, int base, int Ns)
{
    constexpr int CB = (int)sizeof(cpx<R>);
#pragma unroll
    for (int m = 1; m < r; m <<= 1) w[m] = buf_load<R>(t, kb, (base + (m - 1) * Ns) * CB);
}

template <int SIGN, typename R, int r>
__device__ __forceinline__ void apply_twiddles(cpx<R> (&x)[r], cpx<R> (&w)[r])
{
#pragma unroll
    for (int m = 1; m < r; ++m) {
        if (m & (m - 1)) {
            const int hi = 1 << (31 - __builtin_clz(m));
            w[m] = cmul(w[hi], w[m - hi]);
        }
        x[m] = twmul<SIGN>(x[m], w[m]);
    }
}

// LDS copy of the power-of-two rows of the stage table, for blocks that own
// a whole CU: a table read from L2 waits behind the streaming HBM traffic,
// an LDS read does not.  Stage s >= 1 holds log2(r) rows of Ns entries,
// row b = exponent 2^b.
template <int N, int E>
struct TwLds {
    using SC = Sched<N, E>;
    static constexpr int count(int s) { return ilog2c(SC::radix(s)) * SC::span(s); }
    static constexpr int off(int s)
    {
        int o = 0;
        for (int i = 1; i < s; ++i) o += count(i);
        return o;
    }
    static constexpr int TOTAL = off(SC::NST);
    // global table index of LDS entry i
    __device__ static int src(int i)
    {
        int s = 1;
        while (s + 1 < SC::NST && i >= off(s + 1)) ++s;
        const int j = i - off(s), ns = SC::span(s);
        const int b = j / ns, k = j - b * ns;
        return SC::twoff(s) + ((1 << b) - 1) * ns + k;
    }
};

template <typename R, int r>
__device__ __forceinline__ void load_twiddles_lds(cpx<R> (&w)[r], const cpx<R> *t, int Ns)
{
#pragma unroll
    for (int b = 0; (1 << b) < r; ++b) w[1 << b] = t[b * Ns];
}

template <typename R, int N, int E, int SIGN, int S, bool TWL>
__device__ __forceinline__ void run_stage(cpx<R> (&v)[E], int p, const cpx<R> *__restrict__ tw, const cpx<R> *twl)
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
            cpx<R> w[r];
            if constexpr (TWL) load_twiddles_lds<R, r>(w, twl + TwLds<N, E>::off(S) + k, Ns);
            else
                load_twiddles<R, r>(w, make_rsrc(tw, (long long)SC::TWSIZE * (int)sizeof(cpx<R>)),
                                    k * (int)sizeof(cpx<R>), SC::twoff(S), Ns);
            apply_twiddles<SIGN>(x, w);
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

template <typename R, int N, int E, int TB, bool CFAST, bool SPLIT, int SIGN, int S, bool TWL = false>
__device__ __forceinline__ void run_stages(cpx<R> (&v)[E], int p, int t, const cpx<R> *__restrict__ tw,
                                           void *lds_raw, const cpx<R> *twl = nullptr)
{
    using SC = Sched<N, E>;
    run_stage<R, N, E, SIGN, S, TWL>(v, p, tw, twl);
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
            constexpr int Ns = SC::span(S);
            const int rbase = lds_index<N, E, TB, CFAST>(t, p);
            int wbase[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) wbase[q] = lds_index<N, E, TB, CFAST>(t, stage_dst<N, E, S>(p + q * P, 0));
            if constexpr (!SPLIT) {
                cpx<R> *lds = (cpx<R> *)lds_raw;
                if constexpr (S > 0) __syncthreads();
#pragma unroll
                for (int q = 0; q < Q; ++q)
#pragma unroll
                    for (int m = 0; m < r; ++m) lds[wbase[q] + lds_off<N, E, TB, CFAST>(m * Ns)] = v[q + m * Q];
                __syncthreads();
#pragma unroll
                for (int e = 0; e < E; ++e) v[e] = lds[rbase + lds_off<N, E, TB, CFAST>(P * e)];
            } else {
                R *lds = (R *)lds_raw;
                if constexpr (S > 0) __syncthreads();
#pragma unroll
                for (int q = 0; q < Q; ++q)
#pragma unroll
                    for (int m = 0; m < r; ++m) lds[wbase[q] + lds_off<N, E, TB, CFAST>(m * Ns)] = v[q + m * Q].x;
                __syncthreads();
#pragma unroll
                for (int e = 0; e < E; ++e) v[e].x = lds[rbase + lds_off<N, E, TB, CFAST>(P * e)];
                __syncthreads();
#pragma unroll
                for (int q = 0; q < Q; ++q)
#pragma unroll
                    for (int m = 0; m < r; ++m) lds[wbase[q] + lds_off<N, E, TB, CFAST>(m * Ns)] = v[q + m * Q].y;
                __syncthreads();
#pragma unroll
                for (int e = 0; e < E; ++e) v[e].y = lds[rbase + lds_off<N, E, TB, CFAST>(P * e)];
            }
        }
        run_stages<R, N, E, TB, CFAST, SPLIT, SIGN, S + 1, TWL>(v, p, t, tw, lds_raw, twl);
    }
}

template <typename R, int N, int E, int TB, bool CFAST, bool SPLIT, bool DIF2>
constexpr int axis_exchange_bytes()
{
    using SC = Sched<N, E>;
    return (DIF2 || ((SC::P > 1) && (SC::NST > 1))) ? Exchange<R, N, E, TB, CFAST, SPLIT>::BYTES : 16;
}
// Stage twiddles go to LDS when the exchange alone already limits the block
// to one per CU (more than half of the 160 KiB) and the table still fits.
template <typename R, int N, int E, int TB, bool CFAST, bool SPLIT, bool DIF2>
constexpr bool axis_tw_lds()
{
    constexpr int X = axis_exchange_bytes<R, N, E, TB, CFAST, SPLIT, DIF2>();
    return Sched<N, E>::NST > 1 && X > 81920 && X + TwLds<N, E>::TOTAL * (int)sizeof(cpx<R>) <= 163840;
}
template <typename R, int N, int E, int TB, bool CFAST, bool SPLIT, bool DIF2>
constexpr int axis_lds_bytes()
{
    return axis_exchange_bytes<R, N, E, TB, CFAST, SPLIT, DIF2>() +
           (axis_tw_lds<R, N, E, TB, CFAST, SPLIT, DIF2>() ? TwLds<N, E>::TOTAL * (int)sizeof(cpx<R>) : 0);
}

// One logical block (tile blk) of an axis pass: TB transforms.  POST applies
// the four-step post-twiddle; the output scale is runtime-uniform.
//
// Global accesses go through buffer resources based at the block's first
// element: each thread keeps one 32-bit byte offset and the per-element
// stride e*P*is*16 is a uniform (SGPR) offset, which keeps address
// registers out of the way of the 4*E data registers.  Blocks whose
// footprint exceeds 2^31 bytes run the BIG instantiation (64-bit pointers).
template <typename R, int N, int E, int TB, bool CFAST, bool SPLIT, int SIGN, bool POST, bool BIG, bool DIF2>
__device__ __forceinline__ void axis_block(const cpx<R> *in, cpx<R> *out, const AxisArgs<R> &a, long long blk,
                                           unsigned char *lds)
{
    using SC = Sched<N, E>;
    constexpr int P = SC::P;
    static_assert(!DIF2 || (!CFAST && TB % 2 == 0), "DIF2 pairs rows within a row block");
    constexpr int CB = (int)sizeof(cpx<R>);
    constexpr bool TWL = axis_tw_lds<R, N, E, TB, CFAST, SPLIT, DIF2>();
    constexpr int XB = axis_exchange_bytes<R, N, E, TB, CFAST, SPLIT, DIF2>();

    const int tid = threadIdx.x;
    const int t = CFAST ? (tid % TB) : (tid / P);
    const int p = CFAST ? (tid / TB) : (tid % P);
    // LDS twiddle rows (stage 1 reads them after the first exchange's
    // barriers): read into registers ahead of the data loads, stored once
    // they are issued -- a store right after each read waited for it before
    // the first data load went out
    cpx<R> *twl = reinterpret_cast<cpx<R> *>(lds + XB);
    constexpr int NTH = TB * P;
    constexpr int TWN = TWL ? (TwLds<N, E>::TOTAL + NTH - 1) / NTH : 1;
    typename RawCpx<R>::type twr[TWN];
    if constexpr (TWL) {
#pragma unroll
        for (int j = 0; j < TWN; ++j) {
            const int i = tid + j * NTH;
            twr[j] = raw_ld(a.tw + TwLds<N, E>::src(i < TwLds<N, E>::TOTAL ? i : 0));
        }
    }
    const long long o = blk / a.tiles;
    int tile = (int)(blk - o * a.tiles) + a.rot;
    if (tile >= a.tiles) tile -= a.tiles;
    // DIF2: TB/2 row pairs per block, transform t = 2q + h
    constexpr int TROWS = DIF2 ? TB / 2 : TB;
    const int c0 = tile * TROWS;
    const int c = c0 + (DIF2 ? (t >> 1) : t);
    const long long hoff = DIF2 ? (long long)(t & 1) * a.pair : 0;
    const int tr = DIF2 ? (t >> 1) : t;  // row within the tile
    const bool valid = c < a.C;

    cpx<R> v[E];
    if constexpr (!BIG) {
        const long long ibase = o * a.in_os + (long long)c0 * a.in_cs;
        const __amdgpu_buffer_rsrc_t rin = make_rsrc(in + ibase, (a.in_elems - ibase) * CB);
        const int voff = (int)((tr * a.in_cs + hoff + (long long)p * a.in_is) * CB);
        const int sstep = (int)((long long)P * a.in_is * CB);
        if (valid) {
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = buf_load<R>(rin, voff, e * sstep);
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = cpx<R>{0, 0};
        }
    } else {
        if (valid) {
            const cpx<R> *src = in + o * a.in_os + (long long)c * a.in_cs + hoff;
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = src[(long long)(p + P * e) * a.in_is];
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = cpx<R>{0, 0};
        }
    }

    if constexpr (TWL) {
        asm volatile("" ::: "memory");  // the stores stay after the data loads
#pragma unroll
        for (int j = 0; j < TWN; ++j) {
            const int i = tid + j * NTH;
            if (i < TwLds<N, E>::TOTAL) twl[i] = raw_cpx<R>(twr[j]);
        }
        run_stages<R, N, E, TB, CFAST, SPLIT, SIGN, 0, true>(v, p, t, a.tw, lds, twl);
    } else {
        run_stages<R, N, E, TB, CFAST, SPLIT, SIGN, 0>(v, p, t, a.tw, lds);
    }

    if constexpr (DIF2) {
        // radix-2 DIF step of the next axis between the rows of each pair,
        // exchanged through LDS one component at a time (no extra registers)
        R *l = (R *)lds;
        const int mate = t ^ 1;
        const int rbase = lds_index<N, E, TB, CFAST>(t, p), mbase = lds_index<N, E, TB, CFAST>(mate, p);
#pragma unroll
        for (int comp = 0; comp < 2; ++comp) {
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) l[rbase + lds_off<N, E, TB, CFAST>(P * e)] = comp ? v[e].y : v[e].x;
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const R m = l[mbase + lds_off<N, E, TB, CFAST>(P * e)];
                R &mine = comp ? v[e].y : v[e].x;
                mine = (t & 1) ? (m - mine) : (mine + m);
            }
        }
        if (t & 1) {
            const cpx<R> w = a.dif_tw[c < a.C ? c : 0];
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = twmul<SIGN>(v[e], w);
        }
    }

    if (!valid) return;
    if constexpr (POST)
        post_twiddle<SIGN>(v, o * a.pt_o + (long long)c * a.pt_c, p, P, a.pt_lo, a.pt_hi, a.pt_n, a.pt_shift);
    if (a.scale != R(1)) {  // uniform: only the last pass of an inverse scales
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = scale(v[e], a.scale);
    }
    if constexpr (!BIG) {
        const long long obase = o * a.out_os + (long long)c0 * a.out_cs;
        const __amdgpu_buffer_rsrc_t rout = make_rsrc(out + obase, (a.out_elems - obase) * CB);
        const int voff = (int)((tr * a.out_cs + hoff + (long long)p * a.out_is) * CB);
        const int sstep = (int)((long long)P * a.out_is * CB);
#pragma unroll
        for (int e = 0; e < E; ++e) buf_store<R>(v[e], rout, voff, e * sstep);
    } else {
        cpx<R> *dst = out + o * a.out_os + (long long)c * a.out_cs + hoff;
#pragma unroll
        for (int e = 0; e < E; ++e) dst[(long long)(p + P * e) * a.out_is] = v[e];
    }
}

// One launch = one axis pass, one workgroup per logical block.
template <typename R, int N, int E, int TB, bool CFAST, bool SPLIT, int SIGN, bool POST, bool BIG, bool DIF2 = false>
__global__ void __launch_bounds__(TB *(N / E))
axis_kernel(const cpx<R> *in, cpx<R> *out, const AxisArgs<R> a)
{
    __shared__ __align__(16) unsigned char lds[axis_lds_bytes<R, N, E, TB, CFAST, SPLIT, DIF2>()];
    axis_block<R, N, E, TB, CFAST, SPLIT, SIGN, POST, BIG, DIF2>(in, out, a, xcd_block(a.xcd), lds);
}

}  // namespace tope
