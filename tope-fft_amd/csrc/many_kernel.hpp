// Persistent kernel for a packed batch of transforms of uneven power-of-two
// lengths (no reference counterpart: the reference plans one length per
// topeFFTInit, src/fft1d.c:303-370).
//
// Transforms are grouped by length class (n = 2^L, L = 4..14).  A task is up
// to TB transforms of one class, TB = 1024 / P threads' worth (P = n / E
// threads per transform), so every task moves the same ~16384 elements; the
// task list is ordered by class (largest first) and dealt round-robin to
// the resident blocks, which loop until the list is exhausted (every block
// reaches the end; no inter-block waits).  Each class runs a compile-time radix
// schedule (E = 16: radix-16 stages; E = 8: radix-8 stages) chosen per class
// at plan time, by the autotuner or by default.
#pragma once
#include "axis_kernel.hpp"

namespace tope {

// Two block sizes, by the threads a class's schedule needs per transform
// (P = n/E): 1024 for P > 256 (a 16384-point FP64 transform fills one
// block's registers, one block per CU), 256 for P <= 256 (four blocks per
// CU, so one block's loads overlap another's arithmetic).  A 512-thread
// tier for P = 512 measured slower on the 20,000-transform batch.
constexpr int MANY_THREADS = 1024;
constexpr int MANY_SMALL_THREADS = 256;
constexpr int MANY_LMIN = 4, MANY_LMAX = 14;

// One task: `count` transforms (<= TB of the class schedule) from position
// `first` of the class-grouped order.
struct ManyTask {
    int first;
    int count;
};

// Task j of class L (seg0[L] <= j < seg1[L]) covers the positions
// pos0[L] + k TB .. of the class-grouped order, k = j - seg0[L]: the task's
// range is arithmetic on kernel arguments, and a transform's element offset
// is one read of toffs (offsets in that order), so a block starts its data
// loads one dependent read after taking a task rather than three (task
// record, order entry, offset).
template <typename R>
struct ManyArgs {
    int seg0[16], seg1[16];  // tasks of class L: seg0[L] .. seg1[L]
    int pos0[16], cnt[16];   // class L: first position in the order, transforms
    const long long *toffs;  // element offset of the transform at each position of the order
    const cpx<R> *tw[16];    // stage twiddles of each class (layout of its schedule)
    int sched[16];           // per class: 0 -> E = 16, 1 -> E = 8
    long long elems;         // packed buffer length (elements)
};

template <int L, int TB, typename A>
__device__ __forceinline__ ManyTask many_task(const A &a, int k)
{
    ManyTask tk;
    tk.first = a.pos0[L] + k * TB;
    tk.count = min(TB, a.cnt[L] - k * TB);
    return tk;
}

template <typename R, int N, int E, int SIGN, int NT>
__device__ __forceinline__ void many_rows(const cpx<R> *in, cpx<R> *out, const ManyArgs<R> &a, int k, void *lds)
{
    constexpr int P = N / E;
    constexpr int TB = NT / P;
    using SC = Sched<N, E>;
    constexpr int CB = (int)sizeof(cpx<R>);
    const ManyTask tk = many_task<SC::LN, TB>(a, k);  // task k of the class
    // Opaque to the optimiser: nothing lane-dependent (LDS bases, twiddle
    // and data offsets) is hoisted out of the task loop, where it would stay
    // live across every class and spill under the 128-VGPR cap of a
    // 1024-thread block; it is recomputed per task instead (a few VALU ops).
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int t = tid / P, p = tid % P;
    const bool valid = t < tk.count;
    const long long off = valid ? a.toffs[tk.first + t] : 0;
    const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(in, a.elems * CB);
    const __amdgpu_buffer_rsrc_t wsrc = make_rsrc(out, a.elems * CB);
    const int voff = (int)((off + p) * CB);
    constexpr int sstep = P * CB;
    cpx<R> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = valid ? buf_load<R>(rsrc, voff, e * sstep) : cpx<R>{0, 0};
    run_stages<R, N, E, TB, false, true, SIGN, 0>(v, p, t, a.tw[SC::LN], lds);
    const R sc = SIGN > 0 ? (R)(1.0 / N) : (R)1;  // INVERSE: 1/n, a power of two (exact)
    if (valid) {
#pragma unroll
        for (int e = 0; e < E; ++e) buf_store<R>(scale(v[e], sc), wsrc, voff, e * sstep);
    }
}

// LDS: TB * (N + N/16) slots of one component for every class = NT * 17 words.
template <typename R, int NT>
constexpr int many_lds_bytes() { return NT / 16 * 17 * 16 * (int)sizeof(R); }

// Dispatch one task of class L to its schedule (E = 8 only where n/8 threads
// fit the block).
template <typename R, int L, int SIGN, int NT>
__device__ __forceinline__ void many_class(const cpx<R> *in, cpx<R> *out, const ManyArgs<R> &a, int k, void *lds,
                                           bool e8)
{
    constexpr int N = 1 << L;
    if constexpr (N >= 8 && N / 8 <= NT) {
        if (e8) {
            many_rows<R, N, 8, SIGN, NT>(in, out, a, k, lds);
            return;
        }
    }
    if constexpr ((N < 16 ? 1 : N / 16) <= NT) many_rows<R, N, (N < 16 ? N : 16), SIGN, NT>(in, out, a, k, lds);
}

// Tasks of one class: this block takes global task indices b, b + G, ...
// that fall in [seg0[L], seg1[L]) -- the same round-robin deal as one loop
// over the whole list, but each loop body holds one class's code, which keeps
// the register allocation of the 1024-thread block free of spills.
template <typename R, int L, int SIGN, int NT>
__device__ __forceinline__ void many_segment(const cpx<R> *in, cpx<R> *out, const ManyArgs<R> &a, void *lds)
{
    const int s0 = a.seg0[L], s1 = a.seg1[L];
    if (s0 >= s1) return;
    const int G = gridDim.x, b = blockIdx.x;
    int task = s0 + ((b - s0) % G + G) % G;
    const bool e8 = a.sched[L] != 0;
    for (; task < s1; task += G) {
        many_class<R, L, SIGN, NT>(in, out, a, task - s0, lds, e8);
        __syncthreads();  // the next task reuses the LDS
    }
}

// The classes a kernel holds code for (the host routes each class to one
// kernel through the segment bounds it passes): 16384, 8192 and 4096
// points (the last for its radix-8 schedule, 512 threads per transform) at
// 1024 threads; every class up to 4096 points at 256, capped at 128 VGPRs
// so four blocks share a CU.
template <typename R, int SIGN, int NT>
__global__ void __launch_bounds__(NT, NT >= 1024 ? 1 : 4) many_kernel(const cpx<R> *in, cpx<R> *out,
                                                                     const ManyArgs<R> a)
{
    __shared__ __align__(16) unsigned char lds[many_lds_bytes<R, NT>()];
    if constexpr (NT >= 1024) {
        many_segment<R, 14, SIGN, NT>(in, out, a, lds);
        many_segment<R, 13, SIGN, NT>(in, out, a, lds);
        many_segment<R, 12, SIGN, NT>(in, out, a, lds);
    } else {
        many_segment<R, 12, SIGN, NT>(in, out, a, lds);
        many_segment<R, 11, SIGN, NT>(in, out, a, lds);
        many_segment<R, 10, SIGN, NT>(in, out, a, lds);
        many_segment<R, 9, SIGN, NT>(in, out, a, lds);
        many_segment<R, 8, SIGN, NT>(in, out, a, lds);
        many_segment<R, 7, SIGN, NT>(in, out, a, lds);
        many_segment<R, 6, SIGN, NT>(in, out, a, lds);
        many_segment<R, 5, SIGN, NT>(in, out, a, lds);
        many_segment<R, 4, SIGN, NT>(in, out, a, lds);
    }
}

}  // namespace tope
