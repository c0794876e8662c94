// Mixed-precision transform rows on the matrix cores (gfx950 MFMA).
//
// Storage is complex64.  Every radix-16 stage of a length-N transform
// (N = 16^s) is a dense 16 x 16 complex DFT-matrix product on MFMA with FP32
// accumulation: 16 butterflies of a stage form the 16 columns of the B
// operand, the DFT matrix is the A operand, in its real 16 x 32 form
//     [Yr]   [Fr  -Fi] [Xr]
//     [Yi] = [Fi   Fr] [Xi].
// There is no reference counterpart (the reference is FP64-only and runs
// radix-2/4/8 butterflies, src/kernels1D.cl:15-497); the Stockham data flow,
// stage twiddles and the sign convention are those of the FP64 axis kernel.
//
// Operand precision (MODE):
//   MP_BF16X3  x = x0 + x1 (bf16 truncation split, |x - x0 - x1| < 2^-16|x|),
//              F = F0 + F1 (bf16 round-to-nearest from FP64); products
//              F0x0 + F0x1 + F1x0 on v_mfma_f32_16x16x32_bf16.
//   MP_BF16X6  x = x0 + x1 + x2 (exact for FP32 x), F = F0 + F1 + F2; the six
//              products with i + j <= 2, dropped terms ~2^-24.
//   MP_FP32    plain FP32 operands on v_mfma_f32_16x16x4_f32.
//
// Lane layout.  Lane l of a wave handles butterfly c = l % 16 of a group of
// 16 and holds its inputs m = 4g .. 4g+3 (g = l / 16), real and imaginary
// parts; after the product it holds outputs i = 4g .. 4g+3 of the same
// butterfly.  The K index of the product is permuted so that this holds for
// both instruction shapes:
//   16x16x32 bf16: lane group g supplies K slots 8g..8g+7 = Re x_{4g..4g+3},
//                  Im x_{4g..4g+3}.
//   16x16x4 f32:   instruction kk (0..7) takes K = 4kk + g from lane group g
//                  = Re x_{4g+kk} (kk < 4) or Im x_{4g+kk-4} (kk >= 4).
// A holds the matching columns; D (both shapes) gives lane l the rows
// 4g .. 4g+3 of column c.
#pragma once
#include "axis_kernel.hpp"

namespace tope {

enum { MP_VALU = 0, MP_BF16X3 = 1, MP_BF16X6 = 2, MP_FP32 = 3 };

typedef __bf16 tope_bf16x8 __attribute__((ext_vector_type(8)));
typedef float tope_f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int tope_u32x4v __attribute__((ext_vector_type(4)));

// Per-lane DFT-16 operand constants, generated on the GPU in FP64 and
// rounded once.  For each direction (0 forward, 1 inverse):
//   bf16: part p (0..2) of A_re and A_im, 4 dwords (8 bf16) per lane
//   f32:  A_re, A_im for the 8 instructions, one float per lane each
struct MfmaConsts {
    unsigned int bf[2][2][3][64][4];  // [dir][re/im][part][lane][dword]
    float f32[2][2][8][64];           // [dir][re/im][kk][lane]
};

struct MfmaArgs {
    long long count;       // transforms
    const cpx<float> *tw;  // stage twiddles (FP64-generated, rounded), layout of Sched<N,16>
    const MfmaConsts *k;   // DFT-16 operand constants
    float scale;           // output scale (1/N on inverse, a power of two: exact)
    long long in_elems, out_elems;
};

__device__ __forceinline__ unsigned int f2u(float x) { return __builtin_bit_cast(unsigned int, x); }
__device__ __forceinline__ float u2f(unsigned int x) { return __builtin_bit_cast(float, x); }
// pack the high halves (bf16 truncations) of two floats: lo -> bits 0..15
__device__ __forceinline__ unsigned int pack_hi(unsigned int lo, unsigned int hi)
{
    return __builtin_amdgcn_perm(hi, lo, 0x07060302u);
}

// Split the lane's 8 real operands (Re x0..3, Im x0..3) into bf16 parts.
template <int PARTS>
__device__ __forceinline__ void split_bf16(const float (&x)[8], tope_u32x4v (&part)[PARTS])
{
    unsigned int b0[8], b1[8], b2[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const unsigned int u = f2u(x[s]);
        b0[s] = u;
        const float r = x[s] - u2f(u & 0xffff0000u);  // exact: the low 16 mantissa bits
        const unsigned int ur = f2u(r);
        b1[s] = ur;
        if constexpr (PARTS == 3) b2[s] = f2u(r - u2f(ur & 0xffff0000u));  // exact, <= 8 bits
    }
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        part[0][d] = pack_hi(b0[2 * d], b0[2 * d + 1]);
        part[1][d] = pack_hi(b1[2 * d], b1[2 * d + 1]);
        if constexpr (PARTS == 3) part[2][d] = pack_hi(b2[2 * d], b2[2 * d + 1]);
    }
}

__device__ __forceinline__ tope_f32x4 mfma_bf16(tope_u32x4v a, tope_u32x4v b, tope_f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(tope_bf16x8, a),
                                                   __builtin_bit_cast(tope_bf16x8, b), c, 0, 0, 0);
}

// One radix-16 DFT on MFMA for the 16 butterflies of this wave's group.
// in: x[s] = Re/Im of the lane's 4 inputs (slots 0..3 real, 4..7 imaginary);
// out: y[r] = output 4g + r (complex).
template <int MODE>
__device__ __forceinline__ void dft16_mfma(const float (&x)[8], cpx<float> (&y)[4], const MfmaConsts *__restrict__ K,
                                           int dir, int lane)
{
    tope_f32x4 ar = {0, 0, 0, 0}, ai = {0, 0, 0, 0};
    if constexpr (MODE == MP_FP32) {
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) {
            const float b = x[kk];
            ar = __builtin_amdgcn_mfma_f32_16x16x4f32(K->f32[dir][0][kk][lane], b, ar, 0, 0, 0);
            ai = __builtin_amdgcn_mfma_f32_16x16x4f32(K->f32[dir][1][kk][lane], b, ai, 0, 0, 0);
        }
    } else {
        constexpr int PARTS = MODE == MP_BF16X6 ? 3 : 2;
        tope_u32x4v xp[PARTS];
        split_bf16<PARTS>(x, xp);
        tope_u32x4v fr[PARTS], fi[PARTS];
#pragma unroll
        for (int p = 0; p < PARTS; ++p) {
            const unsigned int *q = K->bf[dir][0][p][lane];
            const unsigned int *w = K->bf[dir][1][p][lane];
            fr[p] = tope_u32x4v{q[0], q[1], q[2], q[3]};
            fi[p] = tope_u32x4v{w[0], w[1], w[2], w[3]};
        }
        // terms with (F part) + (x part) <= PARTS - 1, smallest first
#pragma unroll
        for (int tot = PARTS - 1; tot >= 0; --tot)
#pragma unroll
            for (int pf = 0; pf <= tot; ++pf) {
                const int px = tot - pf;
                ar = mfma_bf16(fr[pf], xp[px], ar);
                ai = mfma_bf16(fi[pf], xp[px], ai);
            }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) y[r] = cpx<float>{ar[r], ai[r]};
}

// LDS index of element i of block-local transform t (one pad slot per 16).
template <int N>
__device__ __forceinline__ int mp_lds(int t, int i) { return t * (N + N / 16) + i + (i >> 4); }

// One block = TPB transforms of length N = 16^NST, WPT waves per transform,
// GPW groups of 16 butterflies per wave per stage.
template <int MODE, int N, int SIGN>
struct MfmaShape {
    static constexpr int NST = ilog2c(N) / 4;
    static constexpr int NB = N / 16;                    // butterflies per stage
    static constexpr int NG = NB / 16;                   // groups per stage
    static constexpr int WPT = NG >= 4 ? 4 : NG;         // waves per transform
    static constexpr int GPW = NG / WPT;                 // groups per wave
    static constexpr int TPB = 4 / WPT;                  // transforms per block (4 waves)
    static constexpr int THREADS = 256;
    static_assert(N == (1 << (4 * NST)) && NST >= 2, "N must be 16^s, s >= 2");
};

template <int MODE, int N, int SIGN>
__global__ void __launch_bounds__(256) mfma_rows_kernel(const cpx<float> *in, cpx<float> *out, const MfmaArgs a)
{
    using SH = MfmaShape<MODE, N, SIGN>;
    constexpr int NB = SH::NB, WPT = SH::WPT, GPW = SH::GPW, TPB = SH::TPB, NST = SH::NST;
    using SC = Sched<N, 16>;
    __shared__ cpx<float> lds[TPB * (N + N / 16)];

    const int tid = threadIdx.x;
    // wave index made provably uniform: buffer descriptors derived from it
    // must live in SGPRs (otherwise the compiler emits waterfall loops)
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int c = lane & 15, g = lane >> 4;
    const int tb = wave / WPT, wt = wave % WPT;
    const long long row = (long long)blockIdx.x * TPB + tb;
    const bool valid = row < a.count;
    const int dir = SIGN < 0 ? 0 : 1;
    constexpr int CB = 8;

    float x[GPW][8];  // per group: Re x_{4g..4g+3}, Im x_{4g..4g+3}
    // stage-0 inputs straight from HBM: element j + NB*m, j = group*16 + c
    {
        const __amdgpu_buffer_rsrc_t rin = make_rsrc(in + row * N, (a.in_elems - row * N) * CB);
#pragma unroll
        for (int q = 0; q < GPW; ++q) {
            const int j = (wt * GPW + q) * 16 + c;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                cpx<float> v = valid ? buf_load<float>(rin, (j + NB * (4 * g + s)) * CB, 0) : cpx<float>{0, 0};
                x[q][s] = v.x;
                x[q][4 + s] = v.y;
            }
        }
    }

#pragma unroll
    for (int S = 0; S < NST; ++S) {
        const int Ns = 1 << (4 * S);
        cpx<float> y[GPW][4];
#pragma unroll
        for (int q = 0; q < GPW; ++q) {
            const int j = (wt * GPW + q) * 16 + c;
            if (S > 0) {
                // input m = 4g + s of butterfly j: W_{16 Ns}^{(j mod Ns) m}
                const int k = j & (Ns - 1);
                const cpx<float> *t = a.tw + SC::twoff(S) + k;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    // input m = 0 takes twiddle 1 (its table slot is clamped
                    // to m = 1 and then replaced, keeping the lanes converged)
                    const int m = 4 * g + s;
                    cpx<float> w = t[(m > 0 ? m - 1 : 0) * Ns];
                    if (m == 0) w = cpx<float>{1.0f, 0.0f};
                    const cpx<float> v = twmul<SIGN>(cpx<float>{x[q][s], x[q][4 + s]}, w);
                    x[q][s] = v.x;
                    x[q][4 + s] = v.y;
                }
            }
            dft16_mfma<MODE>(x[q], y[q], a.k, dir, lane);
        }
        if (S == NST - 1) {
            if (valid) {
                const __amdgpu_buffer_rsrc_t rout = make_rsrc(out + row * N, (a.out_elems - row * N) * CB);
#pragma unroll
                for (int q = 0; q < GPW; ++q) {
                    const int j = (wt * GPW + q) * 16 + c;
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        buf_store<float>(scale(y[q][r], a.scale), rout, (j + NB * (4 * g + r)) * CB, 0);
                }
            }
        } else {
            if (S > 0) __syncthreads();
#pragma unroll
            for (int q = 0; q < GPW; ++q) {
                const int j = (wt * GPW + q) * 16 + c;
                const int base = (j / Ns) * Ns * 16 + (j & (Ns - 1));
#pragma unroll
                for (int r = 0; r < 4; ++r) lds[mp_lds<N>(tb, base + (4 * g + r) * Ns)] = y[q][r];
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < GPW; ++q) {
                const int j = (wt * GPW + q) * 16 + c;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const cpx<float> v = lds[mp_lds<N>(tb, j + NB * (4 * g + s))];
                    x[q][s] = v.x;
                    x[q][4 + s] = v.y;
                }
            }
        }
    }
}

}  // namespace tope
