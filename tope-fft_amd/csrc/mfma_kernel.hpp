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
//   MP_FP16X3  x = x0 + x1 (fp16 round-to-nearest parts, the residual taken
//              exactly in FP32; |x - x0 - x1| <= 2^-22 |x| in the normal
//              range), F = F0 + F1 (fp16 parts from FP64); products
//              F1x0 + F0x1 + F0x0 on v_mfma_f32_16x16x32_f16.  fp16 has
//              5 exponent bits, so each transform is scaled by a power of
//              two 2^k, chosen from the largest |Re|, |Im| of its input,
//              so that the input lies below 2^7: a radix-16 stage grows the
//              largest magnitude by at most 16, and the last stage's input
//              stays below 2^15 < 65504 (N = 16^3).  Elements smaller than
//              2^-14 of the scaled range keep an absolute error <= 2^-25
//              there (fp16 subnormal spacing).  The output is scaled back
//              by 2^-k, exactly.
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

enum { MP_VALU = 0, MP_BF16X3 = 1, MP_BF16X6 = 2, MP_FP32 = 3, MP_FP16X3 = 4 };

typedef __bf16 tope_bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 tope_f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 tope_f16x2 __attribute__((ext_vector_type(2)));
typedef float tope_f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int tope_u32x4v __attribute__((ext_vector_type(4)));

// Per-lane DFT-16 operand constants, generated on the GPU in FP64 and
// rounded once.  For each direction (0 forward, 1 inverse):
//   bf16: part p (0..2) of A_re and A_im, 4 dwords (8 bf16) per lane
//   f32:  A_re, A_im for the 8 instructions, one float per lane each
//   f16:  part p (0..1) of A_re and A_im, same slot layout as bf16
struct MfmaConsts {
    unsigned int bf[2][2][3][64][4];  // [dir][re/im][part][lane][dword]
    float f32[2][2][8][64];           // [dir][re/im][kk][lane]
    unsigned int h[2][2][2][64][4];   // [dir][re/im][part][lane][dword]
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

__device__ __forceinline__ tope_f32x4 mfma_f16(tope_u32x4v a, tope_u32x4v b, tope_f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(tope_f16x8, a),
                                                  __builtin_bit_cast(tope_f16x8, b), c, 0, 0, 0);
}

// Split the lane's 8 real operands into two fp16 parts, packed two per
// dword in slot order: x0 = fp16(x) (round to nearest), x1 = fp16(x - x0).
// x - x0 is exact in FP32 (x0 lies within a factor of two of x, or is 0); it
// is formed by one v_fma_mix_f32 per value, -x0 * 1 + x with x0 read as an
// fp16 half of the packed dword (bit-identical to converting x0 to FP32 and
// subtracting, fp16 subnormals included).
__device__ __forceinline__ void split_f16(const float (&x)[8], tope_u32x4v &p0, tope_u32x4v &p1)
{
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        float a = x[2 * d], b = x[2 * d + 1];
        // keep a, b as FP32 values: otherwise the producing multiply-add is
        // fused into the fp16 conversion and then computed twice
        asm("" : "+v"(a), "+v"(b));
        const tope_f16x2 h = {(_Float16)a, (_Float16)b};
        const unsigned int hb = __builtin_bit_cast(unsigned int, h);
        float ra, rb;
        asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(ra) : "v"(hb), "v"(a));
        asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(rb) : "v"(hb), "v"(b));
        const tope_f16x2 l = {(_Float16)ra, (_Float16)rb};
        p0[d] = hb;
        p1[d] = __builtin_bit_cast(unsigned int, l);
    }
}

// One radix-16 DFT on MFMA for the 16 butterflies of this wave's group.
// in: x[s] = Re/Im of the lane's 4 inputs (slots 0..3 real, 4..7 imaginary);
// out: y[r] = Re of output 4g + r, y[4 + r] = its Im (the same slot layout).
template <int MODE>
__device__ __forceinline__ void dft16_mfma(const float (&x)[8], float (&y)[8], const MfmaConsts *__restrict__ K,
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
    } else if constexpr (MODE == MP_FP16X3) {
        tope_u32x4v x0, x1;
        split_f16(x, x0, x1);
        tope_u32x4v fr[2], fi[2];
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const unsigned int *q = K->h[dir][0][p][lane];
            const unsigned int *w = K->h[dir][1][p][lane];
            fr[p] = tope_u32x4v{q[0], q[1], q[2], q[3]};
            fi[p] = tope_u32x4v{w[0], w[1], w[2], w[3]};
        }
        // small terms first
        ar = mfma_f16(fr[1], x0, ar);
        ai = mfma_f16(fi[1], x0, ai);
        ar = mfma_f16(fr[0], x1, ar);
        ai = mfma_f16(fi[0], x1, ai);
        ar = mfma_f16(fr[0], x0, ar);
        ai = mfma_f16(fi[0], x0, ai);
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
    for (int r = 0; r < 4; ++r) {
        y[r] = ar[r];
        y[4 + r] = ai[r];
    }
}

// LDS word of element i (i < 4096) within a transform's slab: an XOR
// swizzle of i's low six bits by bits 5, 6, 7, 10, 11 (a bijection), chosen
// so that every exchange access of the kernel below -- stage-0 outputs
// (16 j + 4 g + r), stage-1 outputs ((j/16) 256 + j%16 + 16 (4 g + r)) and
// stage inputs (j + NB (4 g + s)) -- puts the 64 lanes of a wave on 64
// distinct banks (and each half wave on 32).  The swizzle is linear over
// GF(2), and bits 8, 9 and 12 up pass through unchanged, so fields in those
// bits can be added afterwards.
__device__ __forceinline__ constexpr int mp_swz(int i)
{
    return i ^ (((i >> 5) & 1) * 0x01) ^ (((i >> 6) & 1) * 0x12) ^ (((i >> 7) & 1) * 0x28) ^
           (((i >> 10) & 1) * 0x10) ^ (((i >> 11) & 1) * 0x20);
}

// One block = TPB transforms of length N = 16^NST, WPT <= WMAX waves per
// transform, GPW groups of 16 butterflies per wave per stage; blocks hold at
// least 4 waves.
template <int N, int WMAX>
struct MfmaShape {
    static constexpr int NST = ilog2c(N) / 4;
    static constexpr int NB = N / 16;                    // butterflies per stage
    static constexpr int NG = NB / 16;                   // groups per stage
    static constexpr int WPT = NG >= WMAX ? WMAX : NG;   // waves per transform
    static constexpr int GPW = NG / WPT;                 // groups per wave
    static constexpr int TPB = WPT >= 4 ? 1 : 4 / WPT;   // transforms per block
    static constexpr int THREADS = 64 * WPT * TPB;
    static_assert(N == (1 << (4 * NST)) && NST >= 2, "N must be 16^s, s >= 2");
};

// Stockham data flow as in the FP64 axis kernel: stage S (span Ns = 16^S)
// reads butterfly j's input m at j + NB m and writes its output r to
// (j / Ns) Ns 16 + (j mod Ns) + r Ns.  Between stages the outputs pass
// through LDS one component at a time (real parts, then imaginary parts),
// which halves the LDS a block holds.  Each group's inputs become its
// outputs in place, so a lane holds 8 GPW floats of transform data.
template <int MODE, int N, int SIGN, int WMAX>
__global__ void __launch_bounds__((MfmaShape<N, WMAX>::THREADS)) __attribute__((amdgpu_waves_per_eu(WMAX == 8 ? 8 : 1))) mfma_rows_kernel(const cpx<float> *in,
                                                                               cpx<float> *out, const MfmaArgs a)
{
    using SH = MfmaShape<N, WMAX>;
    constexpr int NB = SH::NB, WPT = SH::WPT, GPW = SH::GPW, TPB = SH::TPB, NST = SH::NST;
    using SC = Sched<N, 16>;
    __shared__ float lds[TPB * N];
    __shared__ float wave_max[SH::THREADS / 64];
    static_assert(N <= 4096, "the LDS swizzle covers 12 index bits");

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

    float out_scale = a.scale;
    if constexpr (MODE == MP_FP16X3) {
        // power-of-two range scale of this transform (all its waves agree)
        float m = 0.0f;
#pragma unroll
        for (int q = 0; q < GPW; ++q)
#pragma unroll
            for (int s = 0; s < 8; ++s) m = fmaxf(m, fabsf(x[q][s]));
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        if constexpr (WPT > 1) {
            if (lane == 0) wave_max[wave] = m;
            __syncthreads();
#pragma unroll
            for (int w = 0; w < WPT; ++w) m = fmaxf(m, wave_max[tb * WPT + w]);
        }
        // m = f 2^e, f in [0.5, 1): the input times 2^(7-e) lies below 2^7;
        // e is clamped so that both scale factors are normal floats
        int e = __builtin_amdgcn_frexp_expf(m);
        e = e < -119 ? -119 : (e > 126 ? 126 : e);
        const float up = __builtin_ldexpf(1.0f, 7 - e);
        out_scale *= __builtin_ldexpf(1.0f, e - 7);
#pragma unroll
        for (int q = 0; q < GPW; ++q)
#pragma unroll
            for (int s = 0; s < 8; ++s) x[q][s] *= up;
    }

#pragma unroll
    for (int S = 0; S < NST; ++S) {
        const int Ns = 1 << (4 * S);
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
            float y[8];
            dft16_mfma<MODE>(x[q], y, a.k, dir, lane);
#pragma unroll
            for (int s = 0; s < 8; ++s) x[q][s] = y[s];
        }
        if (S == NST - 1) {
            if (valid) {
                const __amdgpu_buffer_rsrc_t rout = make_rsrc(out + row * N, (a.out_elems - row * N) * CB);
                // the forward transform of the unscaled modes stores as is
                constexpr bool SCALED = MODE == MP_FP16X3 || SIGN > 0;
#pragma unroll
                for (int q = 0; q < GPW; ++q) {
                    const int j = (wt * GPW + q) * 16 + c;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        cpx<float> v{x[q][r], x[q][4 + r]};
                        if constexpr (SCALED) v = scale(v, out_scale);
                        buf_store<float>(v, rout, (j + NB * (4 * g + r)) * CB, 0);
                    }
                }
            }
        } else {
            // Write index of output r of group q: dst = base_q + (4g + r) Ns,
            // where base_q + 4g Ns and r Ns occupy disjoint bits, so
            // swz(dst) = swz(base_q + 4g Ns) ^ swz(r Ns); q enters base_q
            // as q * 256 (bits 8, 9: unswizzled), an additive offset.
            // Read index of input s: src = (j + 4g NB) | s NB; for NB = 256
            // s NB lies in bits 8, 9 (additive), otherwise it is XORed in.
            float *slab = lds + tb * N;
            static_assert(GPW * 256 <= 1024, "group index must stay in bits 8, 9");
            const int j0 = wt * GPW * 16 + c;
            int wadr[4], radr[GPW];
            {
                const int wb = mp_swz((j0 / Ns) * Ns * 16 + (j0 & (Ns - 1)) + 4 * g * Ns);
#pragma unroll
                for (int r = 0; r < 4; ++r) wadr[r] = wb ^ mp_swz(r * Ns);
            }
#pragma unroll
            for (int q = 0; q < GPW; ++q) radr[q] = mp_swz(j0 + q * 16 + NB * 4 * g);
#pragma unroll
            for (int part = 0; part < 2; ++part) {
                if (S > 0 || part > 0) __syncthreads();  // the previous reads of lds are done
#pragma unroll
                for (int q = 0; q < GPW; ++q)
#pragma unroll
                    for (int r = 0; r < 4; ++r) slab[wadr[r] + q * 256] = x[q][4 * part + r];
                __syncthreads();
#pragma unroll
                for (int q = 0; q < GPW; ++q)
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const int idx = NB == 256 ? radr[q] + NB * s : radr[q] ^ mp_swz(NB * s);
                        x[q][4 * part + s] = slab[idx];
                    }
            }
        }
    }
}


// ---- rows of N = R 16^2 (R = 2, 4, 8: N = 512, 1024, 2048) -------------
// A leading radix-R stage on the VALU (thread t of a transform takes
// butterflies t + T u), then the two radix-16 stages on MFMA as above.

// LDS word of element i of a transform's slab: XOR of fixed masks selected
// by index bits 5 and up into bits 0..5 (a bijection, linear over GF(2)),
// found by an exhaustive bank model of this kernel's three access patterns
// (radix-R outputs j R + r, stage inputs j + NB (4g + s), first-stage
// outputs (j / R) 16 R + j % R + (4g + r) R): every access puts a wave's 64
// lanes on 64 distinct banks and each half wave on 32.
template <int N>
__device__ __forceinline__ constexpr int rem_swz(int i)
{
    constexpr int m512[] = {21, 15, 48, 28, 0, 0};
    constexpr int m1024[] = {11, 21, 31, 61, 26, 0};
    constexpr int m2048[] = {19, 36, 9, 29, 21, 42};
    const int *m = N == 512 ? m512 : N == 1024 ? m1024 : m2048;
    int w = i;
#pragma unroll
    for (int b = 0; b < 6; ++b) w ^= ((i >> (5 + b)) & 1) * m[b];
    return w;
}

template <int N>
struct MfmaRemShape {
    using SC = Sched<N, 16>;
    static constexpr int R = SC::REM;
    static constexpr int NB = N / 16;                  // butterflies per MFMA stage
    static constexpr int NG = NB / 16;                 // groups per MFMA stage
    // up to 4 groups per lane (16 complex values): few waves per transform
    // and many transforms in flight per CU
    static constexpr int WPT = NG >= 8 ? NG / 4 : 1;   // waves per transform
    static constexpr int GPW = NG / WPT;               // groups per wave
    static constexpr int TPB = 4 / WPT;                // transforms per block (4 waves)
    static constexpr int T = 64 * WPT;                 // threads per transform
    static constexpr int U = N / R / T;                // radix-R butterflies per thread
    static constexpr int THREADS = 64 * WPT * TPB;
    static_assert(R > 1 && SC::NST == 3 && U * R == 4 * GPW, "N = R 16^2, R = 2, 4, 8");
};

template <int MODE, int N, int SIGN>
__global__ void __launch_bounds__((MfmaRemShape<N>::THREADS)) mfma_rows_rem_kernel(const cpx<float> *in,
                                                                                   cpx<float> *out, const MfmaArgs a)
{
    using SH = MfmaRemShape<N>;
    using SC = typename SH::SC;
    constexpr int R = SH::R, NB = SH::NB, WPT = SH::WPT, GPW = SH::GPW, TPB = SH::TPB, T = SH::T, U = SH::U;
    constexpr int CB = 8;
    __shared__ float lds[TPB * N];
    __shared__ float wave_max[SH::THREADS / 64];

    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int c = lane & 15, g = lane >> 4;
    const int tb = wave / WPT, wt = wave % WPT;
    const int t = wt * 64 + lane;  // thread within the transform
    const long long row = (long long)blockIdx.x * TPB + tb;
    const bool valid = row < a.count;
    const int dir = SIGN < 0 ? 0 : 1;
    float *slab = lds + tb * N;

    // radix-R inputs: butterfly j = t + T u takes elements j + (N/R) m; the
    // R U = 4 GPW values of a thread fill x[e / 4][e % 4] (re), [4 + e % 4] (im)
    float x[GPW][8];
    {
        const __amdgpu_buffer_rsrc_t rin = make_rsrc(in + row * N, (a.in_elems - row * N) * CB);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int m = 0; m < R; ++m) {
                const int e = u * R + m;
                const cpx<float> v = valid ? buf_load<float>(rin, (t + T * u + (N / R) * m) * CB, 0) : cpx<float>{0, 0};
                x[e / 4][e % 4] = v.x;
                x[e / 4][4 + e % 4] = v.y;
            }
    }

    float out_scale = a.scale;
    if constexpr (MODE == MP_FP16X3) {
        float m = 0.0f;
#pragma unroll
        for (int q = 0; q < GPW; ++q)
#pragma unroll
            for (int s = 0; s < 8; ++s) m = fmaxf(m, fabsf(x[q][s]));
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        if constexpr (WPT > 1) {
            if (lane == 0) wave_max[wave] = m;
            __syncthreads();
#pragma unroll
            for (int w = 0; w < WPT; ++w) m = fmaxf(m, wave_max[tb * WPT + w]);
        }
        // the last stage's input is at most N/16 times the largest input
        constexpr int TOP = 15 - ilog2c(N / 16);
        int e = __builtin_amdgcn_frexp_expf(m);
        e = e < TOP - 126 ? TOP - 126 : (e > TOP + 126 ? TOP + 126 : e);
        const float up = __builtin_ldexpf(1.0f, TOP - e);
        out_scale *= __builtin_ldexpf(1.0f, e - TOP);
#pragma unroll
        for (int q = 0; q < GPW; ++q)
#pragma unroll
            for (int s = 0; s < 8; ++s) x[q][s] *= up;
    }

    // radix-R stage (span 1: no twiddles); output r of butterfly j goes to j R + r
#pragma unroll
    for (int u = 0; u < U; ++u) {
        cpx<float> v[R];
#pragma unroll
        for (int m = 0; m < R; ++m) {
            const int e = u * R + m;
            v[m] = cpx<float>{x[e / 4][e % 4], x[e / 4][4 + e % 4]};
        }
        dft<R, SIGN>(v);
#pragma unroll
        for (int m = 0; m < R; ++m) {
            const int e = u * R + m;
            x[e / 4][e % 4] = v[m].x;
            x[e / 4][4 + e % 4] = v[m].y;
        }
    }
    // lane parts of the LDS words (every index is lane part ^ constant part
    // in disjoint bits, and rem_swz is linear)
    const int j0 = wt * GPW * 16 + c;  // butterfly of group q: j0 + 16 q
    const int w_rem = rem_swz<N>(t * R);
    const int w_in = rem_swz<N>(j0 + NB * 4 * g);
    const int w_s1 = rem_swz<N>(((j0 / R) << ilog2c(16 * R)) + (j0 & (R - 1)) + 4 * g * R);
#pragma unroll
    for (int part = 0; part < 2; ++part) {
        if (part > 0) __syncthreads();
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int e = u * R + r;
                slab[w_rem ^ rem_swz<N>(T * R * u + r)] = x[e / 4][4 * part + e % 4];
            }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < GPW; ++q)
#pragma unroll
            for (int s = 0; s < 4; ++s) x[q][4 * part + s] = slab[w_in ^ rem_swz<N>(16 * q + NB * s)];
    }

#pragma unroll
    for (int S = 1; S < 3; ++S) {
        const int Ns = SC::span(S);
#pragma unroll
        for (int q = 0; q < GPW; ++q) {
            const int j = j0 + 16 * q;
            // input m = 4g + s of butterfly j: W_{16 Ns}^{(j mod Ns) m}
            const cpx<float> *tw = a.tw + SC::twoff(S) + (j & (Ns - 1));
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int m = 4 * g + s;
                cpx<float> w = tw[(m > 0 ? m - 1 : 0) * Ns];
                if (m == 0) w = cpx<float>{1.0f, 0.0f};
                const cpx<float> v = twmul<SIGN>(cpx<float>{x[q][s], x[q][4 + s]}, w);
                x[q][s] = v.x;
                x[q][4 + s] = v.y;
            }
            float y[8];
            dft16_mfma<MODE>(x[q], y, a.k, dir, lane);
#pragma unroll
            for (int s = 0; s < 8; ++s) x[q][s] = y[s];
        }
        if (S == 2) {
            if (valid) {
                const __amdgpu_buffer_rsrc_t rout = make_rsrc(out + row * N, (a.out_elems - row * N) * CB);
                constexpr bool SCALED = MODE == MP_FP16X3 || SIGN > 0;
#pragma unroll
                for (int q = 0; q < GPW; ++q)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        cpx<float> v{x[q][r], x[q][4 + r]};
                        if constexpr (SCALED) v = scale(v, out_scale);
                        buf_store<float>(v, rout, (j0 + 16 * q + NB * (4 * g + r)) * CB, 0);
                    }
            }
        } else {
            // output r of butterfly j (span R): (j / R) 16 R + j % R + (4g + r) R
#pragma unroll
            for (int part = 0; part < 2; ++part) {
                __syncthreads();
#pragma unroll
                for (int q = 0; q < GPW; ++q)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        slab[w_s1 ^ rem_swz<N>(((16 * q / R) << ilog2c(16 * R)) + r * R)] = x[q][4 * part + r];
                __syncthreads();
#pragma unroll
                for (int q = 0; q < GPW; ++q)
#pragma unroll
                    for (int s = 0; s < 4; ++s) x[q][4 * part + s] = slab[w_in ^ rem_swz<N>(16 * q + NB * s)];
            }
        }
    }
}

}  // namespace tope
