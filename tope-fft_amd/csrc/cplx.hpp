// Complex arithmetic and in-register small DFTs for the fused axis kernels.
//
// The radix-2/4/8/16 butterflies here replace the reference's per-stage
// OpenCL bodies DIT2C2C / DIT4C2C / DIT8C2C (src/kernels1D.cl:15-497,
// src/kernels2D.cl:15-537, src/kernels3D.cl:15-564).  The reference applies
// one radix per full-array launch on bit-reversed data; here a whole
// transform stays in registers/LDS and the butterflies take natural-order
// inputs (self-sorting Stockham), so no bit-reversal pass exists.
//
// Sign convention: SIGN = -1 is the reference's FORWARD (W = e^{-2 pi i/N}),
// SIGN = +1 its INVERSE (conjugate twiddles, kernels1D.cl:70).
#pragma once
#include <hip/hip_runtime.h>

namespace tope {

template <typename R>
struct alignas(2 * sizeof(R)) cpx {
    R x, y;
};

template <typename R>
__device__ __forceinline__ cpx<R> mk(R a, R b) { return cpx<R>{a, b}; }
template <typename R>
__device__ __forceinline__ cpx<R> operator+(cpx<R> a, cpx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> operator-(cpx<R> a, cpx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> scale(cpx<R> a, R s) { return {a.x * s, a.y * s}; }

// a * b
template <typename R>
__device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename R>
__device__ __forceinline__ cpx<R> cmulc(cpx<R> a, cpx<R> b)
{
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
// a * w where w is a forward-table twiddle, conjugated when SIGN = +1
template <int SIGN, typename R>
__device__ __forceinline__ cpx<R> twmul(cpx<R> a, cpx<R> w)
{
    if constexpr (SIGN < 0) return cmul(a, w);
    else return cmulc(a, w);
}

// multiply by e^{SIGN * i * pi/2}: -i for forward, +i for inverse
template <int SIGN, typename R>
__device__ __forceinline__ cpx<R> rot4(cpx<R> a)
{
    if constexpr (SIGN < 0) return {a.y, -a.x};
    else return {-a.y, a.x};
}

// multiply by (c + SIGN*i*s) for compile-time constants c, s
template <int SIGN, typename R>
__device__ __forceinline__ cpx<R> rotc(cpx<R> a, R c, R s)
{
    const R ss = SIGN < 0 ? -s : s;
    return {a.x * c - a.y * ss, a.x * ss + a.y * c};
}

// cos/sin of k*pi/8 for k = 1, 2, 3 (full double precision literals)
#define TOPE_C1_16 0.92387953251128675612818318939678828682
#define TOPE_S1_16 0.38268343236508977172845998403039886676
#define TOPE_R2 0.70710678118654752440084436210484903928

template <int SIGN, typename R>
__device__ __forceinline__ void dft2(cpx<R> &a, cpx<R> &b)
{
    cpx<R> t = a;
    a = t + b;
    b = t - b;
}

template <int SIGN, typename R>
__device__ __forceinline__ void dft4(cpx<R> &a0, cpx<R> &a1, cpx<R> &a2, cpx<R> &a3)
{
    cpx<R> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = rot4<SIGN>(a1 - a3);
    a0 = t0 + t2;
    a2 = t0 - t2;
    a1 = t1 + t3;
    a3 = t1 - t3;
}

// b + s * a, componentwise, as fused multiply-adds
template <typename R>
__device__ __forceinline__ cpx<R> fmac(cpx<R> a, R s, cpx<R> b)
{
    if constexpr (sizeof(R) == 8) return {__builtin_fma(s, a.x, b.x), __builtin_fma(s, a.y, b.y)};
    else return {__builtin_fmaf(s, a.x, b.x), __builtin_fmaf(s, a.y, b.y)};
}

// The twiddles (+-1 +- i)/sqrt2 are applied as an exact add/subtract
// (rotc with unit coefficients) whose 1/sqrt2 factor is folded into the
// following butterfly as fused multiply-adds.

// 8-point DFT, natural order in and out: two 4-point DFTs on the even and
// odd samples, then W8^k on the odd half.
template <int SIGN, typename R>
__device__ __forceinline__ void dft8(cpx<R> (&v)[8])
{
    cpx<R> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cpx<R> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<SIGN>(e0, e1, e2, e3);
    dft4<SIGN>(o0, o1, o2, o3);
    const R r2 = (R)TOPE_R2;
    const cpx<R> u1 = rotc<SIGN>(o1, R(1), R(1)), u3 = rotc<SIGN>(o3, R(-1), R(1));
    o2 = rot4<SIGN>(o2);
    v[0] = e0 + o0; v[4] = e0 - o0;
    v[1] = fmac(u1, r2, e1); v[5] = fmac(u1, -r2, e1);
    v[2] = e2 + o2; v[6] = e2 - o2;
    v[3] = fmac(u3, r2, e3); v[7] = fmac(u3, -r2, e3);
}

// 16-point DFT as 4 x 4: X[k1 + 4 k2] = sum_n2 W4^{n2 k2} W16^{n2 k1}
// sum_n1 a[4 n1 + n2] W4^{n1 k1}.
template <int SIGN, typename R>
__device__ __forceinline__ void dft16(cpx<R> (&v)[16])
{
    // column DFTs over n1 for each n2: a[n2], a[n2+4], a[n2+8], a[n2+12]
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) dft4<SIGN>(v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]);
    // now v[n2 + 4*k1] = B[n2][k1]; apply W16^{n2*k1}: W^1, W^3, W^3, W^9
    // directly, W^4 as a quarter turn, W^2 (v6, v9) and W^6 (v11, v14) as
    // r2 times a unit-coefficient rotation
    const R c1 = (R)TOPE_C1_16, s1 = (R)TOPE_S1_16, r2 = (R)TOPE_R2;
    const cpx<R> b5 = rotc<SIGN>(v[5], c1, s1), b13 = rotc<SIGN>(v[13], s1, c1);
    const cpx<R> b7 = rotc<SIGN>(v[7], s1, c1), b15 = rotc<SIGN>(v[15], -c1, -s1);
    const cpx<R> b10 = rot4<SIGN>(v[10]);
    const cpx<R> u6 = rotc<SIGN>(v[6], R(1), R(1)), u9 = rotc<SIGN>(v[9], R(1), R(1));
    const cpx<R> u11 = rotc<SIGN>(v[11], R(-1), R(1)), u14 = rotc<SIGN>(v[14], R(-1), R(1));
    // row DFTs over n2 for each k1 (inputs v[4*k1 + n2]); output a_k2 -> X[k1 + 4*k2]
    cpx<R> o[16];
    {  // k1 = 0
        cpx<R> b0 = v[0], b1 = v[1], b2 = v[2], b3 = v[3];
        dft4<SIGN>(b0, b1, b2, b3);
        o[0] = b0; o[4] = b1; o[8] = b2; o[12] = b3;
    }
    {  // k1 = 1: b2 = r2 u6
        const cpx<R> t0 = fmac(u6, r2, v[4]), t1 = fmac(u6, -r2, v[4]);
        const cpx<R> t2 = b5 + b7, t3 = rot4<SIGN>(b5 - b7);
        o[1] = t0 + t2; o[9] = t0 - t2; o[5] = t1 + t3; o[13] = t1 - t3;
    }
    {  // k1 = 2: b1 = r2 u9, b3 = r2 u11
        const cpx<R> t0 = v[8] + b10, t1 = v[8] - b10;
        const cpx<R> sp = u9 + u11, dm = rot4<SIGN>(u9 - u11);
        o[2] = fmac(sp, r2, t0); o[10] = fmac(sp, -r2, t0); o[6] = fmac(dm, r2, t1); o[14] = fmac(dm, -r2, t1);
    }
    {  // k1 = 3: b2 = r2 u14
        const cpx<R> t0 = fmac(u14, r2, v[12]), t1 = fmac(u14, -r2, v[12]);
        const cpx<R> t2 = b13 + b15, t3 = rot4<SIGN>(b13 - b15);
        o[3] = t0 + t2; o[11] = t0 - t2; o[7] = t1 + t3; o[15] = t1 - t3;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = o[i];
}

template <int RADIX, int SIGN, typename R>
__device__ __forceinline__ void dft(cpx<R> (&v)[RADIX])
{
    if constexpr (RADIX == 1) {
    } else if constexpr (RADIX == 2) {
        dft2<SIGN>(v[0], v[1]);
    } else if constexpr (RADIX == 4) {
        dft4<SIGN>(v[0], v[1], v[2], v[3]);
    } else if constexpr (RADIX == 8) {
        dft8<SIGN>(v);
    } else {
        static_assert(RADIX == 16, "radix must be 1, 2, 4, 8 or 16");
        dft16<SIGN>(v);
    }
}

}  // namespace tope
