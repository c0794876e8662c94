// Mixed-radix axis pass for lengths n = 2^a 3^b 5^c 7^d <= 4096 (FP64, and
// complex64 storage with FP32 arithmetic and the table rounded once).
// The reference runs such 1D sizes as an O(n^2) DFT (src/fft1d.c:148-232,
// src/kernels1D.cl:549-573) and rejects them in 2D/3D (src/fft2d.c:355-362,
// src/fft3d.c:522-533).
//
// Self-sorting Stockham, decimation in time, with a runtime radix list
// (radix 16/8/4/2 for the power-of-two part, then 7, 5, 3): stage s of
// radix r at span Ns treats butterfly j < n/r, reads j + m n/r (m < r),
// multiplies element m by W_{Ns r}^{(j mod Ns) m} and writes output m of its
// r-point DFT to (j / Ns) Ns r + (j mod Ns) + m Ns; starting at Ns = 1 the
// result is in natural order.  A block owns whole lines: its first stage
// reads them from memory, the stages in between run between two LDS
// buffers (a barrier per stage), and its last stage writes them back, so
// the pass is in place.  Every twiddle is an entry of
// the plan's table W_n^i (exactly reduced phases, forward sign) and the
// radix-3/5/7 constants are correctly rounded literals, so no device sin/cos
// is evaluated.
#include <algorithm>

#include "internal.hpp"

namespace tope {


// Integer division by a runtime divisor without the emulated divide: a
// float reciprocal estimate, exact to +-1 for the operands here (< 2^24),
// then one correction.
struct Div32 {
    int d;
    float r;
};
__device__ __forceinline__ Div32 mkdiv(int d) { return Div32{d, 1.0f / (float)d}; }
__device__ __forceinline__ int divmod(int a, const Div32 &d, int &rem)
{
    int q = (int)((float)a * d.r);
    int r = a - q * d.d;
    if (r >= d.d) {
        ++q;
        r -= d.d;
    } else if (r < 0) {
        --q;
        r += d.d;
    }
    rem = r;
    return q;
}
// The same for 64-bit line numbers (< 2^40) through a double reciprocal.
struct Div64 {
    long long d;
    double r;
};
__device__ __forceinline__ long long divmod64(long long a, const Div64 &d, long long &rem)
{
    long long q = (long long)((double)a * d.r);
    long long r = a - q * d.d;
    if (r >= d.d) {
        ++q;
        r -= d.d;
    } else if (r < 0) {
        --q;
        r += d.d;
    }
    rem = r;
    return q;
}

// The block's copy of the roots in LDS, in two levels: W_n^i = hi[i >> 6]
// lo[i & 63] with lo[j] = W_n^j (j < 64) and hi[j] = W_n^{64 j}; 128 entries
// whatever n <= 4096 (the per-butterfly reads of the global W_n table cost
// 20 % of a 3000-point pass).
constexpr int ROOT_LO = 64;
constexpr int ROOT_SLOTS = ROOT_LO + MIXED_MAX_N / ROOT_LO;
template <int SIGN, typename R>
__device__ __forceinline__ cpx<R> root(const cpx<R> *__restrict__ w, int i)
{
    const cpx<R> v = cmul(w[ROOT_LO + (i >> 6)], w[i & (ROOT_LO - 1)]);
    return SIGN < 0 ? v : cpx<R>{v.x, -v.y};
}

// cos(2 pi m / r) and sin(2 pi m / r), m < r, for the odd radices: FP64
// values correctly rounded from 50-digit evaluations (complex64 rounds them
// once more).
template <int r>
__device__ __forceinline__ double odd_cos(int m)
{
    constexpr double c3[3] = {0x1p+0, -0x1p-1, -0x1p-1};
    constexpr double c5[5] = {0x1p+0, 0x1.3c6ef372fe950p-2, -0x1.9e3779b97f4a8p-1, -0x1.9e3779b97f4a8p-1,
                              0x1.3c6ef372fe950p-2};
    constexpr double c7[7] = {0x1p+0,
                              0x1.3f3a0e28bedd1p-1,
                              -0x1.c7b90e3024582p-3,
                              -0x1.cd4bca9cb5c71p-1,
                              -0x1.cd4bca9cb5c71p-1,
                              -0x1.c7b90e3024582p-3,
                              0x1.3f3a0e28bedd1p-1};
    return r == 3 ? c3[m] : r == 5 ? c5[m] : c7[m];
}
template <int r>
__device__ __forceinline__ double odd_sin(int m)
{
    constexpr double s3[3] = {0.0, 0x1.bb67ae8584caap-1, -0x1.bb67ae8584caap-1};
    constexpr double s5[5] = {0.0, 0x1.e6f0e134454ffp-1, 0x1.2cf2304755a5ep-1, -0x1.2cf2304755a5ep-1,
                              -0x1.e6f0e134454ffp-1};
    constexpr double s7[7] = {0.0,
                              0x1.904c37505de4bp-1,
                              0x1.f329c0558e969p-1,
                              0x1.bc4c04d71abc1p-2,
                              -0x1.bc4c04d71abc1p-2,
                              -0x1.f329c0558e969p-1,
                              -0x1.904c37505de4bp-1};
    return r == 3 ? s3[m] : r == 5 ? s5[m] : s7[m];
}

// Odd prime radix r by the conjugate-pair form: with a_j = x_j + x_{r-j},
// b_j = x_j - x_{r-j} and W_r^{jk} = c + i s,
//   y_k = x_0 + sum_{j=1}^{(r-1)/2} (a_j c + i b_j s).
template <int r, int SIGN, typename R>
__device__ __forceinline__ void dft_odd(cpx<R> (&x)[r])
{
    constexpr int h = (r - 1) / 2;
    cpx<R> a[h], b[h], wr[r];
#pragma unroll
    for (int j = 1; j <= h; ++j) {
        a[j - 1] = x[j] + x[r - j];
        b[j - 1] = x[j] - x[r - j];
    }
    // W_r^m = cos - i sin forward, its conjugate inverse
#pragma unroll
    for (int m = 1; m < r; ++m)
        wr[m] = cpx<R>{(R)odd_cos<r>(m), (R)(SIGN < 0 ? -odd_sin<r>(m) : odd_sin<r>(m))};
    cpx<R> y[r];
    y[0] = x[0];
#pragma unroll
    for (int j = 0; j < h; ++j) y[0] = y[0] + a[j];
#pragma unroll
    for (int k = 1; k < r; ++k) {
        cpx<R> acc = x[0];
#pragma unroll
        for (int j = 1; j <= h; ++j) {
            const cpx<R> c = wr[(j * k) % r];
            acc.x += a[j - 1].x * c.x - b[j - 1].y * c.y;
            acc.y += a[j - 1].y * c.x + b[j - 1].x * c.y;
        }
        y[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < r; ++k) x[k] = y[k];
}

// LDS slot of element i of the block: one pad slot per 16 elements breaks
// the power-of-two strides of the stage scatter (r, Ns) across the banks.
__device__ __forceinline__ int lslot(int i) { return i + (i >> 4); }

// Rows of a contiguous axis: where the first stage reads and the last
// stage writes, straight from the input lines and to the output lines,
// scaled (consecutive lanes hold consecutive elements).
template <typename R>
struct RowOut {
    const cpx<R> *in;
    cpx<R> *out;
    long long l0, lines, os, cs, is;
    Div64 dC;
    Div32 dtb;
    R scale;
    long long oos, ocs, ois;  // output geometry
    // four-step post-twiddle (pt_n != 0)
    const cpx<R> *pt_lo, *pt_hi;
    long long pt_n, pt_o, pt_c;
    Div32 dsplit;
};

// One stage of compile-time radix r over the block's lines, LDS to LDS,
// (FROMGLOBAL) the input lines to LDS, or (TOGLOBAL) LDS to the output
// lines.  CF (strided axes, global stages only): neighbouring lanes take
// neighbouring lines, so a global access of a lane group covers adjacent
// columns.
template <typename R, int r, int SIGN, int NT, bool TOGLOBAL, bool FROMGLOBAL = false, bool CF = false>
__device__ __forceinline__ void mixed_stage(const cpx<R> *src, cpx<R> *dst, int n, int tb, int Ns,
                                            const cpx<R> *__restrict__ roots, const RowOut<R> &go)
{
    const int nb = n / r;  // butterflies per line
    const int tws = n / (Ns * r);  // table indices stay below n <= 4096: 32-bit
    const Div32 dnb = mkdiv(nb), dns = mkdiv(Ns);
    for (int bi = threadIdx.x; bi < tb * nb; bi += NT) {
        int j, kk, t;
        if constexpr (CF) j = divmod(bi, go.dtb, t);
        else t = divmod(bi, dnb, j);
        const int jn = divmod(j, dns, kk);  // j = jn Ns + kk
        cpx<R> x[r];
        if constexpr (FROMGLOBAL) {
            const long long l = go.l0 + t;
            if (l < go.lines) {
                long long c;
                const long long o = divmod64(l, go.dC, c);
                const cpx<R> *row = go.in + o * go.os + c * go.cs;
#pragma unroll
                for (int m = 0; m < r; ++m) x[m] = CF ? row[(long long)(j + m * nb) * go.is] : row[j + m * nb];
            } else {
#pragma unroll
                for (int m = 0; m < r; ++m) x[m] = cpx<R>{0, 0};
            }
        } else {
            const int si = t * n + j;
#pragma unroll
            for (int m = 0; m < r; ++m) x[m] = src[lslot(si + m * nb)];
        }
        if (Ns > 1) {
#pragma unroll
            for (int m = 1; m < r; ++m) x[m] = cmul(x[m], root<SIGN>(roots, kk * m * tws));
        }
        if constexpr (r == 3 || r == 5 || r == 7) dft_odd<r, SIGN>(x);
        else dft<r, SIGN>(x);
        if constexpr (TOGLOBAL) {
            const long long l = go.l0 + t;
            if (l < go.lines) {
                long long c;
                const long long o = divmod64(l, go.dC, c);
                cpx<R> *row = go.out + o * go.oos + c * go.ocs;
                const int k0 = jn * Ns * r + kk;
                // four-step post-twiddle base: the exponent (base k) < pt_n
                const int base = go.pt_n ? (int)(o * go.pt_o + c * go.pt_c) : 0;
#pragma unroll
                for (int m = 0; m < r; ++m) {
                    cpx<R> v{x[m].x * go.scale, x[m].y * go.scale};
                    if (go.pt_n) {
                        int lo;
                        const int hi = divmod(base * (k0 + m * Ns), go.dsplit, lo);
                        const cpx<R> w = cmul(go.pt_lo[lo], go.pt_hi[hi]);
                        v = SIGN < 0 ? cmul(v, w) : cmulc(v, w);
                    }
                    row[(long long)(k0 + m * Ns) * go.ois] = v;
                }
            }
        } else {
            const int di = t * n + jn * Ns * r + kk;
#pragma unroll
            for (int m = 0; m < r; ++m) dst[lslot(di + m * Ns)] = x[m];
        }
    }
}

template <typename R, int SIGN, int NT, bool TOGLOBAL, bool FROMGLOBAL = false, bool CF = false>
__device__ __forceinline__ void mixed_stage_r(int r, const cpx<R> *src, cpx<R> *dst, int n, int tb, int Ns,
                                              const cpx<R> *__restrict__ roots, const RowOut<R> &go)
{
    switch (r) {
    case 2: mixed_stage<R, 2, SIGN, NT, TOGLOBAL, FROMGLOBAL, CF>(src, dst, n, tb, Ns, roots, go); break;
    case 3: mixed_stage<R, 3, SIGN, NT, TOGLOBAL, FROMGLOBAL, CF>(src, dst, n, tb, Ns, roots, go); break;
    case 4: mixed_stage<R, 4, SIGN, NT, TOGLOBAL, FROMGLOBAL, CF>(src, dst, n, tb, Ns, roots, go); break;
    case 5: mixed_stage<R, 5, SIGN, NT, TOGLOBAL, FROMGLOBAL, CF>(src, dst, n, tb, Ns, roots, go); break;
    case 7: mixed_stage<R, 7, SIGN, NT, TOGLOBAL, FROMGLOBAL, CF>(src, dst, n, tb, Ns, roots, go); break;
    case 8: mixed_stage<R, 8, SIGN, NT, TOGLOBAL, FROMGLOBAL, CF>(src, dst, n, tb, Ns, roots, go); break;
    default: mixed_stage<R, 16, SIGN, NT, TOGLOBAL, FROMGLOBAL, CF>(src, dst, n, tb, Ns, roots, go); break;
    }
}

// CAP: elements of one LDS buffer (two buffers of CAP + CAP/16 slots); NT
// threads: 512 for CAP = 2048 (two blocks, 16 waves per CU: 4 elements per
// thread keep the latency of the stage loop covered; 256 threads measured
// 10-12 % slower at 100^3 and 125^3), 1024 for the single-line 4096-element
// blocks that own a CU.  Four waves per SIMD: the FP64 kernel needs 162
// VGPRs on its own (one 512-thread block per CU); capped at 128 (with
// spills) two blocks share a CU, 9-21 % faster (1000^2 44.0 -> 37.1 us,
// 3000^2 360 -> 328 us with the LDS roots, 1D 10^6 46.5 -> 36.7 us).
template <typename R, int SIGN, int CAP, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4)))
mixed_kernel(const cpx<R> *in, cpx<R> *out, const MixedArgs a)
{
    constexpr int SLOTS = CAP + CAP / 16;
    __shared__ cpx<R> lds[2 * SLOTS + ROOT_SLOTS];
    const int n = a.n, tb = a.tb;
    // the two-level roots (read from the second stage on, after a barrier)
    cpx<R> *roots = lds + 2 * SLOTS;
    {
        const cpx<R> *g = static_cast<const cpx<R> *>(a.roots);
        const int nhi = (n + ROOT_LO - 1) / ROOT_LO;
        for (int i = threadIdx.x; i < ROOT_LO + nhi; i += NT) {
            const int e = i < ROOT_LO ? i : (i - ROOT_LO) * ROOT_LO;
            roots[i] = e < n ? g[e] : cpx<R>{1, 0};
        }
    }
    const int span_elems = tb * n;
    const bool cfast = a.is != 1;  // strided axis: neighbouring lanes take neighbouring lines
    const bool cfast_out = a.ois != 1;  // the same for the last stage's stores
    // strided lines share 128-byte segments with the neighbouring blocks'
    // lines: keep neighbours on one XCD (one L2)
    const long long l0 = xcd_block(cfast ? 1 : 0) * tb;

    // element idx of the block -> (line t, index k), and its global offset
    const Div32 dtb = mkdiv(tb), dn = mkdiv(n);
    const Div64 dC{a.C, 1.0 / (double)a.C};
    auto split = [&](int idx, int &t, int &k) {
        if (cfast) k = divmod(idx, dtb, t);
        else t = divmod(idx, dn, k);
    };
    auto gaddr = [&](long long l, int k) {
        long long c;
        const long long o = divmod64(l, dC, c);
        return o * a.os + c * a.cs + (long long)k * a.is;
    };

    // the first stage of a pass with two or more stages reads the input
    // lines itself (no initial LDS fill and barrier) and the last stage
    // stores straight to the output (no final LDS round trip); on strided
    // axes those two stages put neighbouring lines on neighbouring lanes
    const RowOut<R> go{in,
                       out,
                       l0,
                       a.lines,
                       a.os,
                       a.cs,
                       a.is,
                       dC,
                       dtb,
                       (R)a.scale,
                       a.oos,
                       a.ocs,
                       a.ois,
                       static_cast<const cpx<R> *>(a.pt_lo),
                       static_cast<const cpx<R> *>(a.pt_hi),
                       a.pt_n,
                       a.pt_o,
                       a.pt_c,
                       mkdiv(a.pt_split > 0 ? a.pt_split : 1)};
    const bool direct_first = a.nst >= 2;
    cpx<R> *src = lds, *dst = lds + SLOTS;
    int Ns = 1;
    int s = 0;
    if (direct_first) {
        if (cfast) mixed_stage_r<R, SIGN, NT, false, true, true>(a.radix[0], src, dst, n, tb, Ns, roots, go);
        else mixed_stage_r<R, SIGN, NT, false, true>(a.radix[0], src, dst, n, tb, Ns, roots, go);
        __syncthreads();
        src = lds + SLOTS;
        dst = lds;
        Ns = a.radix[0];
        s = 1;
    } else {
        for (int idx = threadIdx.x; idx < span_elems; idx += NT) {
            int t, k;
            split(idx, t, k);
            const long long l = l0 + t;
            lds[lslot(t * n + k)] = l < a.lines ? in[gaddr(l, k)] : cpx<R>{0, 0};
        }
        __syncthreads();
    }
    for (; s < a.nst; ++s) {
        const int r = a.radix[s];
        if (s == a.nst - 1) {
            if (cfast_out) mixed_stage_r<R, SIGN, NT, true, false, true>(r, src, dst, n, tb, Ns, roots, go);
            else mixed_stage_r<R, SIGN, NT, true>(r, src, dst, n, tb, Ns, roots, go);
            return;
        }
        mixed_stage_r<R, SIGN, NT, false>(r, src, dst, n, tb, Ns, roots, go);
        __syncthreads();
        cpx<R> *tmp = src;
        src = dst;
        dst = tmp;
        Ns *= r;
    }
}

bool mixed_supported(long long n)
{
    if (n < 2 || n > MIXED_MAX_N) return false;
    for (long long p : {2, 3, 5, 7})
        while (n % p == 0) n /= p;
    return n == 1;
}

MixedArgs mixed_plan(long long n, long long O, long long os, long long C, long long cs, long long is)
{
    MixedArgs a{};
    a.n = (int)n;
    a.O = O;
    a.os = os;
    a.C = C;
    a.cs = cs;
    a.is = is;
    a.lines = O * C;
    long long m = n;
    int ns = 0;
    while (m % 16 == 0) a.radix[ns++] = 16, m /= 16;
    if (m % 8 == 0) a.radix[ns++] = 8, m /= 8;
    if (m % 4 == 0) a.radix[ns++] = 4, m /= 4;
    if (m % 2 == 0) a.radix[ns++] = 2, m /= 2;
    for (int p : {7, 5, 3})
        while (m % p == 0) a.radix[ns++] = p, m /= p;
    a.nst = ns;
    // lines per block: as many as fit 2048 elements (two 32 KiB LDS buffers,
    // two blocks per CU), one line of up to 4096 otherwise
    long long tb = std::max<long long>(1, 2048 / n);
    // strided lines of 1025..2048 would run one per block (16-byte column
    // segments): two or three per 4096-element block instead (2048 x 2000:
    // 89 -> 81.5 us; wider tiles for shorter lines measured slower)
    if (is != 1 && n > 1024 && n <= 2048) tb = MIXED_MAX_N / n;
    if (tb > a.lines) tb = a.lines;
    a.tb = (int)tb;
    a.scale = 1.0;
    a.oos = os;
    a.ocs = cs;
    a.ois = is;
    return a;
}

bool mixed4_split(long long n, long long *n1)
{
    if (n <= MIXED_MAX_N || n > (long long)MIXED_MAX_N * MIXED_MAX_N) return false;
    long long m = n;
    for (long long p : {2, 3, 5, 7})
        while (m % p == 0) m /= p;
    if (m != 1) return false;
    // the smallest divisor d >= sqrt(n) with d, n / d <= MIXED_MAX_N
    for (long long d = 1; d <= MIXED_MAX_N; ++d) {
        if (n % d || d * d < n || n / d > MIXED_MAX_N) continue;
        *n1 = d;
        return true;
    }
    return false;
}

unsigned mixed_grid(const MixedArgs &a) { return (unsigned)((a.lines + a.tb - 1) / a.tb); }

template <typename R>
void launch_mixed_t(const void *in, void *out, const MixedArgs &a, int sign, const dim3 g, hipStream_t st)
{
    const cpx<R> *i = static_cast<const cpx<R> *>(in);
    cpx<R> *o = static_cast<cpx<R> *>(out);
    if ((long long)a.tb * a.n <= 2048) {
        if (sign < 0) hipLaunchKernelGGL((mixed_kernel<R, -1, 2048, 512>), g, dim3(512), 0, st, i, o, a);
        else hipLaunchKernelGGL((mixed_kernel<R, 1, 2048, 512>), g, dim3(512), 0, st, i, o, a);
    } else {
        if (sign < 0) hipLaunchKernelGGL((mixed_kernel<R, -1, MIXED_MAX_N, 1024>), g, dim3(1024), 0, st, i, o, a);
        else hipLaunchKernelGGL((mixed_kernel<R, 1, MIXED_MAX_N, 1024>), g, dim3(1024), 0, st, i, o, a);
    }
}

hipError_t launch_mixed(int prec, const void *in, void *out, const MixedArgs &a, int sign, hipStream_t st)
{
    const long long grid = (a.lines + a.tb - 1) / a.tb;
    if (grid > 0x7fffffffLL || a.nst < 1 || a.nst > 16 || a.tb < 1 || (long long)a.tb * a.n > MIXED_MAX_N)
        return hipErrorInvalidValue;
    const dim3 g((unsigned)grid);
    if (prec == PREC_F64) launch_mixed_t<double>(in, out, a, sign, g, st);
    else launch_mixed_t<float>(in, out, a, sign, g, st);
    return hipGetLastError();
}

}  // namespace tope
