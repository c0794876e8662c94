// Real-input transforms (R2C / C2R) on top of the complex axis passes.  No
// reference counterpart: the reference declares R2C (topefft.h:9) and stores
// the plan type (src/fft1d.c:310, src/fft2d.c:329) but always runs C2C.
//
// A real row x of even length n = 2m is read as m complex values
// z[k] = x[2k] + i x[2k+1] (the same bytes) and transformed by the m-point
// complex pass; the half spectrum follows from Z = DFT_m(z):
//   E[k] = (Z[k] + conj Z[m-k]) / 2,  O[k] = (Z[k] - conj Z[m-k]) / (2i),
//   X[k] = E[k] + W_n^k O[k],  k = 0..m  (Z indices mod m).
// The inverse recovers E, O from X and its mirror,
//   E[k] = (X[k] + conj X[m-k]) / 2,  O[k] = (X[k] - conj X[m-k]) W_n^-k / 2,
// forms Z[k] = E[k] + i O[k] (k < m) and the inverse m-point complex pass
// returns z, i.e. the even and odd real samples.  W_n^k (k <= m) comes from a
// table with exactly reduced phases, generated in FP64 (rounded once for
// complex64 storage).
//
// Packed DC/Nyquist columns (`pack`, complex64 plans with column passes):
// X[0] and X[m] of every row are real, so the row step stores X[0] + i X[m]
// in column 0 and leaves column m alone; the column passes then transform
// an even number of columns, m (the paired-column tiles need whole pairs),
// and real_pack_kernel splits the two Hermitian spectra F0, Fm out of
// FZ = F0 + i Fm afterwards:
//   F0[K] = (FZ[K] + conj FZ[-K]) / 2,  Fm[K] = (FZ[K] - conj FZ[-K]) / (2i),
// K = (ky, kz) and -K = (-ky mod ny, -kz mod nz).  C2R packs first,
// FZ = F0 + i Fm; after the inverse column passes column 0 holds
// x0 + i xm, both real, and the row step reads X[0] = x0, X[m] = xm from it.
#include "internal.hpp"

namespace tope {

// One thread per mirror pair (k, m - k), k = 0..m/2, of one row: forward
// reads Z (row pitch m) and writes X (m + 1 values per row, row pitch p).
template <typename R>
__global__ void r2c_post_kernel(const cpx<R> *__restrict__ Z, cpx<R> *__restrict__ X,
                                const cpx<R> *__restrict__ w, long long m, long long p, long long rows, int pack)
{
    const long long pairs = m / 2 + 1;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= pairs * rows) return;
    const long long row = idx / pairs, k = idx % pairs;
    const cpx<R> *z = Z + row * m;
    cpx<R> *x = X + row * p;
    const cpx<R> zk = z[k % m], zm = z[(m - k) % m];
    if (pack && k == 0) {  // X[0] + i X[m] = (Re Z0 + Im Z0) + i (Re Z0 - Im Z0)
        x[0] = cpx<R>{zk.x + zk.y, zk.x - zk.y};
        return;
    }
    // X[k] from a = Z[k], b = conj Z[m-k]; X[m-k] from a' = Z[m-k], b' = conj Z[k]
    auto out = [&](cpx<R> a, cpx<R> bc, long long kk) {
        const cpx<R> b{bc.x, -bc.y};
        const cpx<R> e{R(0.5) * (a.x + b.x), R(0.5) * (a.y + b.y)};
        // O = (a - b) / (2i) = (Im(a-b), -Re(a-b)) / 2
        const cpx<R> o{R(0.5) * (a.y - b.y), R(-0.5) * (a.x - b.x)};
        const cpx<R> wo = cmul(w[kk], o);
        x[kk] = cpx<R>{e.x + wo.x, e.y + wo.y};
    };
    out(zk, zm, k);
    if (m - k != k) out(zm, zk, m - k);
}

// Inverse: X (row pitch p) -> Z (row pitch m), scaled by s; one thread per pair.
template <typename R>
__global__ void c2r_pre_kernel(const cpx<R> *__restrict__ X, cpx<R> *__restrict__ Z,
                               const cpx<R> *__restrict__ w, long long m, long long p, long long rows, R s, int pack)
{
    const long long pairs = m / 2 + 1;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= pairs * rows) return;
    const long long row = idx / pairs, k = idx % pairs;
    const cpx<R> *x = X + row * p;
    cpx<R> *z = Z + row * m;
    cpx<R> xk = x[k], xm = x[m - k];
    if (pack && k == 0) {  // column 0 holds X[0] + i X[m]
        xm = cpx<R>{xk.y, R(0)};
        xk = cpx<R>{xk.x, R(0)};
    }
    auto out = [&](cpx<R> a, cpx<R> bc, long long kk) {
        const cpx<R> b{bc.x, -bc.y};
        const cpx<R> e{R(0.5) * (a.x + b.x), R(0.5) * (a.y + b.y)};
        const cpx<R> d{R(0.5) * (a.x - b.x), R(0.5) * (a.y - b.y)};
        const cpx<R> o = cmulc(d, w[kk]);  // d * conj(W^k)
        // Z = E + i O
        z[kk] = cpx<R>{s * (e.x - o.y), s * (e.y + o.x)};
    };
    if (k < m) out(xk, xm, k);
    if (m - k != k && m - k < m) out(xm, xk, m - k);
}

// Fused rows: the m-point complex row transform and the half-spectrum step
// in one launch (no intermediate array).  TB rows per block, P = N/E threads
// per row, thread p owning elements p + P e as in the row axis passes.
//   R2C (SIGN = -1): load z (the real row read as m complex values), run
//   the radix stages, exchange Z through LDS so each thread also holds its
//   mirrors Z[m - k], form X[k] = E[k] + W^k O[k] for its k and X[m] (thread
//   0), store the row of m + 1 values at `pitch`.
//   C2R (SIGN = +1): load X[k] and X[m - k] for its k, form
//   Z[k] = s (E[k] + i O[k]), run the inverse radix stages, store z.
template <typename R, int N, int E, int TB, int SIGN>
__global__ void __launch_bounds__(TB *(N / E))
real_rows_kernel(const cpx<R> *__restrict__ in, cpx<R> *__restrict__ out,
                 const cpx<R> *__restrict__ stw, const cpx<R> *__restrict__ w, long long pitch,
                 long long rows, R s, int pack)
{
    using SC = Sched<N, E>;
    constexpr int P = SC::P;
    constexpr long long m = N;
    __shared__ __align__(16) unsigned char lds[Exchange<R, N, E, TB, false, true>::BYTES];
    const int tid = threadIdx.x;
    const int t = tid / P, p = tid % P;
    const long long row = (long long)blockIdx.x * TB + t;
    const bool valid = row < rows;
    cpx<R> v[E];
    if constexpr (SIGN < 0) {
        const cpx<R> *z = in + row * m;
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = valid ? z[p + P * e] : cpx<R>{0, 0};
    } else {
        const cpx<R> *X = in + row * pitch;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const long long k = p + P * e;
            cpx<R> xk{0, 0}, xm{0, 0};
            if (valid) {
                xk = X[k];
                xm = X[m - k];
                if (pack && k == 0) {  // column 0 holds X[0] + i X[m]
                    xm = cpx<R>{xk.y, R(0)};
                    xk = cpx<R>{xk.x, R(0)};
                }
            }
            const cpx<R> b{xm.x, -xm.y};
            const cpx<R> ev{R(0.5) * (xk.x + b.x), R(0.5) * (xk.y + b.y)};
            const cpx<R> d{R(0.5) * (xk.x - b.x), R(0.5) * (xk.y - b.y)};
            const cpx<R> o = cmulc(d, w[k]);
            v[e] = cpx<R>{s * (ev.x - o.y), s * (ev.y + o.x)};
        }
    }
    run_stages<R, N, E, TB, false, true, SIGN, 0>(v, p, t, stw, lds);
    if constexpr (SIGN > 0) {
        if (!valid) return;
        cpx<R> *z = out + row * m;
#pragma unroll
        for (int e = 0; e < E; ++e) z[p + P * e] = v[e];
    } else {
        // mirrors Z[(m - k) mod m] through LDS, one component at a time
        R *l = reinterpret_cast<R *>(lds);
        const int base = lds_index<N, E, TB, false>(t, 0);
        cpx<R> mir[E];
#pragma unroll
        for (int comp = 0; comp < 2; ++comp) {
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int k = p + P * e;
                l[base + k + (k >> 4)] = comp ? v[e].y : v[e].x;
            }
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int k = (int)((m - (p + P * e)) & (m - 1));
                const R val = l[base + k + (k >> 4)];
                if (comp) mir[e].y = val;
                else mir[e].x = val;
            }
        }
        if (!valid) return;
        cpx<R> *X = out + row * pitch;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const long long k = p + P * e;
            const cpx<R> a = v[e], b{mir[e].x, -mir[e].y};
            const cpx<R> ev{R(0.5) * (a.x + b.x), R(0.5) * (a.y + b.y)};
            const cpx<R> o{R(0.5) * (a.y - b.y), R(-0.5) * (a.x - b.x)};
            const cpx<R> wo = cmul(w[k], o);
            X[k] = cpx<R>{ev.x + wo.x, ev.y + wo.y};
        }
        // k = m: E[0] - O[0] = Re Z0 - Im Z0 (packed: X[0] + i X[m] in column 0)
        if (p == 0) {
            if (pack) X[0] = cpx<R>{v[0].x + v[0].y, v[0].x - v[0].y};
            else X[m] = cpx<R>{v[0].x - v[0].y, R(0)};
        }
    }
}

template <typename R, int N, int E, int TB>
static hipError_t launch_rows_t(int dir, const void *in, void *out, const void *stw, const void *w,
                                long long pitch, long long rows, double s, int pack, hipStream_t st)
{
    const long long grid = (rows + TB - 1) / TB;
    if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
    const cpx<R> *i = static_cast<const cpx<R> *>(in);
    cpx<R> *o = static_cast<cpx<R> *>(out);
    const cpx<R> *t = static_cast<const cpx<R> *>(stw);
    const cpx<R> *ww = static_cast<const cpx<R> *>(w);
    if (dir == 0)
        hipLaunchKernelGGL((real_rows_kernel<R, N, E, TB, -1>), dim3((unsigned)grid), dim3(TB * (N / E)), 0, st, i, o, t,
                           ww, pitch, rows, (R)s, pack);
    else
        hipLaunchKernelGGL((real_rows_kernel<R, N, E, TB, 1>), dim3((unsigned)grid), dim3(TB * (N / E)), 0, st, i, o, t,
                           ww, pitch, rows, (R)s, pack);
    return hipGetLastError();
}

bool real_rows_supported(long long m) { return m >= 64 && m <= 4096 && (m & (m - 1)) == 0; }

template <typename R>
static hipError_t launch_real_rows_t(int dir, const void *in, void *out, const void *stw, const void *w, long long m,
                                     long long pitch, long long rows, double s, int pack, hipStream_t st)
{
    switch (m) {
    case 64: return launch_rows_t<R, 64, 16, 16>(dir, in, out, stw, w, pitch, rows, s, pack, st);
    case 128: return launch_rows_t<R, 128, 16, 16>(dir, in, out, stw, w, pitch, rows, s, pack, st);
    case 256: return launch_rows_t<R, 256, 16, 16>(dir, in, out, stw, w, pitch, rows, s, pack, st);
    case 512: return launch_rows_t<R, 512, 16, 8>(dir, in, out, stw, w, pitch, rows, s, pack, st);
    case 1024: return launch_rows_t<R, 1024, 16, 4>(dir, in, out, stw, w, pitch, rows, s, pack, st);
    case 2048: return launch_rows_t<R, 2048, 16, 2>(dir, in, out, stw, w, pitch, rows, s, pack, st);
    case 4096: return launch_rows_t<R, 4096, 16, 1>(dir, in, out, stw, w, pitch, rows, s, pack, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_real_rows(int prec, int dir, const void *in, void *out, const void *stw, const void *w, long long m,
                            long long pitch, long long rows, double s, int pack, hipStream_t st)
{
    return prec == PREC_F64 ? launch_real_rows_t<double>(dir, in, out, stw, w, m, pitch, rows, s, pack, st)
                            : launch_real_rows_t<float>(dir, in, out, stw, w, m, pitch, rows, s, pack, st);
}

template <typename R>
static void launch_real_pass_t(int dir, const void *in, void *out, const void *w, long long m, long long pitch,
                               long long rows, double s, int pack, long long grid, hipStream_t st)
{
    const cpx<R> *i = static_cast<const cpx<R> *>(in);
    cpx<R> *o = static_cast<cpx<R> *>(out);
    const cpx<R> *ww = static_cast<const cpx<R> *>(w);
    if (dir == 0)
        hipLaunchKernelGGL(r2c_post_kernel<R>, dim3((unsigned)grid), dim3(256), 0, st, i, o, ww, m, pitch, rows, pack);
    else
        hipLaunchKernelGGL(c2r_pre_kernel<R>, dim3((unsigned)grid), dim3(256), 0, st, i, o, ww, m, pitch, rows, (R)s,
                           pack);
}

hipError_t launch_real_pass(int prec, int dir, const void *in, void *out, const void *w, long long m, long long pitch,
                            long long rows, double s, int pack, hipStream_t st)
{
    const long long threads = (m / 2 + 1) * rows;
    const long long grid = (threads + 255) / 256;
    if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
    if (prec == PREC_F64) launch_real_pass_t<double>(dir, in, out, w, m, pitch, rows, s, pack, grid, st);
    else launch_real_pass_t<float>(dir, in, out, w, m, pitch, rows, s, pack, grid, st);
    return hipGetLastError();
}

// dir 0 (after the forward column passes): column 0 of every row K holds
// FZ[K]; one thread per pair {K, -K} writes F0 to column 0 and Fm to column
// m of both rows.  dir 1 (before the inverse column passes): one thread per
// row packs FZ = F0 + i Fm into column 0.  Rows K = ky + ny kz, pitch p.
template <typename R>
__global__ void real_pack_kernel(cpx<R> *__restrict__ X, long long m, long long p, long long ny, long long nz,
                                 int dir)
{
    const long long rows = ny * nz;
    const long long K = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (K >= rows) return;
    if (dir != 0) {
        const cpx<R> f0 = X[K * p], fm = X[K * p + m];
        X[K * p] = cpx<R>{f0.x - fm.y, f0.y + fm.x};
        return;
    }
    const long long ky = K % ny, kz = K / ny;
    const long long J = (ky ? ny - ky : 0) + ny * (kz ? nz - kz : 0);  // -K
    if (J < K) return;  // the pair's lower row does both
    const cpx<R> a = X[K * p], b = X[J * p];
    // F0[K] = (a + conj b) / 2, Fm[K] = (a - conj b) / (2i); rows swap for J
    X[K * p] = cpx<R>{R(0.5) * (a.x + b.x), R(0.5) * (a.y - b.y)};
    X[K * p + m] = cpx<R>{R(0.5) * (a.y + b.y), R(-0.5) * (a.x - b.x)};
    if (J != K) {
        X[J * p] = cpx<R>{R(0.5) * (b.x + a.x), R(0.5) * (b.y - a.y)};
        X[J * p + m] = cpx<R>{R(0.5) * (b.y + a.y), R(-0.5) * (b.x - a.x)};
    }
}

hipError_t launch_real_pack(int prec, int dir, void *X, long long m, long long pitch, long long ny, long long nz,
                            hipStream_t st)
{
    const long long grid = (ny * nz + 255) / 256;
    if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
    if (prec == PREC_F64)
        hipLaunchKernelGGL(real_pack_kernel<double>, dim3((unsigned)grid), dim3(256), 0, st, static_cast<cpx<double> *>(X),
                           m, pitch, ny, nz, dir);
    else
        hipLaunchKernelGGL(real_pack_kernel<float>, dim3((unsigned)grid), dim3(256), 0, st, static_cast<cpx<float> *>(X),
                           m, pitch, ny, nz, dir);
    return hipGetLastError();
}

}  // namespace tope
