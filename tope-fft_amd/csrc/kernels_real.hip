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
// returns z, i.e. the even and odd real samples.  W_n^k (k <= m) comes from an
// FP64 table with exactly reduced phases.
#include "internal.hpp"

namespace tope {

// One thread per mirror pair (k, m - k), k = 0..m/2, of one row: forward
// reads Z (row pitch m) and writes X (m + 1 values per row, row pitch p).
__global__ void r2c_post_kernel(const cpx<double> *__restrict__ Z, cpx<double> *__restrict__ X,
                                const cpx<double> *__restrict__ w, long long m, long long p, long long rows)
{
    const long long pairs = m / 2 + 1;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= pairs * rows) return;
    const long long row = idx / pairs, k = idx % pairs;
    const cpx<double> *z = Z + row * m;
    cpx<double> *x = X + row * p;
    const cpx<double> zk = z[k % m], zm = z[(m - k) % m];
    // X[k] from a = Z[k], b = conj Z[m-k]; X[m-k] from a' = Z[m-k], b' = conj Z[k]
    auto out = [&](cpx<double> a, cpx<double> bc, long long kk) {
        const cpx<double> b{bc.x, -bc.y};
        const cpx<double> e{0.5 * (a.x + b.x), 0.5 * (a.y + b.y)};
        // O = (a - b) / (2i) = (Im(a-b), -Re(a-b)) / 2
        const cpx<double> o{0.5 * (a.y - b.y), -0.5 * (a.x - b.x)};
        const cpx<double> wo = cmul(w[kk], o);
        x[kk] = cpx<double>{e.x + wo.x, e.y + wo.y};
    };
    out(zk, zm, k);
    if (m - k != k) out(zm, zk, m - k);
}

// Inverse: X (row pitch p) -> Z (row pitch m), scaled by s; one thread per pair.
__global__ void c2r_pre_kernel(const cpx<double> *__restrict__ X, cpx<double> *__restrict__ Z,
                               const cpx<double> *__restrict__ w, long long m, long long p, long long rows, double s)
{
    const long long pairs = m / 2 + 1;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= pairs * rows) return;
    const long long row = idx / pairs, k = idx % pairs;
    const cpx<double> *x = X + row * p;
    cpx<double> *z = Z + row * m;
    const cpx<double> xk = x[k], xm = x[m - k];
    auto out = [&](cpx<double> a, cpx<double> bc, long long kk) {
        const cpx<double> b{bc.x, -bc.y};
        const cpx<double> e{0.5 * (a.x + b.x), 0.5 * (a.y + b.y)};
        const cpx<double> d{0.5 * (a.x - b.x), 0.5 * (a.y - b.y)};
        const cpx<double> o = cmulc(d, w[kk]);  // d * conj(W^k)
        // Z = E + i O
        z[kk] = cpx<double>{s * (e.x - o.y), s * (e.y + o.x)};
    };
    if (k < m) out(xk, xm, k);
    if (m - k != k && m - k < m) out(xm, xk, m - k);
}

hipError_t launch_real_pass(int dir, const void *in, void *out, const void *w, long long m, long long pitch,
                            long long rows, double s, hipStream_t st)
{
    const long long threads = (m / 2 + 1) * rows;
    const long long grid = (threads + 255) / 256;
    if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
    if (dir == 0)
        hipLaunchKernelGGL(r2c_post_kernel, dim3((unsigned)grid), dim3(256), 0, st, static_cast<const cpx<double> *>(in),
                           static_cast<cpx<double> *>(out), static_cast<const cpx<double> *>(w), m, pitch, rows);
    else
        hipLaunchKernelGGL(c2r_pre_kernel, dim3((unsigned)grid), dim3(256), 0, st, static_cast<const cpx<double> *>(in),
                           static_cast<cpx<double> *>(out), static_cast<const cpx<double> *>(w), m, pitch, rows, s);
    return hipGetLastError();
}

}  // namespace tope
