/*
 * CPU oracle for the topeFFT transform path.  TEST INFRASTRUCTURE ONLY.
 *
 * Only the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark load this library, and only as the checker (or, for the baseline,
 * as the CPU restatement that is timed beside the GPU).  The product library
 * never links it and must fail loudly when its HIP code is missing.
 *
 * Contents
 *   oracle_dft1d / oracle_dftnd   naive O(n^2) DFT with exact integer phase
 *                                 reduction and long-double accumulation.
 *                                 Restates the formula of the reference kernel
 *                                 `DFT` (src/kernels1D.cl:549-573) with two
 *                                 fixes: the phase index (n*k) is reduced
 *                                 modulo N in integers before the cos/sin
 *                                 (kernels1D.cl:561-562 evaluates an
 *                                 unreduced argument) and the direction is
 *                                 honoured (the kernel ignores `dir`, :552).
 *                                 INVERSE divides by N as the drivers do
 *                                 (src/fft2d.c:178-189).
 *   oracle_dft_bins               direct DFT at selected output bins, O(N) per
 *                                 bin, for arrays too large for the full
 *                                 naive DFT.
 *   refcpu_fft{1,2,3}d            restatement of the reference ALGORITHM:
 *                                 radix rule (src/fft1d.c:316-329,
 *                                 src/fft2d.c:335-354, src/fft3d.c:492-521),
 *                                 quarter-wave twiddle table
 *                                 (src/kernels1D.cl:7-13), bit-reversal table
 *                                 (kernels1D.cl:526-547), in-place swap pass
 *                                 (kernels1D.cl:505-524, kernels2D.cl:546-592),
 *                                 one full-array radix-2/4/8 DIT pass per
 *                                 stage (kernels1D.cl:15-497 and the 2D/3D
 *                                 axis addressing kernels2D.cl:41-70,
 *                                 kernels3D.cl:42-78), axis order X,Y,Z
 *                                 (fft2d.c:162-163, fft3d.c:244-246), divide
 *                                 by the volume on INVERSE and the n==2 host
 *                                 case (fft1d.c:16-28).  OpenMP-parallel over
 *                                 the reference's work-items.  It is the CPU
 *                                 baseline that the benchmark times.
 *
 * Pinning (tests/test_oracle.py): numpy/pocketfft golden fixtures in
 * tests/golden/, closed forms (delta, constant, planted exponentials, the
 * ramp of test/1d.c:70 and the flat sinusoid of test/2d.c:77), brute force
 * at n <= 16 against a pure-Python sum, Parseval, linearity and round trips.
 *
 * Layout everywhere: interleaved complex128, x fastest:
 * d[2*((z*NY + y)*NX + x)] = re, d[... + 1] = im  (test/3d.c:79).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_FORWARD 1
#define ORACLE_INVERSE 0

typedef long double ld;

/* ------------------------------------------------------------------------ */
/* exact-phase root-of-unity table: w[j] = exp(sign * 2*pi*i * j / n)       */
/* ------------------------------------------------------------------------ */
static const ld PI_L = 3.141592653589793238462643383279502884L;

static void root_table(ld *wr, ld *wi, int64_t n, int dir)
{
    /* j/n is reduced to [0, 1/8] by the octant symmetries so that the
     * long-double cos/sin only ever see small arguments. */
    const ld sign = (dir == ORACLE_FORWARD) ? -1.0L : 1.0L;
    for (int64_t j = 0; j < n; ++j) {
        /* angle = 2*pi*j/n, computed from the integer ratio */
        int64_t e8 = 8 * j;           /* angle in units of pi/4: e8/n */
        int64_t oct = e8 / n;         /* octant 0..7 */
        int64_t rem = e8 - oct * n;   /* 0 <= rem < n */
        /* angle = (oct + rem/n) * pi/4 */
        ld frac = (ld)rem / (ld)n;    /* in [0,1) */
        ld c, s;
        ld a;
        /* reduce to a base angle in [0, pi/4] and map back */
        switch (oct) {
        case 0: a = frac * PI_L / 4; c = cosl(a); s = sinl(a); break;
        case 1: a = (1 - frac) * PI_L / 4; c = sinl(a); s = cosl(a); break;
        case 2: a = frac * PI_L / 4; c = -sinl(a); s = cosl(a); break;
        case 3: a = (1 - frac) * PI_L / 4; c = -cosl(a); s = sinl(a); break;
        case 4: a = frac * PI_L / 4; c = -cosl(a); s = -sinl(a); break;
        case 5: a = (1 - frac) * PI_L / 4; c = -sinl(a); s = -cosl(a); break;
        case 6: a = frac * PI_L / 4; c = sinl(a); s = -cosl(a); break;
        default: a = (1 - frac) * PI_L / 4; c = cosl(a); s = -sinl(a); break;
        }
        wr[j] = c;
        wi[j] = sign * s;
    }
}

/* ------------------------------------------------------------------------ */
/* naive DFT                                                                */
/* ------------------------------------------------------------------------ */

/* One line of length n with element stride `stride` (complex units).
 * out may not alias in. */
static void dft_line(const double *in, double *out, int64_t n, int64_t stride,
                     const ld *wr, const ld *wi, ld scale)
{
    for (int64_t k = 0; k < n; ++k) {
        ld accr = 0, acci = 0;
        int64_t idx = 0; /* (j*k) mod n, updated incrementally */
        for (int64_t j = 0; j < n; ++j) {
            ld xr = in[2 * j * stride], xi = in[2 * j * stride + 1];
            accr += xr * wr[idx] - xi * wi[idx];
            acci += xr * wi[idx] + xi * wr[idx];
            idx += k;
            if (idx >= n) idx -= n;
        }
        out[2 * k * stride] = (double)(accr * scale);
        out[2 * k * stride + 1] = (double)(acci * scale);
    }
}

/* out = DFT(in) of length n (FORWARD unnormalised, INVERSE scaled by 1/n). */
void oracle_dft1d(const double *in, double *out, int64_t n, int dir)
{
    ld *wr = malloc(sizeof(ld) * n), *wi = malloc(sizeof(ld) * n);
    root_table(wr, wi, n, dir);
    ld scale = (dir == ORACLE_INVERSE) ? 1.0L / (ld)n : 1.0L;
    if (n >= 1024) {
        /* split the output bins over threads */
        #pragma omp parallel for schedule(dynamic, 16)
        for (int64_t k = 0; k < n; ++k) {
            ld accr = 0, acci = 0;
            int64_t idx = 0;
            for (int64_t j = 0; j < n; ++j) {
                ld xr = in[2 * j], xi = in[2 * j + 1];
                accr += xr * wr[idx] - xi * wi[idx];
                acci += xr * wi[idx] + xi * wr[idx];
                idx += k;
                if (idx >= n) idx -= n;
            }
            out[2 * k] = (double)(accr * scale);
            out[2 * k + 1] = (double)(acci * scale);
        }
    } else {
        dft_line(in, out, n, 1, wr, wi, scale);
    }
    free(wr);
    free(wi);
}

/* Naive separable DFT of an nx*ny*nz array (x fastest): the 1D naive DFT is
 * applied along x, then y, then z.  in and out may alias. */
void oracle_dftnd(const double *in, double *out, int64_t nx, int64_t ny,
                  int64_t nz, int dir)
{
    int64_t total = nx * ny * nz;
    double *a = malloc(sizeof(double) * 2 * total);
    double *b = malloc(sizeof(double) * 2 * total);
    memcpy(a, in, sizeof(double) * 2 * total);
    int64_t dims[3] = {nx, ny, nz};
    int64_t strides[3] = {1, nx, nx * ny};
    for (int ax = 0; ax < 3; ++ax) {
        int64_t n = dims[ax], st = strides[ax];
        if (n == 1) continue;
        ld *wr = malloc(sizeof(ld) * n), *wi = malloc(sizeof(ld) * n);
        root_table(wr, wi, n, dir);
        ld scale = (dir == ORACLE_INVERSE) ? 1.0L / (ld)n : 1.0L;
        int64_t lines = total / n;
        #pragma omp parallel for schedule(static)
        for (int64_t l = 0; l < lines; ++l) {
            /* line l: decompose into the two other coordinates */
            int64_t base;
            if (ax == 0) base = l * nx;
            else if (ax == 1) base = (l / nx) * nx * ny + (l % nx);
            else base = l;
            dft_line(a + 2 * base, b + 2 * base, n, st, wr, wi, scale);
        }
        free(wr);
        free(wi);
        double *t = a; a = b; b = t;
    }
    memcpy(out, a, sizeof(double) * 2 * total);
    free(a);
    free(b);
}

/* Direct DFT evaluated only at the flat output indices bins[0..nb) of an
 * nx*ny*nz array.  out[2*b], out[2*b+1] receive bin b.  The phase of each
 * axis is reduced exactly modulo that axis' length. */
void oracle_dft_bins(const double *in, int64_t nx, int64_t ny, int64_t nz,
                     const int64_t *bins, int64_t nb, int dir, double *out)
{
    ld *wx = malloc(sizeof(ld) * 2 * nx), *wy = malloc(sizeof(ld) * 2 * ny),
       *wz = malloc(sizeof(ld) * 2 * nz);
    ld *tr = malloc(sizeof(ld) * nx), *ti = malloc(sizeof(ld) * nx);
    root_table(tr, ti, nx, dir);
    for (int64_t i = 0; i < nx; ++i) { wx[2 * i] = tr[i]; wx[2 * i + 1] = ti[i]; }
    free(tr); free(ti);
    tr = malloc(sizeof(ld) * ny); ti = malloc(sizeof(ld) * ny);
    root_table(tr, ti, ny, dir);
    for (int64_t i = 0; i < ny; ++i) { wy[2 * i] = tr[i]; wy[2 * i + 1] = ti[i]; }
    free(tr); free(ti);
    tr = malloc(sizeof(ld) * nz); ti = malloc(sizeof(ld) * nz);
    root_table(tr, ti, nz, dir);
    for (int64_t i = 0; i < nz; ++i) { wz[2 * i] = tr[i]; wz[2 * i + 1] = ti[i]; }
    free(tr); free(ti);
    ld scale = (dir == ORACLE_INVERSE) ? 1.0L / ((ld)nx * ny * nz) : 1.0L;

    for (int64_t b = 0; b < nb; ++b) {
        int64_t kx = bins[b] % nx, ky = (bins[b] / nx) % ny, kz = bins[b] / (nx * ny);
        ld sr = 0, si = 0;
        #pragma omp parallel for reduction(+:sr, si) schedule(static)
        for (int64_t zy = 0; zy < ny * nz; ++zy) {
            int64_t y = zy % ny, z = zy / ny;
            const double *row = in + 2 * (zy * nx);
            ld rr = 0, ri = 0;
            int64_t idx = 0;
            for (int64_t x = 0; x < nx; ++x) {
                ld xr = row[2 * x], xi = row[2 * x + 1];
                rr += xr * wx[2 * idx] - xi * wx[2 * idx + 1];
                ri += xr * wx[2 * idx + 1] + xi * wx[2 * idx];
                idx += kx;
                idx %= nx;
            }
            int64_t iy = (ky * y) % ny, iz = (kz * z) % nz;
            ld cr = wy[2 * iy] * wz[2 * iz] - wy[2 * iy + 1] * wz[2 * iz + 1];
            ld ci = wy[2 * iy] * wz[2 * iz + 1] + wy[2 * iy + 1] * wz[2 * iz];
            sr += rr * cr - ri * ci;
            si += rr * ci + ri * cr;
        }
        out[2 * b] = (double)(sr * scale);
        out[2 * b + 1] = (double)(si * scale);
    }
    free(wx); free(wy); free(wz);
}

/* ------------------------------------------------------------------------ */
/* refcpu: restatement of the reference algorithm                           */
/* ------------------------------------------------------------------------ */

/* The twiddle constant the reference kernels use (kernels1D.cl:4). */
static const double REF_2PI = 6.283185307179586476925286766559;
static const double REF_R2 = 0.707106781186547524400844362104; /* kernels1D.cl:5 */

static int ilog2_i(int64_t n)
{
    int l = 0;
    while (((int64_t)1 << l) < n) ++l;
    return l;
}

/* Radix rule of fft1d.c:316-329 / fft2d.c:335-354: 8 if log%3==0, else 4 if
 * log is even, else 2.  Returns -1 for a length that is not a power of two. */
int refcpu_radix(int64_t n)
{
    if (n < 1 || (n & (n - 1))) return -1;
    int l = ilog2_i(n);
    if (l % 3 == 0) return 8;
    if (l % 2 == 0) return 4;
    return 2;
}

/* Quarter-wave table tw[b] = (cos(2*pi*b/n), -sin(2*pi*b/n)), b < n/4, with
 * the argument formed as the kernel forms it (kernels1D.cl:11). */
static void ref_twiddles(double *tw, int64_t n)
{
    for (int64_t b = 0; b < n / 4; ++b) {
        tw[2 * b] = cos(REF_2PI * (double)b / (double)n);
        tw[2 * b + 1] = -sin(REF_2PI * (double)b / (double)n);
    }
}

/* W_n^c for 0 <= c < n from the quarter-wave table: the quadrant q = c/(n/4)
 * multiplies the table entry by (-i)^q (kernels1D.cl:62-69). */
static inline void ref_w(const double *tw, int64_t n, int64_t c, int dir,
                         double *wr, double *wi)
{
    int64_t red = n / 4;
    int64_t q = c / red, b = c % red;
    double x = tw[2 * b], y = tw[2 * b + 1];
    double r, i;
    switch (q) {
    case 0: r = x; i = y; break;
    case 1: r = y; i = -x; break;
    case 2: r = -x; i = -y; break;
    default: r = -y; i = x; break;
    }
    if (dir == ORACLE_INVERSE) i = -i;
    *wr = r;
    *wi = i;
}

static void ref_bitrev(int *rev, int64_t n)
{
    int l = ilog2_i(n);
    for (int64_t g = 0; g < n; ++g) {
        int s = 0;
        for (int b = 0; b < l; ++b)
            if (g & ((int64_t)1 << b)) s |= 1 << (l - 1 - b);
        rev[g] = s;
    }
}

static inline int rev_digits(int j, int r)
{
    /* reverse the log2(r) bits of j */
    int l = (r == 8) ? 3 : (r == 4) ? 2 : 1, s = 0;
    for (int b = 0; b < l; ++b)
        if (j & (1 << b)) s |= 1 << (l - 1 - b);
    return s;
}

/* In-register r-point DFT (r = 2,4,8) with natural-order output, forward
 * sign e^{-2 pi i/r}, conjugate for INVERSE; computed as a radix-2 network
 * like the reference's DIT4/DIT8 bodies (kernels1D.cl:134-153, :321-396). */
static void small_dft(double *xr, double *xi, int r, int dir)
{
    double s = (dir == ORACLE_FORWARD) ? -1.0 : 1.0;
    if (r == 2) {
        double ar = xr[0], ai = xi[0];
        xr[0] = ar + xr[1]; xi[0] = ai + xi[1];
        xr[1] = ar - xr[1]; xi[1] = ai - xi[1];
        return;
    }
    if (r == 4) {
        double t0r = xr[0] + xr[2], t0i = xi[0] + xi[2];
        double t1r = xr[0] - xr[2], t1i = xi[0] - xi[2];
        double t2r = xr[1] + xr[3], t2i = xi[1] + xi[3];
        double dr = xr[1] - xr[3], di = xi[1] - xi[3];
        /* multiply (dr, di) by s*i */
        double t3r = -s * di, t3i = s * dr;
        xr[0] = t0r + t2r; xi[0] = t0i + t2i;
        xr[2] = t0r - t2r; xi[2] = t0i - t2i;
        xr[1] = t1r + t3r; xi[1] = t1i + t3i;
        xr[3] = t1r - t3r; xi[3] = t1i - t3i;
        return;
    }
    /* r == 8: two 4-point DFTs of the even / odd samples, then W8^k */
    double er[4] = {xr[0], xr[2], xr[4], xr[6]}, ei[4] = {xi[0], xi[2], xi[4], xi[6]};
    double orr[4] = {xr[1], xr[3], xr[5], xr[7]}, oi[4] = {xi[1], xi[3], xi[5], xi[7]};
    small_dft(er, ei, 4, dir);
    small_dft(orr, oi, 4, dir);
    for (int k = 0; k < 4; ++k) {
        double wr, wi;
        switch (k) {
        case 0: wr = 1; wi = 0; break;
        case 1: wr = REF_R2; wi = s * REF_R2; break;
        case 2: wr = 0; wi = s; break;
        default: wr = -REF_R2; wi = s * REF_R2; break;
        }
        double pr = orr[k] * wr - oi[k] * wi, pi = orr[k] * wi + oi[k] * wr;
        xr[k] = er[k] + pr; xi[k] = ei[k] + pi;
        xr[k + 4] = er[k] - pr; xi[k + 4] = ei[k] - pi;
    }
}

/* One transform axis of length n over `lines` lines: line l starts at
 * base(l) and has element stride `st`.  Bit-reversal swap, then log_r(n)
 * DIT stages, each a full pass over all lines. */
typedef struct {
    int64_t nx, ny, nz;
    int axis;
} ref_geom;

static inline int64_t line_base(const ref_geom *g, int64_t l)
{
    if (g->axis == 0) return l * g->nx;                                  /* type 1 */
    if (g->axis == 1) return (l / g->nx) * g->nx * g->ny + (l % g->nx);  /* type 2 */
    return l;                                                            /* type 3 */
}

static void ref_axis(double *d, const ref_geom *g, int dir)
{
    int64_t dims[3] = {g->nx, g->ny, g->nz};
    int64_t strides[3] = {1, g->nx, g->nx * g->ny};
    int64_t n = dims[g->axis], st = strides[g->axis];
    int64_t lines = (g->nx * g->ny * g->nz) / n;
    int r = refcpu_radix(n);
    if (n < 2) return;
    int *rev = malloc(sizeof(int) * n);
    ref_bitrev(rev, n);
    int64_t tws = n / 4 > 0 ? n / 4 : 1;
    double *tw = malloc(sizeof(double) * 2 * tws);
    if (n >= 4) ref_twiddles(tw, n);

    /* swap pass (kernels2D.cl:546-592): lower index of each pair swaps */
    #pragma omp parallel for schedule(static)
    for (int64_t l = 0; l < lines; ++l) {
        double *b = d + 2 * line_base(g, l);
        for (int64_t i = 0; i < n; ++i) {
            int64_t j = rev[i];
            if (i < j) {
                double tr = b[2 * j * st], ti = b[2 * j * st + 1];
                b[2 * j * st] = b[2 * i * st];
                b[2 * j * st + 1] = b[2 * i * st + 1];
                b[2 * i * st] = tr;
                b[2 * i * st + 1] = ti;
            }
        }
    }

    if (n == 2) {
        /* a single radix-2 stage with twiddle 1 */
        #pragma omp parallel for schedule(static)
        for (int64_t l = 0; l < lines; ++l) {
            double *b = d + 2 * line_base(g, l);
            double xr[2] = {b[0], b[2 * st]}, xi[2] = {b[1], b[2 * st + 1]};
            small_dft(xr, xi, 2, dir);
            b[0] = xr[0]; b[1] = xi[0]; b[2 * st] = xr[1]; b[2 * st + 1] = xi[1];
        }
    } else {
        int lr = (r == 8) ? 3 : (r == 4) ? 2 : 1;
        int stages = ilog2_i(n) / lr;
        int64_t items = n / r; /* work-items per line (kernels1D.cl:21, fft2d.c:36-47) */
        for (int s = 1; s <= stages; ++s) {
            int64_t powX = (int64_t)1 << (lr * s), powXm1 = powX / r;
            #pragma omp parallel for schedule(static)
            for (int64_t w = 0; w < lines * items; ++w) {
                int64_t l = w / items, id = w % items;
                double *b = d + 2 * line_base(g, l);
                int64_t yIndex = id / powXm1, kIndex = id % powXm1;
                int64_t coeff = kIndex * (n / powX);
                double xr[8], xi[8], ur[8], ui[8];
                for (int j = 0; j < r; ++j) {
                    int64_t e = kIndex + yIndex * powX + j * powXm1;
                    xr[j] = b[2 * e * st];
                    xi[j] = b[2 * e * st + 1];
                }
                /* slot j holds the sub-DFT of digit a = rev(j); it is
                 * multiplied by W_n^{a*coeff} (kernels1D.cl:57-108,
                 * :211-319), skipped when kIndex == 0 */
                for (int a = 0; a < r; ++a) {
                    int j = rev_digits(a, r);
                    double vr = xr[j], vi = xi[j];
                    if (kIndex != 0 && a != 0) {
                        double wr, wi;
                        ref_w(tw, n, a * coeff, dir, &wr, &wi);
                        double tr = vr * wr - vi * wi;
                        vi = vr * wi + vi * wr;
                        vr = tr;
                    }
                    ur[a] = vr;
                    ui[a] = vi;
                }
                small_dft(ur, ui, r, dir);
                for (int q = 0; q < r; ++q) {
                    int64_t e = kIndex + yIndex * powX + q * powXm1;
                    b[2 * e * st] = ur[q];
                    b[2 * e * st + 1] = ui[q];
                }
            }
        }
    }
    free(rev);
    free(tw);
}

static void ref_divide(double *d, int64_t total)
{
    /* kernels2D.cl:539-544: data /= volume */
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < 2 * total; ++i) d[i] /= (double)total;
}

/* In-place transforms with the reference's conventions.  Return 0, or -1
 * for a size the reference's radix path does not handle. */
int refcpu_fft1d(double *d, int64_t n, int dir)
{
    if (refcpu_radix(n) < 0) return -1;
    if (n == 2) {
        /* host special case fft1d.c:16-28 */
        double hr = d[0], hi = d[1];
        d[0] += d[2]; d[1] += d[3];
        d[2] = hr - d[2]; d[3] = hi - d[3];
        if (dir == ORACLE_INVERSE) { d[0] /= 2; d[2] /= 2; d[1] /= 2; d[3] /= 2; }
        return 0;
    }
    ref_geom g = {n, 1, 1, 0};
    ref_axis(d, &g, dir);
    if (dir == ORACLE_INVERSE) ref_divide(d, n);
    return 0;
}

int refcpu_fft2d(double *d, int64_t nx, int64_t ny, int dir)
{
    if (refcpu_radix(nx) < 0 || refcpu_radix(ny) < 0) return -1;
    ref_geom g = {nx, ny, 1, 0};
    ref_axis(d, &g, dir);
    g.axis = 1;
    ref_axis(d, &g, dir);
    if (dir == ORACLE_INVERSE) ref_divide(d, nx * ny);
    return 0;
}

int refcpu_fft3d(double *d, int64_t nx, int64_t ny, int64_t nz, int dir)
{
    if (refcpu_radix(nx) < 0 || refcpu_radix(ny) < 0 || refcpu_radix(nz) < 0) return -1;
    ref_geom g = {nx, ny, nz, 0};
    ref_axis(d, &g, dir);
    g.axis = 1;
    ref_axis(d, &g, dir);
    g.axis = 2;
    ref_axis(d, &g, dir);
    if (dir == ORACLE_INVERSE) ref_divide(d, nx * ny * nz);
    return 0;
}

/* Number of OpenMP threads the oracle uses (for the baseline's core count). */
#ifdef _OPENMP
#include <omp.h>
int oracle_threads(void) { return omp_get_max_threads(); }
#else
int oracle_threads(void) { return 1; }
#endif
