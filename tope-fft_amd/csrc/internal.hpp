// Internal interfaces shared by the kernel translation units and the C-ABI.
#pragma once
#include <hip/hip_runtime.h>

#include "axis_kernel.hpp"

struct topeFFT;

namespace tope {

enum { PREC_F64 = 0, PREC_F32 = 1 };

// Launches one axis pass: `args` points to AxisArgs<double> or
// AxisArgs<float> according to the variant's precision.
typedef void (*AxisLaunchFn)(const void *in, void *out, const void *args, unsigned grid, hipStream_t s,
                             int sign);

// One compiled instantiation of axis_kernel.
struct AxisVariant {
    int prec;     // PREC_F64 / PREC_F32
    int n;        // transform length
    int e;        // elements per thread
    int tb;       // transforms per block
    bool cfast;   // lanes walk transforms first (strided-axis tiles)
    bool split;   // real/imag LDS exchange in two rounds
    bool dif2;    // row pairs + radix-2 DIF step of the next axis
    bool dit2;    // column tiles split into even/odd halves (axis_dit2.hpp):
                  // stage table of (n/2, e), W_n^k table (k < n/2) in dif_tw
    bool ct3;     // three-step column tile (axis_ct3.hpp): its own twiddle table
    int threads;  // block size
    int twsize;   // stage twiddle entries for (n, e), (n/2, e) when dit2, CT3_TW when ct3
    int lds_bytes;
    AxisLaunchFn fn;
    const char *name;
};

const AxisVariant *axis_variants(int *count);
// Same list, same order, compiled with 64-bit addressing.
const AxisVariant *axis_variants_big(int *count);

// Runtime mirror of Sched<N, E> (the device derives the same layout at
// compile time; api.cpp checks that both agree for every variant).
struct HostSched {
    int nst;
    int radix[16];
    int span[16];
    int off[16];
    int twsize;
};
HostSched host_sched(int n, int e);

// FP64 generation of the stage twiddle table of schedule (n, e): entry
// off[s] + (m-1)*span[s] + k = W_{span[s]*radix[s]}^{k*m}, forward sign,
// rounded once to the storage precision.
hipError_t gen_stage_twiddles(void *out, int prec, int n, int e, hipStream_t s);
// out[i] = W_M^{i*mult} for i < count (FP64 generation, rounded once).
hipError_t gen_power_table(void *out, int prec, long long M, long long mult, long long count, hipStream_t s);

// O(n^2) DFT of a length-n vector (any n): out[k] = scale * sum_j in[j]
// W_n^{sign (j*k mod n)}; `roots` holds W_n^i, i < n (forward sign).
hipError_t launch_naive_dft(int prec, const void *in, void *out, const void *roots, long long n, int sign,
                            double scale, hipStream_t s);

// Mixed-precision MFMA rows (mfma_kernel.hpp).
struct MfmaArgs;
hipError_t gen_mfma_constants(void *out, hipStream_t s);
size_t mfma_constants_bytes();
bool mfma_supported(int n);
hipError_t launch_mfma_rows(int mode, int n, const void *in, void *out, const MfmaArgs &a, int sign, hipStream_t s);
// Real-input transforms (kernels_real.hip): dir 0 = R2C half-spectrum
// post-processing (Z, row pitch m -> X, m + 1 values per row at row pitch
// `pitch`), 1 = C2R pre-processing (X -> Z, scaled by s); w = W_n^k,
// k = 0..m, n = 2m.  pack: column 0 carries X[0] + i X[m] (column m
// unused) for plans whose column passes run on the packed form.
hipError_t launch_real_pass(int prec, int dir, const void *in, void *out, const void *w, long long m, long long pitch,
                            long long rows, double s, int pack, hipStream_t st);
// The m-point row transform fused with that step (64 <= m <= 4096, powers of
// two): dir 0 reads the real rows (pitch m complex) and writes the half
// spectrum, dir 1 the reverse (scaled by s); stw is the stage twiddle table
// of schedule (m, 16).
bool real_rows_supported(long long m);
hipError_t launch_real_rows(int prec, int dir, const void *in, void *out, const void *stw, const void *w, long long m,
                            long long pitch, long long rows, double s, int pack, hipStream_t st);
// Packed DC/Nyquist columns of the half spectrum X (rows K = ky + ny kz at
// row pitch `pitch`): dir 0 splits column 0's FZ = F0 + i Fm into column 0
// (F0) and column m (Fm) after the forward column passes; dir 1 packs
// F0 + i Fm into column 0 before the inverse ones.
hipError_t launch_real_pack(int prec, int dir, void *X, long long m, long long pitch, long long ny, long long nz,
                            hipStream_t st);
// Bluestein chirp-z passes for non-power-of-two axes (kernels_chirp.hip),
// FP64: n the axis length, M the power-of-two convolution length; the lines
// (o < O, c < C) of the axis sit at o os + c cs, elements is apart, and
// line l is padded into row l of an M-element-row scratch array.
enum { CHIRP_FILTER = 0, CHIRP_PRE = 1, CHIRP_MUL = 2, CHIRP_POST = 3 };
// Mixed-radix (2, 3, 4, 5, 7, 8, 16) in-place axis pass for 7-smooth
// lengths up to MIXED_MAX_N (kernels_mixed.hip), FP64: line l = o C + c
// (o < O, c < C) starts at o os + c cs, its elements are `is` apart; roots
// holds W_n^i, i < n (forward sign).
constexpr int MIXED_MAX_N = 4096;
struct MixedArgs {
    long long O, os, C, cs, is, lines;
    int n, tb, nst;
    int radix[16];
    const void *roots;  // W_n^i, i < n, in the pass's precision (forward sign)
    double scale;
    // output geometry (mixed_plan sets it equal to the input's)
    long long oos, ocs, ois;
    // optional four-step post-twiddle: output k of line (o, c) times
    // W_pt_n^{(o pt_o + c pt_c) k}, the exponent below pt_n <= 2^24 without
    // reduction; W_pt_n^a = pt_lo[a mod pt_split] * pt_hi[a / pt_split]
    const void *pt_lo, *pt_hi;
    long long pt_n, pt_o, pt_c;
    int pt_split;
};
bool mixed_supported(long long n);
// 7-smooth n above MIXED_MAX_N that splits as n = n1 n2 with both factors
// at most MIXED_MAX_N (mixed-radix four-step); *n1 >= n2 is the split used.
bool mixed4_split(long long n, long long *n1);
MixedArgs mixed_plan(long long n, long long O, long long os, long long C, long long cs, long long is);
unsigned mixed_grid(const MixedArgs &a);
hipError_t launch_mixed(int prec, const void *in, void *out, const MixedArgs &a, int sign, hipStream_t st);
struct ChirpGeo {
    long long O, os, C, cs, is;
};
hipError_t launch_chirp(int op, const void *in, void *out, const void *table, const ChirpGeo &g, long long n,
                        long long M, double s, int inverse, hipStream_t st);
// Mixed-precision column tiles (mfma_tile.hpp): 16 unit-stride lines per block.
struct MfmaTileArgs;
bool mfma_tile_supported(int n);
hipError_t launch_mfma_tile(int mode, int n, const void *in, void *out, const MfmaTileArgs &a, unsigned grid, int sign,
                            hipStream_t s);

// Framework plumbing (api.cpp) for modules outside it.
hipStream_t fw_stream(struct topeFFT *f);
void *fw_handle(struct topeFFT *f);
void fw_retain(struct topeFFT *f);
void fw_release(struct topeFFT *f, void *owner);
// Row variant the planner would pick for one transform of length n.
const AxisVariant *row_variant(int prec, int n);
// One-shot variant the planner would pick for `lines` transforms of length
// n (per outer index, `outer` of them); cfast: lanes walk the transforms
// first.  Never an even/odd-split variant.
const AxisVariant *axis_variant_for(int prec, int n, bool cfast, long long lines, long long outer);

// Allocation-free entry point (noalloc.hip): per-device address of the
// static scratch array (taken once per device at topeFFTInit), and the
// plan-free 1D launch path over it (returns a TOPE status).
hipError_t noalloc_setup(int dev);
void *noalloc_ptr(int dev);
int noalloc_exec_1d(int dev, void *data, long long n, long long count, int inverse, hipStream_t s);

// Persistent uneven batch (many_kernel.hpp); args is ManyArgs<double|float>.
hipError_t launch_many(int prec, int threads, const void *in, void *out, const void *args, int grid, int sign,
                       hipStream_t s);

}  // namespace tope
