// Allocation-free 1D transforms on caller-owned device memory.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc):
// topeExec1DNoAlloc runs the same axis-pass kernels as a 1D plan's
// topePlanExecPtr, with no plan.  Everything those kernels need beyond the
// caller's array -- the stage twiddle tables, the four-step's post-twiddle
// tables and its intermediate array -- lives in one static device array,
// `scratch`, of kNaBytes per GPU.  The tables are regenerated on every call
// (they are a few KiB and depend on n), so a call never depends on what an
// earlier call left there; calls that overlap on one GPU share the array and
// race.
#include <mutex>

#include "../../include/topefft.h"
#include "internal.hpp"

namespace tope {

// Tables: at most two stage tables of 4096 entries and two power tables of
// 2^11 entries (16 B each) fit in the first kNaTab bytes; the four-step's
// intermediate array (count * n complex128) takes the rest.
constexpr size_t kNaTab = 256u << 10;
constexpr size_t kNaBytes = kNaTab + (size_t)TOPE_NOALLOC_MAX_ELEMENTS * 16;

static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices];  // device address of `scratch` on each GPU

}  // namespace

// Called by topeFFTInit with `dev` current: the device address of this
// GPU's copy of the array (the host-side name is only a stand-in).
hipError_t noalloc_setup(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[dev]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[dev] = p;
    return hipSuccess;
}

void *noalloc_ptr(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_ptr[dev];
}

namespace {

AxisArgs<double> args_for(const AxisVariant *v, long long O, long long os_in, long long os_out, long long C,
                          long long cs_in, long long cs_out, long long is_in, long long is_out, const void *tw,
                          long long elems)
{
    AxisArgs<double> a{};
    a.in_os = os_in;
    a.out_os = os_out;
    a.in_cs = cs_in;
    a.out_cs = cs_out;
    a.in_is = is_in;
    a.out_is = is_out;
    a.C = (int)C;
    a.tiles = (int)((C + v->tb - 1) / v->tb);
    a.tw = static_cast<const cpx<double> *>(tw);
    a.scale = 1.0;
    a.xcd = v->cfast ? 1 : 0;
    a.in_elems = a.out_elems = elems;
    (void)O;
    return a;
}

}  // namespace

// `count` length-n complex128 transforms back to back in `data` (device),
// in place on stream s; the same schedule as a 1D plan of length n (one
// fused pass up to 4096, the four-step beyond) with the scratch array in
// place of the plan's tables and second buffer.  Returns a TOPE status.
int noalloc_exec_1d(int dev, void *data, long long n, long long count, int inverse, hipStream_t s)
{
    unsigned char *scr = static_cast<unsigned char *>(noalloc_ptr(dev));
    if (!scr) return TOPE_ERROR_NOT_INITIALIZED;
    if (n < 1 || (n & (n - 1)) || n > (1LL << 24) || count < 1) return TOPE_ERROR_INVALID_SIZE;
    const long long elems = n * count;
    if (n > 4096 && elems > TOPE_NOALLOC_MAX_ELEMENTS) return TOPE_ERROR_SCRATCH_LIMIT;
    if (n == 1) return TOPE_SUCCESS;
    const int sign = inverse ? 1 : -1;
    const double sc = inverse ? 1.0 / (double)n : 1.0;
    unsigned char *tab = scr;
    void *mid = scr + kNaTab;
    hipError_t err = hipSuccess;
    if (n <= 4096) {
        const AxisVariant *v = axis_variant_for(PREC_F64, (int)n, false, count, 1);
        if (!v) return TOPE_ERROR_NO_KERNEL;
        err = gen_stage_twiddles(tab, PREC_F64, (int)n, v->e, s);
        if (err != hipSuccess) return (int)err;
        AxisArgs<double> a = args_for(v, 1, 0, 0, count, n, n, 1, 1, tab, elems);
        a.scale = sc;
        if (a.tiles <= 0) return TOPE_ERROR_INVALID_SIZE;
        v->fn(data, data, &a, (unsigned)a.tiles, s, sign);
        err = hipGetLastError();
        return err == hipSuccess ? TOPE_SUCCESS : (int)err;
    }
    // four-step, i = n1 + N1 n2: step 1 transforms n2 (length N2, stride N1)
    // for each (transform, n1) into the scratch array, multiplying output k2
    // by W_n^{n1 k2}; step 2 transforms n1 (length N1) for each (transform,
    // k2) and writes X[k2 + N2 k1] back into `data`
    int L = 0;
    while ((1LL << L) < n) ++L;
    const int lo_bits = (L + 1) / 2;
    const long long N1 = 1LL << lo_bits, N2 = n / N1;
    const AxisVariant *v1 = axis_variant_for(PREC_F64, (int)N2, true, N1, count);
    const AxisVariant *v2 = axis_variant_for(PREC_F64, (int)N1, true, N2, count);
    // (the even/odd-split tiles take tables of their own; lengths up to 2048
    // here never select them)
    if (!v1 || !v2 || v1->dit2 || v2->dit2) return TOPE_ERROR_NO_KERNEL;
    // table layout: stage tables of step 1 and step 2 (<= 4096 entries each),
    // then W_n^i (i < N1) and W_n^{i N1} (i < n / N1)
    unsigned char *tw1 = tab, *tw2 = tab + (64u << 10), *lo = tab + (128u << 10), *hi = tab + (192u << 10);
    err = gen_stage_twiddles(tw1, PREC_F64, (int)N2, v1->e, s);
    if (err == hipSuccess) err = gen_stage_twiddles(tw2, PREC_F64, (int)N1, v2->e, s);
    if (err == hipSuccess) err = gen_power_table(lo, PREC_F64, n, 1, N1, s);
    if (err == hipSuccess) err = gen_power_table(hi, PREC_F64, n, N1, n / N1, s);
    if (err != hipSuccess) return (int)err;
    AxisArgs<double> a1 = args_for(v1, count, n, n, N1, 1, 1, N1, N1, tw1, elems);
    a1.pt_lo = reinterpret_cast<const cpx<double> *>(lo);
    a1.pt_hi = reinterpret_cast<const cpx<double> *>(hi);
    a1.pt_o = 0;
    a1.pt_c = 1;
    a1.pt_n = n;
    a1.pt_shift = lo_bits;
    AxisArgs<double> a2 = args_for(v2, count, n, n, N2, N1, 1, 1, N2, tw2, elems);
    a2.scale = sc;
    const long long g1 = count * a1.tiles, g2 = count * a2.tiles;
    if (g1 <= 0 || g2 <= 0 || g1 > 0x7fffffffLL || g2 > 0x7fffffffLL) return TOPE_ERROR_INVALID_SIZE;
    v1->fn(data, mid, &a1, (unsigned)g1, s, sign);
    v2->fn(mid, data, &a2, (unsigned)g2, s, sign);
    err = hipGetLastError();
    return err == hipSuccess ? TOPE_SUCCESS : (int)err;
}

}  // namespace tope
