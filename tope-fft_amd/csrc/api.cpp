// C-ABI of the topeFFT library for MI355X: framework bootstrap, plan/exec/
// destroy for 1D/2D/3D complex FP64 transforms, and the device-resident and
// timing extensions declared in include/topefft.h.
//
// Reference counterparts: topeFFTInit src/util.c:21-236, tope1D* src/fft1d.c,
// tope2D* src/fft2d.c, tope3D* src/fft3d.c.  The reference plans one
// single-radix launch per stage plus a bit-reversal swap and a divide pass
// per axis, with a host clFinish after each; a plan here is a short list of
// fused axis passes (one per axis for lengths up to 4096, two for longer
// axes via a four-step split) enqueued back to back on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/topefft.h"
#include "internal.hpp"
#include "mfma_kernel.hpp"
#include "mfma_tile.hpp"
#include "axis_ct3.hpp"

using namespace tope;

namespace {

constexpr int kMaxFused = 4096;          // longest axis one pass holds on chip
constexpr long long kMaxAxis = 1LL << 24;  // four-step limit (4096 x 4096)
constexpr long long kMaxLen = 1LL << 30;   // three-pass limit of a power-of-two axis here (16 GiB FP64)
constexpr long long kChirpMin = 64;        // longer non-power-of-two 1D lengths use Bluestein

struct EventPair {
    hipEvent_t start = nullptr, stop = nullptr;
    bool recorded = false;
};

struct FrameworkImpl {
    hipStream_t stream = nullptr;  // the stream every plan op is enqueued on
    hipStream_t own = nullptr;     // the framework's own stream (stream unless a caller's is set)
    int live_plans = 0;
    EventPair last;  // last exec's device time (profileThis)
};

enum PassKind {
    PASS_AXIS = 0,
    PASS_DFT = 1,
    PASS_MFMA = 2,
    PASS_MFMA_TILE = 3,
    PASS_REAL = 4,
    PASS_CHIRP = 5,
    PASS_MIXED = 6,
    PASS_RPACK = 7  // real plans: unpack (FORWARD) / pack (INVERSE) of the DC/Nyquist columns
};

struct Pass {
    int kind = PASS_AXIS;
    const AxisVariant *var = nullptr;  // PASS_AXIS: entry of either variant list
    AxisArgs<double> a{};              // FP64 axis pass
    AxisArgs<float> a32{};             // FP32 axis pass
    MfmaArgs m{};                      // PASS_MFMA
    MfmaTileArgs mt{};                 // PASS_MFMA_TILE
    int mfma_mode = 0, mfma_n = 0;
    unsigned grid = 0;
    int in_buf = 0, out_buf = 0;
    long long reps = 1, rep_in = 0, rep_out = 0;  // repeated launches (rare geometry)
    long long base = 0;                           // element offset of the pass's geometry in both buffers
    bool last = false;                            // receives the inverse 1/N scale
    long long dft_n = 0;
    const void *dft_roots = nullptr;
    const void *real_w = nullptr;  // PASS_REAL: W_n^k, k = 0..m (FP64)
    const void *real_stw = nullptr;  // PASS_REAL fused with the row transform: stage table of (m, 16)
    int fixed_dir = -1;            // >= 0: the pass always runs in this direction (Bluestein's inner FFTs)
    int chirp_op = 0;              // PASS_CHIRP: CHIRP_PRE / CHIRP_MUL / CHIRP_POST
    long long chirp_n = 0, chirp_M = 0;
    ChirpGeo chirp_g{1, 0, 1, 0, 1};
    const void *chirp_tab[2] = {nullptr, nullptr};  // transformed filters, forward / inverse
    MixedArgs mx{};                                 // PASS_MIXED
    const char *name = "";
};

struct PlanImpl {
    FrameworkImpl *fw = nullptr;
    int prec = PREC_F64;
    size_t elem = 16;  // bytes per complex element
    int ndim = 0;
    long long dims[3] = {1, 1, 1};
    long long total = 0;
    long long norm = 0;  // elements of one transform: INVERSE divides by it
    size_t bytes = 0;      // the transform's data (host transfers)
    size_t buf_bytes = 0;  // device arrays when larger than bytes (Bluestein's padded length)
    void *buf[2] = {nullptr, nullptr};
    void *ct3_tw = nullptr;  // twiddle table of the three-step column tiles
    bool pair_tiles = false;  // a pass moves complex64 column pairs (16-byte aligned arrays)
    int final_buf = 0;
    std::vector<Pass> passes;
    std::vector<void *> allocs;
    std::map<std::pair<int, int>, void *> stage_tw;  // (n, e + 64 * prec) -> table
    struct AxisTw {
        int n = 0, e = 0;
        void *table = nullptr;
    } axis_tw[3];
    // real-input plans: passes run in inv_order on INVERSE (C2R), the
    // half-spectrum rows have m + 1 elements, the real array N doubles
    bool real = false;
    long long real_m = 0, real_rows = 0, real_pitch = 0;
    double real_scale = 1.0;
    bool real_pack = false;  // DC and Nyquist columns packed through the column passes
    size_t real_bytes = 0;
    std::vector<int> inv_order;
    bool real_scratch0 = false;  // a column pass uses buffer 0 (P->bytes) as scratch
    EventPair run;  // this plan's exec events
    // per-pass timing
    bool timing = false;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
};

thread_local std::string g_msg;

int map_err(hipError_t e) { return e == hipSuccess ? TOPE_SUCCESS : (int)e; }

FrameworkImpl *fw_of(struct topeFFT *f)
{
    if (!f || !f->impl) return nullptr;
    return static_cast<FrameworkImpl *>(f->impl);
}

bool is_pow2(long long n) { return n >= 1 && (n & (n - 1)) == 0; }
int ilog2ll(long long n)
{
    int l = 0;
    while ((1LL << l) < n) ++l;
    return l;
}

// Reference radix rule and stage count (src/fft1d.c:316-329, :366-369;
// src/fft2d.c:335-354, :435-439), kept in the public plan fields.
void ref_radix(long long n, int *radix, int *stages)
{
    if (!is_pow2(n)) {
        *radix = -1;
        *stages = 0;
        return;
    }
    int l = ilog2ll(n);
    if (l % 3 == 0) {
        *radix = 8;
        *stages = l / 3;
    } else if (l % 2 == 0) {
        *radix = 4;
        *stages = l / 2;
    } else {
        *radix = 2;
        *stages = l;
    }
}

// Timing events skip the system-scope fence (cache writeback/invalidate)
// that a default event performs when it completes: with it, every event
// between two passes costs ~10 us of idle GPU.
// Events that bracket work whose results a host copy reads next keep the
// system-scope fence; timing-only events (per-pass timing) drop it, since a
// fence between back-to-back kernels leaves an idle gap of several us.
hipError_t ensure_event(hipEvent_t *e, bool fenced = true)
{
    if (*e) return hipSuccess;
    return hipEventCreateWithFlags(e, fenced ? hipEventDefault : hipEventDisableSystemFence);
}

const AxisVariant *find_variant_named(const char *name)
{
    int cnt = 0;
    const AxisVariant *v = axis_variants(&cnt);
    for (int i = 0; i < cnt; ++i)
        if (std::strcmp(v[i].name, name) == 0) return &v[i];
    return nullptr;
}

// Variants measured fastest on MI355X for their (length, layout) at the
// benchmark shapes: 2D 4096^2 rows 84 us vs 89-93 us for the others, 4-column
// split tiles 114 us vs 140 us for 2-column tiles; 3D 256^3 columns 78-83 us
// with 32-column split tiles vs 83-87 us (16, 64); 3D 512^3 FP64 Z pass
// 887 us with 16-column split tiles vs 1007 us (8) and 925 us (32), complex64
// Z 443 us with 32-column split tiles vs 701 us (8).  256-point rows in
// 4-row tiles, only for shapes of at least 8192 such blocks: 256^3 202.4 ->
// 197.3 us (complex64 104.7 -> 103.5 us), but 256 x 4096 (1024 blocks, the
// array in cache) 28.3 -> 34.1 us.
struct Preferred {
    const char *name;
    long long min_blocks;
    bool strided_only = false;  // only when input and output elements are both strided
};
constexpr long long kMinBlocks = 256;  // one block per CU at least, when the shape allows
const Preferred kPreferred[] = {
    // the three-step tile's lane groups read rows 16 apart: strided lines only
    // (2D 4096^2 Y pass 103.8-104.2 us against 110-112 us for the one-shot tile)
    {"axis_double_n4096_e16_tb4_col_ct3", kMinBlocks, true},
    {"axis_double_n4096_e16_tb1_row_split", kMinBlocks},
    {"axis_double_n4096_e16_tb4_col_split", kMinBlocks},
    {"axis_double_n2048_e16_tb4_col_split", kMinBlocks},
    {"axis_double_n256_e16_tb32_col_split", kMinBlocks},
    {"axis_double_n512_e16_tb16_col_split", kMinBlocks},
    {"axis_double_n256_e16_tb4_row", 8192},
    // complex64 4096-point columns in pairs (2D 4096^2 Y pass)
    {"axis_float_n4096_e16_tb8_col_ct3", kMinBlocks, true},
    {"axis_float_n512_e16_tb32_col_split", kMinBlocks},
    {"axis_float_n256_e16_tb32_col_split", kMinBlocks},
    {"axis_float_n4096_e16_tb1_row_split", kMinBlocks},
    {"axis_float_n256_e16_tb4_row", 8192},
};

// Variant choice: a measured preference if the line count fills its tile;
// otherwise strided tiles take the widest tile that the line count fills
// (wider tiles give longer contiguous segments per row), preferring the
// full-complex exchange at equal width.  TOPEFFT_VARIANT_<n>_<row|col>=<name>
// overrides (used by tuning sweeps).
// plain_only: one-shot variants only (no even/odd split or three-step
// tile, which need tables of their own), for callers that provide no tables.
// strided_io: the line elements are strided in the input and the output.
// pair_ok: the complex64 three-step tile's conditions hold (adjacent
// columns, an even column count, even strides).
const AxisVariant *select_variant(int prec, int n, bool cfast, long long lines, bool dif2 = false,
                                  long long outer = 1, bool plain_only = false, bool strided_io = false,
                                  bool pair_ok = false)
{
    auto pair_tile = [](const AxisVariant *v) { return v->ct3 && v->prec == PREC_F32; };
    char key[64];
    std::snprintf(key, sizeof key, "TOPEFFT_VARIANT_%d_%s%s", n, cfast ? "col" : "row", dif2 ? "_dif2" : "");
    if (const char *ov = std::getenv(key)) {
        const AxisVariant *v = find_variant_named(ov);
        if (v && v->prec == prec && v->n == n && v->dif2 == dif2 && !(plain_only && (v->dit2 || v->ct3)) &&
            !(pair_tile(v) && !pair_ok))
            return v;
    }
    // blocks a tile width gives for this shape (small shapes: keep the CUs busy)
    auto blocks = [&](const AxisVariant *v) { return outer * ((lines + v->tb - 1) / v->tb); };
    for (const Preferred &pr : kPreferred) {
        const AxisVariant *v = find_variant_named(pr.name);
        if (v && plain_only && (v->dit2 || v->ct3)) continue;
        if (v && pair_tile(v) && !pair_ok) continue;
        if (pr.strided_only && !strided_io) continue;
        if (v && v->prec == prec && v->n == n && v->cfast == cfast && v->dif2 == dif2 && v->tb <= lines &&
            blocks(v) >= pr.min_blocks)
            return v;
    }
    int cnt = 0;
    const AxisVariant *vs = axis_variants(&cnt);
    // the even/odd-split tiles compete only through the preference table;
    // single-column tiles of a strided axis (16-byte row segments) only
    // when every wider tile leaves more than half of the CUs without a
    // block (1024 x 512: 512-point columns in 1-column tiles 8.1 us against
    // 6.4 us in 8-column tiles, 128 blocks; 1D 2^16: 256-point columns 3.5
    // against 4.2 us in 4-column tiles, 64 blocks)
    auto base_match = [&](const AxisVariant *v) {
        return v->prec == prec && v->n == n && v->cfast == cfast && v->dif2 == dif2 && !v->dit2 && !v->ct3;
    };
    long long most_wide = 0;
    {
        int c0 = 0;
        const AxisVariant *v0 = axis_variants(&c0);
        for (int i = 0; i < c0; ++i)
            if (base_match(&v0[i]) && v0[i].tb > 1) most_wide = std::max(most_wide, blocks(&v0[i]));
    }
    const bool allow_single = !cfast || most_wide < kMinBlocks / 2;
    auto matches = [&](const AxisVariant *v) { return base_match(v) && (allow_single || v->tb > 1); };
    // Small shapes: when no tile gives every CU a block, take the tile that
    // gives the most blocks (full-complex exchange first); otherwise only
    // tiles that still give every CU a block compete.
    long long most = 0;
    for (int i = 0; i < cnt; ++i)
        if (matches(&vs[i])) most = std::max(most, blocks(&vs[i]));
    const AxisVariant *best = nullptr;
    if (most > 0 && most < kMinBlocks) {
        for (int i = 0; i < cnt; ++i) {
            const AxisVariant *v = &vs[i];
            if (matches(v) && blocks(v) == most && (!best || (best->split && !v->split))) best = v;
        }
        return best;
    }
    for (int i = 0; i < cnt; ++i) {
        const AxisVariant *v = &vs[i];
        if (!matches(v) || blocks(v) < kMinBlocks) continue;
        if (!best) {
            best = v;
            continue;
        }
        const bool fits = v->tb <= lines, bfits = best->tb <= lines;
        if (fits && !bfits) best = v;
        else if (fits == bfits) {
            if (fits) {
                if (v->tb > best->tb || (v->tb == best->tb && !v->split && best->split)) best = v;
            } else if (v->tb < best->tb) {
                best = v;
            }
        }
    }
    return best;
}

int launch_pass(PlanImpl *P, size_t i, hipStream_t s, int dir);

struct Geo {  // lines (o < O, c < C) of length n, element stride is
    long long O, os, C, cs, is;
};

// Line enumeration for a four-step step: the axis' own lines (o, c) plus
// the new indices of the step (count, input stride, output stride; `fresh`
// marks the one the step's post-twiddle multiplies by).  Two indices go
// into the kernel's (O, C) slots, C taking the smallest output stride so
// strided tiles stay contiguous; a third non-trivial index (never the fresh
// one) becomes repeated launches.  More than three is not expressible.
struct Fold {
    Geo in{1, 0, 1, 0, 0}, out{1, 0, 1, 0, 0};
    long long reps = 1, rep_in = 0, rep_out = 0;
    bool new_in_o = false, new_in_c = false;
    bool ok = true;
};

struct FoldIdx {
    long long count, si, so;
    bool fresh;
};

Fold fold_lines(const Geo &g, std::vector<FoldIdx> extra)
{
    std::vector<FoldIdx> v;
    if (g.O > 1) v.push_back({g.O, g.os, g.os, false});
    if (g.C > 1) v.push_back({g.C, g.cs, g.cs, false});
    for (const FoldIdx &x : extra)
        if (x.count > 1) v.push_back(x);
    Fold r;
    if (v.size() > 3) {
        r.ok = false;
        return r;
    }
    if (v.size() == 3) {  // a non-fresh index becomes repeated launches
        size_t k = v[0].fresh ? 1 : 0;
        r.reps = v[k].count;
        r.rep_in = v[k].si;
        r.rep_out = v[k].so;
        v.erase(v.begin() + (long)k);
    }
    if (v.size() == 2 && v[0].so < v[1].so) std::swap(v[0], v[1]);
    if (v.size() == 2) {
        r.in.O = r.out.O = v[0].count;
        r.in.os = v[0].si;
        r.out.os = v[0].so;
        r.new_in_o = v[0].fresh;
        v.erase(v.begin());
    }
    if (v.size() == 1) {
        r.in.C = r.out.C = v[0].count;
        r.in.cs = v[0].si;
        r.out.cs = v[0].so;
        r.new_in_c = v[0].fresh;
    }
    return r;
}

Fold fold(const Geo &g, long long cnt, long long s_in, long long s_out)
{
    return fold_lines(g, {{cnt, s_in, s_out, true}});
}

class Planner {
   public:
    Planner(PlanImpl *p, hipStream_t s) : P(p), st(s) {}

    hipError_t stage_table(int n, int e, void **out)
    {
        auto key = std::make_pair(n, e + 64 * P->prec);
        auto it = P->stage_tw.find(key);
        if (it != P->stage_tw.end()) {
            *out = it->second;
            return hipSuccess;
        }
        HostSched h = host_sched(n, e);
        void *d = nullptr;
        size_t sz = (size_t)(h.twsize > 0 ? h.twsize : 1) * P->elem;
        hipError_t err = hipMalloc(&d, sz);
        if (err != hipSuccess) return err;
        P->allocs.push_back(d);
        err = gen_stage_twiddles(d, P->prec, n, e, st);
        if (err != hipSuccess) return err;
        P->stage_tw[key] = d;
        *out = d;
        return hipSuccess;
    }

    hipError_t power_table(long long M, long long mult, long long count, void **out)
    {
        void *d = nullptr;
        hipError_t err = hipMalloc(&d, (size_t)count * P->elem);
        if (err != hipSuccess) return err;
        P->allocs.push_back(d);
        err = gen_power_table(d, P->prec, M, mult, count, st);
        *out = d;
        return err;
    }

    // One fused pass: lines (O, C) of length n; input geometry gi in buffer
    // bi, output geometry go in buffer bo.
    int add_pass(int n, const Geo &gi, const Geo &go, int bi, int bo, long long reps, long long rep_in,
                 long long rep_out, const void *pt_lo = nullptr, const void *pt_hi = nullptr, long long pt_o = 0,
                 long long pt_c = 0, long long pt_n = 0, int pt_shift = 0, int axis = -1, long long pair = 0,
                 const void *dif_tw = nullptr)
    {
        const bool dif2 = pair != 0;
        const bool cfast = !(gi.is == 1 && go.is == 1);
        // the complex64 three-step tile moves two adjacent columns as one
        // 16-byte element: every pair must start 16-byte aligned
        auto even = [](long long x) { return (x & 1) == 0; };
        const bool pair_ok = !dif2 && gi.cs == 1 && go.cs == 1 && even(gi.C) && even(gi.is) &&
                             even(go.is) && (gi.O == 1 || (even(gi.os) && even(go.os))) &&
                             (reps == 1 || (even(rep_in) && even(rep_out)));
        const AxisVariant *v = select_variant(P->prec, n, cfast, dif2 ? 2 * gi.C : gi.C, dif2, gi.O * reps, false,
                                              gi.is != 1 && go.is != 1, pair_ok);
        if (!v) return TOPE_ERROR_NO_KERNEL;
        // even/odd-split tiles run the stages of an n/2-point transform and
        // combine the halves with W_n^k, k < n/2 (passed as dif_tw)
        const int tn = v->dit2 ? v->n / 2 : v->n;
        void *tw = nullptr;
        hipError_t err = hipSuccess;
        if (v->ct3) {
            // W_256^j then W_4096^i (j, i < 128), one table per plan
            if (!P->ct3_tw) {
                void *d = nullptr;
                err = hipMalloc(&d, (size_t)CT3_TW * P->elem);
                if (err == hipSuccess) err = gen_power_table(d, P->prec, 256, 1, 128, st);
                if (err == hipSuccess)
                    err = gen_power_table(static_cast<char *>(d) + 128 * P->elem, P->prec, 4096, 1, 128, st);
                if (d) P->allocs.push_back(d);
                if (err != hipSuccess) return map_err(err);
                P->ct3_tw = d;
            }
            tw = P->ct3_tw;
        } else {
            HostSched h = host_sched(tn, v->e);
            if (h.twsize != v->twsize) return TOPE_ERROR_NO_KERNEL;  // host/device layouts disagree
            // (kind stays PASS_AXIS)
            err = stage_table(tn, v->e, &tw);
            if (err != hipSuccess) return map_err(err);
        }
        if (v->dit2) {
            void *wk = nullptr;
            err = power_table(v->n, 1, v->n / 2, &wk);
            if (err != hipSuccess) return map_err(err);
            dif_tw = wk;
        }
        if (axis >= 0 && !P->axis_tw[axis].table) {
            // the axis's exported table is the stage table of (n, e) for
            // every variant (the three-step tile keeps its own beside it)
            void *stw = tw;
            if (v->ct3) {
                err = stage_table(tn, v->e, &stw);
                if (err != hipSuccess) return map_err(err);
            }
            P->axis_tw[axis].n = tn;
            P->axis_tw[axis].e = v->e;
            P->axis_tw[axis].table = stw;
        }
        Pass ps;
        ps.var = v;
        ps.name = v->name;
        if (v->ct3 && v->prec == PREC_F32) P->pair_tiles = true;
        AxisArgs<double> &a = ps.a;
        a.in_os = gi.os;
        a.in_cs = gi.cs;
        a.in_is = gi.is;
        a.out_os = go.os;
        a.out_cs = go.cs;
        a.out_is = go.is;
        const int rows_per_tile = dif2 ? v->tb / 2 : v->tb;
        a.C = (int)gi.C;
        a.tiles = (int)((gi.C + rows_per_tile - 1) / rows_per_tile);
        a.pair = pair;
        a.dif_tw = static_cast<const cpx<double> *>(dif_tw);
        a.tw = static_cast<const cpx<double> *>(tw);
        a.pt_lo = static_cast<const cpx<double> *>(pt_lo);
        a.pt_hi = static_cast<const cpx<double> *>(pt_hi);
        a.pt_o = pt_o;
        a.pt_c = pt_c;
        a.pt_n = pt_n;
        a.pt_shift = pt_shift;
        a.scale = 1.0;
        a.xcd = (v->cfast && !std::getenv("TOPEFFT_NO_XCD_REMAP")) ? 1 : 0;
        // elements the buffers hold (Bluestein's padded rows can exceed the data)
        a.in_elems = a.out_elems = std::max<long long>(P->total, (long long)(P->buf_bytes / P->elem));
        {
            // largest byte offset inside one block's footprint, in and out
            const long long span_in =
                ((long long)(rows_per_tile - 1) * gi.cs + pair + (long long)(n - 1) * gi.is + 1) * (long long)P->elem;
            const long long span_out =
                ((long long)(rows_per_tile - 1) * go.cs + pair + (long long)(n - 1) * go.is + 1) * (long long)P->elem;
            if (span_in >= 0x7fffffffLL || span_out >= 0x7fffffffLL) {
                int cnt = 0, cntb = 0;
                const AxisVariant *base = axis_variants(&cnt);
                const AxisVariant *bigv = axis_variants_big(&cntb);
                v = &bigv[v - base];
                ps.var = v;
            }
        }
        {
            AxisArgs<float> &b = ps.a32;
            b.in_os = a.in_os;
            b.in_cs = a.in_cs;
            b.in_is = a.in_is;
            b.out_os = a.out_os;
            b.out_cs = a.out_cs;
            b.out_is = a.out_is;
            b.C = a.C;
            b.tiles = a.tiles;
            b.tw = static_cast<const cpx<float> *>(tw);
            b.pt_lo = static_cast<const cpx<float> *>(pt_lo);
            b.pt_hi = static_cast<const cpx<float> *>(pt_hi);
            b.pt_o = pt_o;
            b.pt_c = pt_c;
            b.pt_n = pt_n;
            b.pt_shift = pt_shift;
            b.scale = 1.0f;
            b.xcd = a.xcd;
            b.in_elems = b.out_elems = a.in_elems;
            b.pair = pair;
            b.dif_tw = static_cast<const cpx<float> *>(dif_tw);
        }
        const long long grid = gi.O * (long long)a.tiles;
        if (grid <= 0 || grid > 0x7fffffffLL) return TOPE_ERROR_INVALID_SIZE;
        ps.grid = (unsigned)grid;
        ps.in_buf = bi;
        ps.out_buf = bo;
        ps.reps = reps;
        ps.rep_in = rep_in;
        ps.rep_out = rep_out;
        P->passes.push_back(ps);
        return TOPE_SUCCESS;
    }

    // Transform along an axis of length n whose lines are g; the data is in
    // buffer `cur` and ends in buffer `cur` (updated for out-of-place steps).
    int axis(long long n, const Geo &g, int *cur, int axis_id)
    {
        if (n == 1) return TOPE_SUCCESS;
        if (n <= kMaxFused) return add_pass((int)n, g, g, *cur, *cur, 1, 0, 0, nullptr, nullptr, 0, 0, 0, 0, axis_id);
        if (n > kMaxLen) return TOPE_ERROR_INVALID_SIZE;
        if (n > kMaxAxis) return three_step(n, g, cur, axis_id);
        // four-step: i = n1 + N1*n2.  Step 1 transforms n2 (length N2, stride
        // N1*is) per (line, n1), multiplying output k2 by W_n^{n1*k2}, into
        // the other buffer (in place when `flip`); step 2 transforms n1
        // (length N1, stride is) per (line, k2) and writes X[k2 + N2*k1]
        // into the buffer step 1 did not write, so without `flip` the result
        // ends where the data started (no swap, no copy for callers' arrays)
        // and with it in the other buffer.
        const int L = ilog2ll(n);
        const long long N1 = 1LL << ((L + 1) / 2), N2 = n / N1;
        const int lo_bits = (L + 1) / 2;
        void *lo = nullptr, *hi = nullptr;
        hipError_t err = power_table(n, 1, 1LL << lo_bits, &lo);
        if (err != hipSuccess) return map_err(err);
        err = power_table(n, 1LL << lo_bits, n >> lo_bits, &hi);
        if (err != hipSuccess) return map_err(err);

        // step 1 (into the other buffer, same layout): lines are (old o, old c, n1)
        Fold f1 = fold(g, N1, g.is, g.is);
        Geo s1i = f1.in, s1o = f1.out;
        s1i.is = s1o.is = N1 * g.is;
        const int b1 = flip ? *cur : *cur ^ 1;
        int rc = add_pass((int)N2, s1i, s1o, *cur, b1, f1.reps, f1.rep_in, f1.rep_out, lo, hi,
                          f1.new_in_o ? 1 : 0, f1.new_in_c ? 1 : 0, n, lo_bits, axis_id);
        if (rc) return rc;
        // step 2 (transposed): lines are (old o, old c, k2)
        Fold f2 = fold(g, N2, N1 * g.is, g.is);
        Geo s2i = f2.in, s2o = f2.out;
        s2i.is = g.is;
        s2o.is = N2 * g.is;
        rc = add_pass((int)N1, s2i, s2o, b1, b1 ^ 1, f2.reps, f2.rep_in, f2.rep_out);
        *cur = b1 ^ 1;
        return rc;
    }

    // Power-of-two axis longer than 4096^2: n = A B C (each at most 4096),
    // i = a + A b + A B c, k = k_c + C k_b + C B k_a.  Three passes:
    //   1. DFT over c (stride A B is) for each (line, a + A b), output k_c
    //      times W_n^{(a + A b) k_c}, into the other buffer, same layout (in
    //      place when `flip`);
    //   2. in place there: DFT over b (stride A is) for each (line, a, k_c),
    //      output k_b times W_{AB}^{a k_b};
    //   3. DFT over a (stride is) for each (line, k_b, k_c), written to the
    //      other buffer as X[k_c + C k_b + C B k_a]: without `flip` the data
    //      ends where it started.
    // Steps 2 and 3 are the four-step of the inner A B-point transforms.
    // The axis' own lines may use one index (O or C) at most.
    int three_step(long long n, const Geo &g, int *cur, int axis_id)
    {
        const int L = ilog2ll(n);
        const int cb = (L + 2) / 3, bb = (L - cb + 1) / 2, ab = L - cb - bb;
        const long long A = 1LL << ab, B = 1LL << bb, C = 1LL << cb, AB = A * B;
        if (C > kMaxFused || B > kMaxFused || A > kMaxFused) return TOPE_ERROR_INVALID_SIZE;
        const long long is = g.is;
        const Fold f1 = fold_lines(g, {{AB, is, is, true}});
        const Fold f2 = fold_lines(g, {{C, AB * is, AB * is, false}, {A, is, is, true}});
        const Fold f3 = fold_lines(g, {{B, A * is, C * is, false}, {C, AB * is, is, false}});
        if (!f1.ok || !f2.ok || !f3.ok) return TOPE_ERROR_INVALID_SIZE;
        // W_n^x = lo[x & (2^h - 1)] hi[x >> h]; the same for W_AB
        auto split_tables = [&](long long m, void **lo, void **hi, int *h) {
            const int lm = ilog2ll(m);
            *h = (lm + 1) / 2;
            hipError_t err = power_table(m, 1, 1LL << *h, lo);
            if (err == hipSuccess) err = power_table(m, 1LL << *h, m >> *h, hi);
            return err;
        };
        void *lo1 = nullptr, *hi1 = nullptr, *lo2 = nullptr, *hi2 = nullptr;
        int h1 = 0, h2 = 0;
        hipError_t err = split_tables(n, &lo1, &hi1, &h1);
        if (err == hipSuccess) err = split_tables(AB, &lo2, &hi2, &h2);
        if (err != hipSuccess) return map_err(err);
        Geo i1 = f1.in, o1 = f1.out;
        i1.is = o1.is = AB * is;
        const int b1 = flip ? *cur : *cur ^ 1;
        int rc = add_pass((int)C, i1, o1, *cur, b1, f1.reps, f1.rep_in, f1.rep_out, lo1, hi1,
                          f1.new_in_o ? 1 : 0, f1.new_in_c ? 1 : 0, n, h1, axis_id);
        if (rc) return rc;
        Geo i2 = f2.in, o2 = f2.out;
        i2.is = o2.is = A * is;
        rc = add_pass((int)B, i2, o2, b1, b1, f2.reps, f2.rep_in, f2.rep_out, lo2, hi2,
                      f2.new_in_o ? 1 : 0, f2.new_in_c ? 1 : 0, AB, h2);
        if (rc) return rc;
        Geo i3 = f3.in, o3 = f3.out;
        i3.is = is;
        o3.is = C * B * is;
        rc = add_pass((int)A, i3, o3, b1, b1 ^ 1, f3.reps, f3.rep_in, f3.rep_out);
        *cur = b1 ^ 1;
        return rc;
    }

    // Non-power-of-two axis of length n (lines g) by Bluestein's identity
    // (kernels_chirp.hip): chirp pre-multiply of every line into its own
    // zero-padded row of M elements in the other buffer, forward FFT_M over
    // the rows, product with the transformed filter of the direction,
    // inverse FFT_M (both inner FFTs keep their direction whatever the outer
    // one), chirp post-multiply back into the line.  The data stays in
    // buffer *cur.  The filters are transformed here, at plan time, by a
    // one-row FFT_M.
    // 7-smooth axis longer than 4096: the four-step i = n1 + N1 n2 over two
    // mixed-radix passes (n = N1 N2, both at most 4096).  Step 1 transforms
    // n2 (length N2, stride N1 is) for each (line, n1) into the other buffer
    // in the same layout, multiplying output k2 by W_n^{n1 k2}; step 2
    // transforms n1 (length N1, stride is) for each (line, k2) and writes
    // X[k2 + N2 k1] back, so the data ends where it started.
    int mixed_four_step(long long n, const Geo &g, int *cur, int axis_id)
    {
        long long N1 = 0;
        if (!mixed4_split(n, &N1)) return TOPE_ERROR_INVALID_SIZE;
        const long long N2 = n / N1;
        long long S = 1;
        while (S * S < n) ++S;
        void *r1 = nullptr, *r2 = nullptr, *lo = nullptr, *hi = nullptr;
        hipError_t err = power_table(N2, 1, N2, &r1);
        if (err == hipSuccess) err = power_table(N1, 1, N1, &r2);
        if (err == hipSuccess) err = power_table(n, 1, S, &lo);
        if (err == hipSuccess) err = power_table(n, S, (n + S - 1) / S, &hi);
        if (err != hipSuccess) return map_err(err);
        auto mixed_pass = [&](long long len, const Fold &f, long long in_is, long long out_is, int bi, int bo,
                              const void *roots, bool post) {
            Pass ps;
            ps.kind = PASS_MIXED;
            ps.mx = mixed_plan(len, f.in.O, f.in.os, f.in.C, f.in.cs, in_is);
            ps.mx.oos = f.out.os;
            ps.mx.ocs = f.out.cs;
            ps.mx.ois = out_is;
            ps.mx.roots = roots;
            if (post) {
                ps.mx.pt_lo = lo;
                ps.mx.pt_hi = hi;
                ps.mx.pt_n = n;
                ps.mx.pt_o = f.new_in_o ? 1 : 0;
                ps.mx.pt_c = f.new_in_c ? 1 : 0;
                ps.mx.pt_split = (int)S;
            }
            ps.grid = mixed_grid(ps.mx);
            ps.in_buf = bi;
            ps.out_buf = bo;
            ps.reps = f.reps;
            ps.rep_in = f.rep_in;
            ps.rep_out = f.rep_out;
            ps.name = P->prec == PREC_F64 ? "mixed_radix_f64" : "mixed_radix_f32";
            P->passes.push_back(ps);
        };
        // step 1: lines (old o, old c, n1)
        mixed_pass(N2, fold(g, N1, g.is, g.is), N1 * g.is, N1 * g.is, *cur, *cur ^ 1, r1, true);
        // step 2: lines (old o, old c, k2)
        mixed_pass(N1, fold(g, N2, N1 * g.is, g.is), g.is, N2 * g.is, *cur ^ 1, *cur, r2, false);
        if (axis_id >= 0 && !P->axis_tw[axis_id].table) {
            P->axis_tw[axis_id].n = (int)N2;
            P->axis_tw[axis_id].e = 0;
            P->axis_tw[axis_id].table = r1;
        }
        return TOPE_SUCCESS;
    }

    // Axis of 7-smooth length n <= 4096 that is not a power of two: one
    // in-place mixed-radix pass (kernels_mixed.hip), twiddles and radix-3/5/7
    // constants from a table of W_n^i.
    int mixed_axis(long long n, const Geo &g, int cur, int axis_id)
    {
        void *roots = nullptr;
        hipError_t err = power_table(n, 1, n, &roots);
        if (err != hipSuccess) return map_err(err);
        Pass ps;
        ps.kind = PASS_MIXED;
        ps.mx = mixed_plan(n, g.O, g.os, g.C, g.cs, g.is);
        ps.mx.roots = roots;
        ps.grid = mixed_grid(ps.mx);
        ps.in_buf = ps.out_buf = cur;
        ps.name = P->prec == PREC_F64 ? "mixed_radix_f64" : "mixed_radix_f32";
        P->passes.push_back(ps);
        if (axis_id >= 0 && !P->axis_tw[axis_id].table) {
            P->axis_tw[axis_id].n = (int)n;
            P->axis_tw[axis_id].e = 0;
            P->axis_tw[axis_id].table = roots;
        }
        return TOPE_SUCCESS;
    }

    static long long chirp_len(long long n)
    {
        long long M = 1;
        while (M < 2 * n - 1) M <<= 1;
        return M;
    }

    int bluestein_axis(long long n, const Geo &g, int *cur)
    {
        const long long M = chirp_len(n), lines = g.O * g.C;
        const int data = *cur, scratch = *cur ^ 1;
        const size_t tb = (size_t)M * 16;
        void *tab[2] = {nullptr, nullptr};
        for (int d = 0; d < 2; ++d) {
            hipError_t err = hipMalloc(&tab[d], tb);
            if (err != hipSuccess) return map_err(err);
            P->allocs.push_back(tab[d]);
        }
        auto chirp_pass = [&](int op, int bi, int bo) {
            Pass ps;
            ps.kind = PASS_CHIRP;
            ps.chirp_op = op;
            ps.chirp_n = n;
            ps.chirp_M = M;
            ps.chirp_g = ChirpGeo{g.O, g.os, g.C, g.cs, g.is};
            ps.chirp_tab[0] = tab[0];
            ps.chirp_tab[1] = tab[1];
            ps.in_buf = bi;
            ps.out_buf = bo;
            ps.name = op == CHIRP_PRE ? "chirp_pre" : op == CHIRP_MUL ? "chirp_mul" : "chirp_post";
            P->passes.push_back(ps);
        };
        // the rows live in the scratch buffer; an inner FFT longer than one
        // pass uses the data buffer as its intermediate: the pre-multiply
        // has read every line of the axis (the whole array) by then
        chirp_pass(CHIRP_PRE, data, scratch);
        int rc_ = scratch;
        const size_t f0 = P->passes.size();
        int rc = axis(M, Geo{1, 0, lines, M, 1}, &rc_, -1);
        if (rc) return rc;
        for (size_t i = f0; i < P->passes.size(); ++i) P->passes[i].fixed_dir = FORWARD;
        chirp_pass(CHIRP_MUL, rc_, rc_);
        const size_t f1 = P->passes.size();
        rc = axis(M, Geo{1, 0, lines, M, 1}, &rc_, -1);
        if (rc) return rc;
        for (size_t i = f1; i < P->passes.size(); ++i) P->passes[i].fixed_dir = INVERSE;
        chirp_pass(CHIRP_POST, rc_, data);

        // transformed filters: b into the scratch buffer, a one-row forward
        // FFT_M (planned, run, then dropped from the pass list), copy out
        const size_t g0 = P->passes.size();
        int fc = scratch;
        rc = axis(M, Geo{1, 0, 1, M, 1}, &fc, -1);
        if (rc) return rc;
        const size_t g1 = P->passes.size();
        for (int d = 0; d < 2 && !rc; ++d) {
            hipError_t err =
                launch_chirp(CHIRP_FILTER, nullptr, P->buf[scratch], nullptr, ChirpGeo{1, 0, 1, 0, 1}, n, M, 1.0, d, st);
            for (size_t i = g0; i < g1 && err == hipSuccess; ++i)
                if (launch_pass(P, i, st, FORWARD)) err = hipErrorUnknown;
            if (err == hipSuccess) err = hipMemcpyAsync(tab[d], P->buf[fc], tb, hipMemcpyDeviceToDevice, st);
            if (err != hipSuccess) rc = map_err(err);
        }
        P->passes.resize(g0);
        return rc;
    }

    // DFT-16 operand constants of the MFMA passes (one table per plan).
    hipError_t mfma_consts(const MfmaConsts **out)
    {
        if (!kc_) {
            void *d = nullptr;
            hipError_t err = hipMalloc(&d, mfma_constants_bytes());
            if (err != hipSuccess) return err;
            P->allocs.push_back(d);
            err = gen_mfma_constants(d, st);
            if (err != hipSuccess) return err;
            kc_ = static_cast<const MfmaConsts *>(d);
        }
        *out = kc_;
        return hipSuccess;
    }

    // Axis pass on the matrix cores when a kernel covers the geometry
    // (complex64 storage): contiguous rows of n = 256 .. 4096 back to back
    // (row kernels), or unit-stride column tiles of 16 lines with n = 256,
    // 512, 1024 (tile kernel).  Otherwise the FP32 VALU axis pass.
    int axis_mp(long long n, const Geo &g, int *cur, int axis_id, int mode)
    {
        if (mode == MP_VALU || n < 256 || n > 4096) return axis(n, g, cur, axis_id);
        const bool rows = g.is == 1 && g.cs == n && (g.O == 1 || g.os == g.C * n) && mfma_supported((int)n);
        const bool cols = g.cs == 1 && g.C % 16 == 0 && mfma_tile_supported((int)n) &&
                          (15 + (n - 1) * g.is) * (long long)P->elem < 0x7fffffffLL;
        if (!rows && !cols) return axis(n, g, cur, axis_id);
        void *tw = nullptr;
        hipError_t err = stage_table((int)n, 16, &tw);
        const MfmaConsts *kc = nullptr;
        if (err == hipSuccess) err = mfma_consts(&kc);
        if (err != hipSuccess) return map_err(err);
        if (axis_id >= 0 && !P->axis_tw[axis_id].table) {
            P->axis_tw[axis_id].n = (int)n;
            P->axis_tw[axis_id].e = 16;
            P->axis_tw[axis_id].table = tw;
        }
        Pass ps;
        ps.mfma_mode = mode;
        ps.mfma_n = (int)n;
        ps.in_buf = ps.out_buf = *cur;
        if (rows) {
            ps.kind = PASS_MFMA;
            ps.m.count = g.O * g.C;
            ps.m.tw = static_cast<const cpx<float> *>(tw);
            ps.m.k = kc;
            ps.m.scale = 1.0f;
            ps.m.in_elems = ps.m.out_elems = P->total;
            ps.name = mode == MP_BF16X3   ? "mfma_rows_bf16x3"
                      : mode == MP_BF16X6 ? "mfma_rows_bf16x6"
                      : mode == MP_FP16X3 ? "mfma_rows_fp16x3"
                                          : "mfma_rows_fp32";
        } else {
            ps.kind = PASS_MFMA_TILE;
            MfmaTileArgs &t = ps.mt;
            t.O = g.O;
            t.os = g.os;
            t.tiles = g.C / 16;
            t.is = g.is;
            t.tw = static_cast<const cpx<float> *>(tw);
            t.k = kc;
            t.scale = 1.0f;
            t.in_elems = t.out_elems = P->total;
            t.xcd = std::getenv("TOPEFFT_NO_XCD_REMAP") ? 0 : 1;
            const long long grid = g.O * t.tiles;
            if (grid <= 0 || grid > 0x7fffffffLL) return TOPE_ERROR_INVALID_SIZE;
            ps.grid = (unsigned)grid;
            ps.name = mode == MP_BF16X3   ? "mfma_cols_bf16x3"
                      : mode == MP_BF16X6 ? "mfma_cols_bf16x6"
                      : mode == MP_FP16X3 ? "mfma_cols_fp16x3"
                                          : "mfma_cols_fp32";
        }
        P->passes.push_back(ps);
        return TOPE_SUCCESS;
    }

    // 2D with a long Y axis (ny <= 4096): the X pass takes rows y and
    // y + ny/2 together and ends with the radix-2 DIF step of Y between
    // them; the Y pass then runs two independent ny/2-point transforms per
    // column (rows 0..ny/2-1 -> even outputs, the rest -> odd outputs) and
    // writes them interleaved into natural order, out of place.  Per-block
    // column tiles are half as large, so two blocks share a CU and overlap
    // their memory and compute phases.
    int plan_2d_dif2(long long nx, long long ny, int *cur)
    {
        const long long half = ny / 2;
        void *dtw = nullptr;
        hipError_t err = power_table(ny, 1, half, &dtw);
        if (err != hipSuccess) return map_err(err);
        const Geo rows{1, 0, half, nx, 1};
        int rc = add_pass((int)nx, rows, rows, *cur, *cur, 1, 0, 0, nullptr, nullptr, 0, 0, 0, 0, 0, half * nx, dtw);
        if (rc) return rc;
        const Geo gin{2, half * nx, nx, 1, nx}, gout{2, nx, nx, 1, 2 * nx};
        rc = add_pass((int)half, gin, gout, *cur, *cur ^ 1, 1, 0, 0, nullptr, nullptr, 0, 0, 0, 0, 1);
        if (rc) return rc;
        *cur ^= 1;
        return TOPE_SUCCESS;
    }

    // true: an axis longer than 4096 may leave its data in the other buffer
    // (complex plans swap their buffers at the end); its first pass then
    // runs in place, which measured 102 us against 195 us for the 2^24
    // four-step's first step (a strided pass writing 64-byte segments into
    // a buffer it did not just read loses the Infinity Cache's merging of
    // the two halves of each 128-byte line).  The real-input plans keep
    // each axis's data where it started.
    bool flip = false;

   private:
    PlanImpl *P;
    hipStream_t st;
    const MfmaConsts *kc_ = nullptr;
};

// 2D plans whose Y axis is at least this long (and fits one pass) use
// plan_2d_dif2.  Off by default: on MI355X the out-of-place half-length Y
// pass measured 181 us against 114 us for the in-place 4096-point column
// pass at 4096^2 (in-place passes re-read lines the previous pass just wrote
// while they are still in the Infinity Cache); TOPEFFT_DIF2_MIN enables it.
long long dif2_min_rows()
{
    if (const char *e = std::getenv("TOPEFFT_DIF2_MIN")) return std::atoll(e);
    return 1LL << 62;
}

hipError_t time_start(EventPair &ev, hipStream_t s, bool fenced = true)
{
    hipError_t e = ensure_event(&ev.start, fenced);
    if (e == hipSuccess) e = ensure_event(&ev.stop, fenced);
    if (e == hipSuccess) e = hipEventRecord(ev.start, s);
    return e;
}

hipError_t time_stop(EventPair &ev, hipStream_t s, tope_ulong *acc_ns)
{
    hipError_t e = hipEventRecord(ev.stop, s);
    if (e != hipSuccess) return e;
    ev.recorded = true;
    if (acc_ns) {
        e = hipEventSynchronize(ev.stop);
        if (e != hipSuccess) return e;
        float ms = 0;
        e = hipEventElapsedTime(&ms, ev.start, ev.stop);
        if (e != hipSuccess) return e;
        *acc_ns += (tope_ulong)((double)ms * 1e6);
    }
    return hipSuccess;
}

hipEvent_t pool_event(PlanImpl *P)
{
    if (P->pool_used == P->pool.size()) {
        hipEvent_t e = nullptr;
        if (ensure_event(&e, false) != hipSuccess) return nullptr;
        P->pool.push_back(e);
    }
    return P->pool[P->pool_used++];
}

// Launch pass i of a plan on stream s.
int launch_pass(PlanImpl *P, size_t i, hipStream_t s, int dir)
{
    const Pass &ps = P->passes[i];
    const int sign = ((ps.fixed_dir >= 0 ? ps.fixed_dir : dir) == FORWARD) ? -1 : 1;
    const double sc = (dir == INVERSE) ? 1.0 / (double)P->norm : 1.0;
    const char *ib = static_cast<const char *>(P->buf[ps.in_buf]);
    char *ob = static_cast<char *>(P->buf[ps.out_buf]);
    const long long eb = (long long)P->elem;
    if (ps.kind == PASS_AXIS) {
        if (ps.var->prec == PREC_F64) {
            AxisArgs<double> a = ps.a;
            a.scale = ps.last ? sc : 1.0;
            a.in_elems -= ps.base;
            a.out_elems -= ps.base;
            for (long long r = 0; r < ps.reps; ++r)
                ps.var->fn(ib + (ps.base + r * ps.rep_in) * eb, ob + (ps.base + r * ps.rep_out) * eb, &a, ps.grid, s,
                           sign);
        } else {
            AxisArgs<float> a = ps.a32;
            a.scale = ps.last ? (float)sc : 1.0f;
            a.in_elems -= ps.base;
            a.out_elems -= ps.base;
            for (long long r = 0; r < ps.reps; ++r)
                ps.var->fn(ib + (ps.base + r * ps.rep_in) * eb, ob + (ps.base + r * ps.rep_out) * eb, &a, ps.grid, s,
                           sign);
        }
    } else if (ps.kind == PASS_MFMA) {
        MfmaArgs m = ps.m;
        m.scale = ps.last ? (float)sc : 1.0f;  // 1/N is a power of two: exact in FP32
        hipError_t err = launch_mfma_rows(ps.mfma_mode, ps.mfma_n, ib, ob, m, sign, s);
        if (err != hipSuccess) return map_err(err);
    } else if (ps.kind == PASS_MIXED) {
        MixedArgs m = ps.mx;
        m.scale = ps.last ? sc : 1.0;
        for (long long r = 0; r < ps.reps; ++r) {
            hipError_t err = launch_mixed(P->prec, ib + r * ps.rep_in * eb, ob + r * ps.rep_out * eb, m, sign, s);
            if (err != hipSuccess) return map_err(err);
        }
    } else if (ps.kind == PASS_CHIRP) {
        // 1/M undoes the unnormalised inner inverse FFT; the plan's 1/N of
        // INVERSE rides on its last pass as for the axis passes
        const int inv = dir == INVERSE ? 1 : 0;
        const double post = (1.0 / (double)ps.chirp_M) * (ps.last ? sc : 1.0);
        hipError_t err = launch_chirp(ps.chirp_op, ib, ob, ps.chirp_tab[inv], ps.chirp_g, ps.chirp_n, ps.chirp_M,
                                      post, inv, s);
        if (err != hipSuccess) return map_err(err);
    } else if (ps.kind == PASS_REAL) {
        const bool fwd = dir == FORWARD;
        const void *src = P->buf[fwd ? ps.in_buf : ps.out_buf];
        void *dst = P->buf[fwd ? ps.out_buf : ps.in_buf];
        hipError_t err =
            ps.real_stw ? launch_real_rows(P->prec, fwd ? 0 : 1, src, dst, ps.real_stw, ps.real_w, P->real_m, P->real_pitch,
                                           P->real_rows, P->real_scale, P->real_pack ? 1 : 0, s)
                        : launch_real_pass(P->prec, fwd ? 0 : 1, src, dst, ps.real_w, P->real_m, P->real_pitch, P->real_rows,
                                           P->real_scale, P->real_pack ? 1 : 0, s);
        if (err != hipSuccess) return map_err(err);
    } else if (ps.kind == PASS_RPACK) {
        hipError_t err = launch_real_pack(P->prec, dir == FORWARD ? 0 : 1, P->buf[ps.in_buf], P->real_m, P->real_pitch,
                                          P->dims[1], P->dims[2], s);
        if (err != hipSuccess) return map_err(err);
    } else if (ps.kind == PASS_MFMA_TILE) {
        MfmaTileArgs m = ps.mt;
        m.scale = ps.last ? (float)sc : 1.0f;
        hipError_t err = launch_mfma_tile(ps.mfma_mode, ps.mfma_n, ib, ob, m, ps.grid, sign, s);
        if (err != hipSuccess) return map_err(err);
    } else {
        hipError_t err = launch_naive_dft(P->prec, ib, ob, ps.dft_roots, ps.dft_n, sign, ps.last ? sc : 1.0, s);
        if (err != hipSuccess) return map_err(err);
    }
    return map_err(hipGetLastError());
}

// Enqueue all passes of one transform on stream s.
int enqueue(PlanImpl *P, hipStream_t s, int dir)
{
    if (dir != FORWARD && dir != INVERSE) return TOPE_ERROR_INVALID_ARG;
    const bool rev = dir == INVERSE && !P->inv_order.empty();
    for (size_t n = 0; n < P->passes.size(); ++n) {
        const size_t i = rev ? (size_t)P->inv_order[n] : n;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (P->timing) {
            e0 = pool_event(P);
            e1 = pool_event(P);
            if (!e0 || !e1) return TOPE_ERROR_UNSUPPORTED;
            (void)hipEventRecord(e0, s);
        }
        int rc = launch_pass(P, i, s, dir);
        if (rc) return rc;
        if (P->timing) {
            (void)hipEventRecord(e1, s);
            P->pending.push_back({(int)i, {e0, e1}});
        }
    }
    if (P->final_buf == 1) std::swap(P->buf[0], P->buf[1]);
    return TOPE_SUCCESS;
}

template <typename Plan>
void sync_fields(Plan *t, PlanImpl *P)
{
    t->data = P->buf[0];
    t->scratch = P->buf[1];
}

PlanImpl *impl_of(void *p) { return static_cast<PlanImpl *>(p); }

enum { SHAPE_BATCH = 4 };

// Common plan construction.  shape 1/2/3: an nx*ny*nz FP64 array (the
// reference plans); SHAPE_BATCH: ny rows of length nx in precision `prec`
// run by `mode`.  Returns a status; on success *out owns the device buffers
// and twiddles and the data has been uploaded.
int build_plan(struct topeFFT *f, int shape, long long nx, long long ny, long long nz, int prec, int mode,
               const void *d, PlanImpl **out, tope_ulong *pre_ns, tope_ulong *mem_ns)
{
    *out = nullptr;
    FrameworkImpl *fw = fw_of(f);
    if (!fw) return TOPE_ERROR_NOT_INITIALIZED;
    if (nx < 1 || ny < 1 || nz < 1) return TOPE_ERROR_INVALID_SIZE;
    const long long total = nx * ny * nz;
    if (total > (1LL << 34)) return TOPE_ERROR_INVALID_SIZE;
    // complex64 axes: powers of two, or 7-smooth up to 4096 (mixed-radix
    // pass) on the vector ALUs; FP64 also takes Bluestein for the others
    // (a batch transforms only its rows: ny is the count)
    long long split_ = 0;
    auto mixed4 = [&](long long n) { return mixed4_split(n, &split_); };
    auto f32_axis = [&](long long n) { return is_pow2(n) || mixed_supported(n) || mixed4(n); };
    const bool batch = shape == SHAPE_BATCH;
    const bool f32_ok = f32_axis(nx) && (batch || (f32_axis(ny) && f32_axis(nz)));
    const bool all_pow2 = is_pow2(nx) && (batch || (is_pow2(ny) && is_pow2(nz)));
    if ((shape == 2 || shape == 3 || batch) && prec != PREC_F64 && !f32_ok) return TOPE_ERROR_UNSUPPORTED;
    if (mode != MP_VALU && !all_pow2) return TOPE_ERROR_UNSUPPORTED;
    if (prec != PREC_F64 && prec != PREC_F32) return TOPE_ERROR_INVALID_ARG;
    if (mode != MP_VALU && (prec != PREC_F32 || (shape != SHAPE_BATCH && shape != 2 && shape != 3)))
        return TOPE_ERROR_UNSUPPORTED;
    if (prec == PREC_F32 && shape == 1 && !is_pow2(nx)) return TOPE_ERROR_UNSUPPORTED;
    if (mode != MP_VALU && mode != MP_BF16X3 && mode != MP_BF16X6 && mode != MP_FP32 && mode != MP_FP16X3)
        return TOPE_ERROR_INVALID_ARG;
    if (mode != MP_VALU && shape == SHAPE_BATCH && !mfma_supported((int)nx)) return TOPE_ERROR_UNSUPPORTED;
    if (!d) return TOPE_ERROR_INVALID_ARG;

    PlanImpl *P = new (std::nothrow) PlanImpl();
    if (!P) return TOPE_ERROR_UNSUPPORTED;
    P->fw = fw;
    P->prec = prec;
    P->elem = prec == PREC_F64 ? 16 : 8;
    P->ndim = shape;
    P->dims[0] = nx;
    P->dims[1] = ny;
    P->dims[2] = nz;
    P->total = total;
    P->norm = shape == SHAPE_BATCH ? nx : total;
    P->bytes = (size_t)total * P->elem;
    // Non-power-of-two axes by Bluestein (1D above kChirpMin, every one in
    // 2D/3D): the rows of M elements per line need a larger scratch buffer
    bool chirpz = false;
    size_t buf_elems = (size_t)total;
    {
        const long long dims3[3] = {nx, ny, nz};
        for (int ax = 0; ax < (shape == SHAPE_BATCH ? 1 : 3); ++ax) {
            const long long n = dims3[ax];
            if (is_pow2(n) && n > kMaxLen) {  // refused before anything is allocated or read
                delete P;
                return TOPE_ERROR_INVALID_SIZE;
            }
            if (n == 1 || is_pow2(n) || mixed_supported(n) || mixed4(n) || (shape == 1 && n <= kChirpMin)) continue;
            // M above 4096 runs its FFTs as the four-step, through the
            // data buffer (consumed by the chirp pre-multiply by then)
            const long long M = Planner::chirp_len(n);
            if (M > kMaxLen) {
                delete P;
                return TOPE_ERROR_INVALID_SIZE;
            }
            chirpz = true;
            buf_elems = std::max(buf_elems, (size_t)(total / n * M));
        }
    }
    P->buf_bytes = buf_elems * P->elem;
    hipStream_t s = fw->stream;

    auto fail = [&](int rc) {
        for (void *a : P->allocs) (void)hipFree(a);
        if (P->buf[0]) (void)hipFree(P->buf[0]);
        if (P->buf[1]) (void)hipFree(P->buf[1]);
        delete P;
        return rc;
    };

    hipError_t err = hipMalloc(&P->buf[0], P->buf_bytes);
    if (err != hipSuccess) return fail(map_err(err));
    if (chirpz) {
        err = hipMalloc(&P->buf[1], P->buf_bytes);
        if (err != hipSuccess) return fail(map_err(err));
    }
    // one axis: the power-of-two passes; 7-smooth lengths up to 4096 by the
    // mixed-radix pass; Bluestein for the others
    auto any_axis = [&](Planner &pl, long long n, const Geo &g, int *cur, int id) {
        if (n > 1 && !is_pow2(n)) {
            if (mixed_supported(n)) return pl.mixed_axis(n, g, *cur, id);
            if (mixed4(n)) return pl.mixed_four_step(n, g, cur, id);
            return pl.bluestein_axis(n, g, cur);
        }
        return pl.axis(n, g, cur, id);
    };
    const bool nonpow2 = !is_pow2(nx) || (shape >= 2 && !is_pow2(ny)) || (shape >= 3 && !is_pow2(nz));

    EventPair pre;
    err = time_start(pre, s);
    if (err != hipSuccess) return fail(map_err(err));

    Planner pl(P, s);
    pl.flip = true;  // the plan swaps its buffers when the data ends in buffer 1
    int cur = 0, rc = TOPE_SUCCESS;
    if (shape == SHAPE_BATCH && mode != MP_VALU) {
        // one MFMA launch over all rows, in place
        void *tw = nullptr, *kc = nullptr;
        err = pl.stage_table((int)nx, 16, &tw);
        if (err == hipSuccess) err = hipMalloc(&kc, mfma_constants_bytes());
        if (err != hipSuccess) return fail(map_err(err));
        P->allocs.push_back(kc);
        err = gen_mfma_constants(kc, s);
        if (err != hipSuccess) return fail(map_err(err));
        Pass ps;
        ps.kind = PASS_MFMA;
        ps.mfma_mode = mode;
        ps.mfma_n = (int)nx;
        ps.m.count = ny;
        ps.m.tw = static_cast<const cpx<float> *>(tw);
        ps.m.k = static_cast<const MfmaConsts *>(kc);
        ps.m.scale = 1.0f;
        ps.m.in_elems = ps.m.out_elems = total;
        ps.name = mode == MP_BF16X3   ? "mfma_rows_bf16x3"
                  : mode == MP_BF16X6 ? "mfma_rows_bf16x6"
                  : mode == MP_FP16X3 ? "mfma_rows_fp16x3"
                                      : "mfma_rows_fp32";
        P->passes.push_back(ps);
        P->axis_tw[0].n = (int)nx;
        P->axis_tw[0].e = 16;
        P->axis_tw[0].table = tw;
    } else if (shape == SHAPE_BATCH) {
        rc = any_axis(pl, nx, Geo{1, 0, ny, nx, 1}, &cur, 0);
    } else if (nonpow2 && !(shape == 1 && !chirpz && !mixed_supported(nx) && !mixed4(nx))) {
        rc = any_axis(pl, nx, Geo{1, 0, ny * nz, nx, 1}, &cur, 0);
        if (!rc && shape >= 2) rc = any_axis(pl, ny, Geo{nz, nx * ny, nx, 1, nx}, &cur, 1);
        if (!rc && shape >= 3) rc = any_axis(pl, nz, Geo{1, 0, nx * ny, 1, nx * ny}, &cur, 2);
    } else if (shape == 1 && !is_pow2(nx)) {
        // short non-power-of-two 1D: GPU DFT with exactly reduced phases
        void *roots = nullptr;
        err = hipMalloc(&roots, (size_t)nx * 16);
        if (err != hipSuccess) return fail(map_err(err));
        P->allocs.push_back(roots);
        err = gen_power_table(roots, PREC_F64, nx, 1, nx, s);
        if (err != hipSuccess) return fail(map_err(err));
        Pass ps;
        ps.kind = PASS_DFT;
        ps.dft_n = nx;
        ps.dft_roots = roots;
        ps.in_buf = 0;
        ps.out_buf = 1;
        ps.name = "naive_dft_f64";
        P->passes.push_back(ps);
        cur = 1;
        P->axis_tw[0].n = (int)nx;
        P->axis_tw[0].e = 0;
        P->axis_tw[0].table = roots;
    } else if (mode != MP_VALU) {
        // complex64 on the matrix cores where a kernel covers the axis
        rc = pl.axis_mp(nx, Geo{1, 0, ny * nz, nx, 1}, &cur, 0, mode);
        if (!rc) rc = pl.axis_mp(ny, Geo{nz, nx * ny, nx, 1, nx}, &cur, 1, mode);
        if (!rc && shape >= 3) rc = pl.axis_mp(nz, Geo{1, 0, nx * ny, 1, nx * ny}, &cur, 2, mode);
    } else if (shape == 2 && nx <= kMaxFused && ny <= kMaxFused && ny >= 4 && ny >= dif2_min_rows()) {
        rc = pl.plan_2d_dif2(nx, ny, &cur);
    } else {
        // X: rows; Y: columns of each plane; Z: columns of the whole volume
        rc = pl.axis(nx, Geo{1, 0, ny * nz, nx, 1}, &cur, 0);
        if (!rc && shape >= 2) rc = pl.axis(ny, Geo{nz, nx * ny, nx, 1, nx}, &cur, 1);
        if (!rc && shape >= 3) rc = pl.axis(nz, Geo{1, 0, nx * ny, 1, nx * ny}, &cur, 2);
    }
    if (rc) return fail(rc);
    if (!P->passes.empty()) P->passes.back().last = true;
    P->final_buf = cur;
    bool need_scratch = false;
    for (const Pass &ps : P->passes)
        if (ps.in_buf == 1 || ps.out_buf == 1) need_scratch = true;
    if (need_scratch && !P->buf[1]) {
        err = hipMalloc(&P->buf[1], P->buf_bytes);
        if (err != hipSuccess) return fail(map_err(err));
    }
    err = time_stop(pre, s, pre_ns);
    if (err != hipSuccess) return fail(map_err(err));

    EventPair up;
    err = time_start(up, s);
    if (err == hipSuccess) err = hipMemcpyAsync(P->buf[0], d, P->bytes, hipMemcpyHostToDevice, s);
    if (err == hipSuccess) err = time_stop(up, s, mem_ns);
    (void)hipEventDestroy(pre.start);
    (void)hipEventDestroy(pre.stop);
    (void)hipEventDestroy(up.start);
    (void)hipEventDestroy(up.stop);
    if (err != hipSuccess) return fail(map_err(err));
    *out = P;
    return TOPE_SUCCESS;
}


void free_plan(PlanImpl *P)
{
    if (!P) return;
    if (P->fw && P->fw->stream) (void)hipStreamSynchronize(P->fw->stream);
    for (void *a : P->allocs) (void)hipFree(a);
    if (P->buf[0]) (void)hipFree(P->buf[0]);
    if (P->buf[1]) (void)hipFree(P->buf[1]);
    for (hipEvent_t e : P->pool) (void)hipEventDestroy(e);
    if (P->run.start) (void)hipEventDestroy(P->run.start);
    if (P->run.stop) (void)hipEventDestroy(P->run.stop);
    delete P;
}

// Real-input plan: x (even, x/2 <= 4096) by y by z (y, z <= 4096), all powers
// of two.  buf[0] holds the real array (x y z doubles, read as (x/2) y z
// complex), buf[1] the half spectrum: rows of x/2 + 1 elements at a pitch of
// x/2 + 8 elements (128-byte aligned rows for the strided passes).  FORWARD
// (R2C): m-point rows in buf[0], post-processing into buf[1], then Y and Z
// in buf[1] over the columns 0..m (the Nyquist column m in a tile of its
// own that runs first); INVERSE (C2R) runs the same passes in reverse order, the
// pre-processing scaled by 1/(m y z) so the result carries 1/(x y z).
int build_real_plan(struct topeFFT *f, long long nx, long long ny, long long nz, int prec, const void *real,
                    PlanImpl **out, tope_ulong *pre_ns, tope_ulong *mem_ns)
{
    FrameworkImpl *fw = fw_of(f);
    if (!fw) return TOPE_ERROR_NOT_INITIALIZED;
    if (!real) return TOPE_ERROR_INVALID_ARG;
    if (nx < 2 || (nx & 1) || ny < 1 || nz < 1) return TOPE_ERROR_INVALID_SIZE;
    if (prec != PREC_F64 && prec != PREC_F32) return TOPE_ERROR_INVALID_ARG;
    // each axis (x/2 for the rows) a power of two (up to 2^30: one pass up to
    // 4096, the four-step up to 2^24, three passes above), or 7-smooth up to
    // 4096 (one mixed-radix pass) or above when it splits into two factors of
    // at most 4096 (two mixed-radix passes)
    auto covered = [](long long n) {
        long long n1 = 0;
        return is_pow2(n) ? n <= kMaxLen : (mixed_supported(n) || mixed4_split(n, &n1));
    };
    if (!covered(nx / 2) || !covered(ny) || !covered(nz)) return TOPE_ERROR_UNSUPPORTED;
    const long long m = nx / 2, rows = ny * nz;
    if ((m + 1) * rows > (1LL << 33)) return TOPE_ERROR_INVALID_SIZE;

    PlanImpl *P = new (std::nothrow) PlanImpl();
    if (!P) return TOPE_ERROR_UNSUPPORTED;
    P->fw = fw;
    P->prec = prec;
    P->elem = prec == PREC_F64 ? 16 : 8;
    P->ndim = nz > 1 ? 3 : (ny > 1 ? 2 : 1);
    P->dims[0] = nx;
    P->dims[1] = ny;
    P->dims[2] = nz;
    const long long pitch = m + 8;
    P->total = pitch * rows;  // elements of the larger (spectrum) buffer
    P->norm = 1;              // axis passes unscaled; the C2R pre-pass scales
    P->bytes = (size_t)P->total * P->elem;
    P->real = true;
    P->real_m = m;
    P->real_rows = rows;
    P->real_pitch = pitch;
    P->real_scale = 1.0 / ((double)m * (double)rows);
    P->real_bytes = (size_t)(nx * rows) * (P->elem / 2);
    hipStream_t s = fw->stream;

    auto fail = [&](int rc) {
        free_plan(P);
        return rc;
    };
    hipError_t err = hipMalloc(&P->buf[0], P->bytes);
    if (err == hipSuccess) err = hipMalloc(&P->buf[1], P->bytes);
    if (err != hipSuccess) return fail(map_err(err));

    EventPair pre;
    err = time_start(pre, s);
    if (err != hipSuccess) return fail(map_err(err));
    Planner pl(P, s);
    int cur = 0;
    // X: m-point complex rows of the real array and the half-spectrum step:
    // one fused pass from buffer 0 into buffer 1 where a kernel covers m,
    // else the row transform in place (one pass, or the two four-step passes
    // above 4096 with buffer 1 as their scratch) and the step on its own.
    // group_end[g] = passes.size() after group g (rows, step, Y, Z): INVERSE
    // runs the groups in reverse order, the passes of a group in order.
    std::vector<size_t> group_end;
    const bool fused = real_rows_supported(m);
    const Geo rows_geo{1, 0, rows, m, 1};
    // one axis of the real plan: power of two, 7-smooth in one mixed-radix
    // pass, or the mixed-radix four-step (the data ends where it started)
    auto real_axis = [&](long long n, const Geo &g, int id) {
        if (is_pow2(n)) return pl.axis(n, g, &cur, id);
        if (mixed_supported(n)) return pl.mixed_axis(n, g, cur, id);
        return pl.mixed_four_step(n, g, &cur, id);
    };
    int rc = fused ? (int)TOPE_SUCCESS : real_axis(m, rows_geo, 0);
    group_end.push_back(P->passes.size());
    if (!rc && cur != 0) rc = TOPE_ERROR_UNSUPPORTED;
    if (!rc) {
        void *w = nullptr;
        err = pl.power_table(nx, 1, m + 1, &w);
        if (err != hipSuccess) return fail(map_err(err));
        Pass ps;
        ps.kind = PASS_REAL;
        ps.real_w = w;
        ps.in_buf = 0;
        ps.out_buf = 1;
        ps.name = "real_half_spectrum";
        if (fused) {
            void *stw = nullptr;
            err = pl.stage_table((int)m, 16, &stw);
            if (err != hipSuccess) return fail(map_err(err));
            ps.real_stw = stw;
            ps.name = "real_rows_fused";
            P->axis_tw[0].n = (int)m;
            P->axis_tw[0].e = 16;
            P->axis_tw[0].table = stw;
        }
        P->passes.push_back(ps);
        group_end.push_back(P->passes.size());
        cur = 1;
        // Y: the columns of every z plane in one pass, Z likewise per y row.
        // FP64: the m + 1 columns kx <= m; the last tile holds only the
        // Nyquist column kx = m (its other lanes are masked and touch no
        // memory) and is rotated to run first, so it overlaps the full tiles
        // instead of trailing them.  complex64: columns kx = 0 and m are
        // real after the row step and travel packed as one complex column
        // (X[0] + i X[m] in column 0), so the passes transform an even
        // number of columns, m, which lets the 4096-point columns run the
        // paired-column tile (2D 4096^2 complex64: 71.5 -> 62.0 us); a last
        // pass unpacks them (INVERSE: packs first).  The packed form
        // measured the same for FP64 (116.7 against 117.0 us at 4096^2: the
        // column pass gains what the unpack pass costs).  Axes above 4096
        // run the four-step, with buffer 0 (the real array's, P->bytes long)
        // as the scratch of its first pass.
        P->real_pack = (ny > 1 || nz > 1) && prec == PREC_F32;
        const long long cols = P->real_pack ? m : m + 1;
        auto axis2 = [&](long long n, Geo g, int id) {
            if (n == 1) return (int)TOPE_SUCCESS;
            const size_t before = P->passes.size();
            int r = real_axis(n, g, id);
            if (r) return r;
            for (size_t i = before; i < P->passes.size(); ++i) {
                Pass &ps = P->passes[i];
                if (ps.kind == PASS_AXIS && !P->real_pack) ps.a.rot = ps.a32.rot = ps.a.tiles - 1;
                if (ps.in_buf == 0 || ps.out_buf == 0) P->real_scratch0 = true;
            }
            group_end.push_back(P->passes.size());
            return (int)TOPE_SUCCESS;
        };
        rc = axis2(ny, Geo{nz, pitch * ny, cols, 1, pitch}, 1);
        if (!rc) rc = axis2(nz, Geo{ny, pitch, cols, 1, pitch * ny}, 2);
        if (!rc && cur != 1) rc = TOPE_ERROR_UNSUPPORTED;
        if (!rc && P->real_pack) {
            Pass pk;
            pk.kind = PASS_RPACK;
            pk.in_buf = pk.out_buf = 1;
            pk.name = "real_pack_dc_nyquist";
            P->passes.push_back(pk);
            group_end.push_back(P->passes.size());
        }
    }
    if (rc) return fail(rc);
    for (size_t g = group_end.size(); g-- > 0;) {
        const size_t begin = g ? group_end[g - 1] : 0;
        for (size_t i = begin; i < group_end[g]; ++i) P->inv_order.push_back((int)i);
    }
    if (P->inv_order.size() != P->passes.size()) return fail(TOPE_ERROR_UNSUPPORTED);
    P->final_buf = 0;
    err = time_stop(pre, s, pre_ns);
    if (err != hipSuccess) return fail(map_err(err));
    EventPair up;
    err = time_start(up, s);
    if (err == hipSuccess) err = hipMemcpyAsync(P->buf[0], real, P->real_bytes, hipMemcpyHostToDevice, s);
    if (err == hipSuccess) err = time_stop(up, s, mem_ns);
    (void)hipEventDestroy(pre.start);
    (void)hipEventDestroy(pre.stop);
    (void)hipEventDestroy(up.start);
    (void)hipEventDestroy(up.stop);
    if (err != hipSuccess) return fail(map_err(err));
    *out = P;
    return TOPE_SUCCESS;
}


void release_framework(struct topeFFT *f)
{
    FrameworkImpl *fw = fw_of(f);
    if (!fw) return;
    if (fw->stream) (void)hipStreamSynchronize(fw->stream);
    if (fw->own) {
        (void)hipStreamSynchronize(fw->own);
        (void)hipStreamDestroy(fw->own);
    }
    if (fw->last.start) (void)hipEventDestroy(fw->last.start);
    if (fw->last.stop) (void)hipEventDestroy(fw->last.stop);
    delete fw;
    f->impl = nullptr;
    f->stream = nullptr;
    f->event = nullptr;
}

// Exec with the reference's semantics: transform, then blocking copy of the
// result into d (src/fft2d.c:193-198).
int exec_host(struct topeFFT *f, PlanImpl *P, void *d, int dir, tope_ulong *kern_ns, tope_ulong *mem_ns)
{
    FrameworkImpl *fw = fw_of(f);
    if (!fw || P->fw != fw) return TOPE_ERROR_NOT_INITIALIZED;
    if (!d) return TOPE_ERROR_INVALID_ARG;
    hipStream_t s = fw->stream;
    hipError_t err = time_start(fw->last, s);
    if (err != hipSuccess) return map_err(err);
    int rc = enqueue(P, s, dir);
    if (rc) return rc;
    err = time_stop(fw->last, s, kern_ns);
    if (err != hipSuccess) return map_err(err);
    EventPair dn;
    err = time_start(dn, s);
    if (err == hipSuccess) err = hipMemcpyAsync(d, P->buf[0], P->bytes, hipMemcpyDeviceToHost, s);
    if (err == hipSuccess) err = time_stop(dn, s, mem_ns);
    (void)hipEventDestroy(dn.start);
    (void)hipEventDestroy(dn.stop);
    return map_err(err);
}

int exec_device(struct topeFFT *f, PlanImpl *P, int dir)
{
    FrameworkImpl *fw = fw_of(f);
    if (!fw || !P || P->fw != fw) return TOPE_ERROR_NOT_INITIALIZED;
    return enqueue(P, fw->stream, dir);
}

void destroy_common(struct topeFFT *f, PlanImpl *P)
{
    FrameworkImpl *fw = P ? P->fw : nullptr;
    free_plan(P);
    if (fw && fw_of(f) == fw) {
        if (--fw->live_plans <= 0) release_framework(f);
    }
}

}  // namespace

// Framework plumbing shared with the uneven-batch module (many.cpp).
namespace tope {
hipStream_t fw_stream(struct topeFFT *f)
{
    FrameworkImpl *fw = fw_of(f);
    return fw ? fw->stream : nullptr;
}
void *fw_handle(struct topeFFT *f) { return fw_of(f); }
void fw_retain(struct topeFFT *f)
{
    if (FrameworkImpl *fw = fw_of(f)) fw->live_plans++;
}
void fw_release(struct topeFFT *f, void *owner)
{
    FrameworkImpl *fw = static_cast<FrameworkImpl *>(owner);
    if (fw && fw_of(f) == fw && --fw->live_plans <= 0) release_framework(f);
}
const AxisVariant *row_variant(int prec, int n) { return select_variant(prec, n, false, 1); }
const AxisVariant *axis_variant_for(int prec, int n, bool cfast, long long lines, long long outer)
{
    return select_variant(prec, n, cfast, lines, false, outer, true);
}
}  // namespace tope

extern "C" {

void topeFFTInit(struct topeFFT *f)
{
    if (!f) return;
    std::memset(f, 0, sizeof(*f));
    int n = 0;
    hipError_t err = hipGetDeviceCount(&n);
    if (err != hipSuccess || n == 0) {
        f->error = TOPE_ERROR_NO_DEVICE;
        return;
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    FrameworkImpl *fw = new (std::nothrow) FrameworkImpl();
    if (!fw) {
        f->error = TOPE_ERROR_UNSUPPORTED;
        return;
    }
    err = hipStreamCreateWithFlags(&fw->own, hipStreamNonBlocking);
    if (err != hipSuccess) {
        delete fw;
        f->error = map_err(err);
        return;
    }
    // device address of the allocation-free entry point's static scratch
    // array on this GPU (once per device)
    err = noalloc_setup(dev);
    if (err != hipSuccess) {
        (void)hipStreamDestroy(fw->own);
        delete fw;
        f->error = map_err(err);
        return;
    }
    fw->stream = fw->own;
    f->impl = fw;
    f->stream = fw->stream;
    f->event = &fw->last;
    f->device = dev;
    f->ret_num_platforms = 1;
    f->ret_num_devices = (unsigned)n;
    if (std::getenv("TOPEFFT_LOG")) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess)
            std::fprintf(stderr, "topeFFT device %d: %s (%s), %.0f MiB, %d CUs, %zu KiB LDS/block\n", dev, prop.name,
                         prop.gcnArchName, prop.totalGlobalMem / 1048576.0, prop.multiProcessorCount,
                         prop.sharedMemPerBlock / 1024);
    }
    f->error = TOPE_SUCCESS;
}

void topeFFTDestroy(struct topeFFT *f)
{
    if (!f) return;
    release_framework(f);
    f->error = TOPE_SUCCESS;
}

tope_ulong profileThis(void *event)
{
    EventPair *ev = static_cast<EventPair *>(event);
    if (!ev || !ev->recorded) return 0;
    if (hipEventSynchronize(ev->stop) != hipSuccess) return 0;
    float ms = 0;
    if (hipEventElapsedTime(&ms, ev->start, ev->stop) != hipSuccess) return 0;
    return (tope_ulong)((double)ms * 1e6);
}

/* ---- 1D ---------------------------------------------------------------- */

void tope1DPlanInit(struct topeFFT *f, struct topePlan1D *t, int x, int type, double *d)
{
    if (!f || !t) return;
    std::memset(t, 0, sizeof(*t));
    t->x = x;
    t->type = type;
    t->dim = 1;
    ref_radix(x, &t->radix, &t->log);
    PlanImpl *P = nullptr;
    int rc = build_plan(f, 1, x, 1, 1, PREC_F64, MP_VALU, d, &P, &t->totalPreKernel, &t->totalMemory);
    f->error = rc;
    if (rc) return;
    P->fw->live_plans++;
    t->impl = P;
    t->dataSize = P->bytes;
    t->twiddle = P->axis_tw[0].table;
    sync_fields(t, P);
}

void tope1DExec(struct topeFFT *f, struct topePlan1D *t, double *d, int dir)
{
    if (!f || !t || !t->impl) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    PlanImpl *P = impl_of(t->impl);
    f->error = exec_host(f, P, d, dir, &t->totalKernel, &t->totalMemory);
    sync_fields(t, P);
}

void tope1DExecDevice(struct topeFFT *f, struct topePlan1D *t, int dir)
{
    if (!f || !t || !t->impl) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    PlanImpl *P = impl_of(t->impl);
    f->error = exec_device(f, P, dir);
    sync_fields(t, P);
}

void tope1DDestroy(struct topeFFT *f, struct topePlan1D *t)
{
    if (!f || !t || !t->impl) return;
    destroy_common(f, impl_of(t->impl));
    t->impl = nullptr;
    t->data = t->scratch = t->twiddle = nullptr;
    f->error = TOPE_SUCCESS;
}

/* ---- 2D ---------------------------------------------------------------- */

void tope2DPlanInit(struct topeFFT *f, struct topePlan2D *t, int x, int y, int type, double *d)
{
    if (!f || !t) return;
    std::memset(t, 0, sizeof(*t));
    t->x = x;
    t->y = y;
    t->type = type;
    t->dim = 2;
    ref_radix(x, &t->radX, &t->logX);
    ref_radix(y, &t->radY, &t->logY);
    PlanImpl *P = nullptr;
    int rc = build_plan(f, 2, x, y, 1, PREC_F64, MP_VALU, d, &P, &t->totalPreKernel, &t->totalMemory);
    f->error = rc;
    if (rc) return;
    P->fw->live_plans++;
    t->impl = P;
    t->dataSize = P->bytes;
    t->twdX = P->axis_tw[0].table;
    t->twdY = P->axis_tw[1].table;
    sync_fields(t, P);
}

void tope2DExec(struct topeFFT *f, struct topePlan2D *t, double *d, int dir)
{
    if (!f || !t || !t->impl) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    PlanImpl *P = impl_of(t->impl);
    f->error = exec_host(f, P, d, dir, &t->totalKernel, &t->totalMemory);
    sync_fields(t, P);
}

void tope2DExecDevice(struct topeFFT *f, struct topePlan2D *t, int dir)
{
    if (!f || !t || !t->impl) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    PlanImpl *P = impl_of(t->impl);
    f->error = exec_device(f, P, dir);
    sync_fields(t, P);
}

void tope2DDestroy(struct topeFFT *f, struct topePlan2D *t)
{
    if (!f || !t || !t->impl) return;
    destroy_common(f, impl_of(t->impl));
    t->impl = nullptr;
    t->data = t->scratch = t->twdX = t->twdY = nullptr;
    f->error = TOPE_SUCCESS;
}

/* ---- 3D ---------------------------------------------------------------- */

void tope3DPlanInit(struct topeFFT *f, struct topePlan3D *t, int x, int y, int z, int type, double *d)
{
    if (!f || !t) return;
    std::memset(t, 0, sizeof(*t));
    t->x = x;
    t->y = y;
    t->z = z;
    t->type = type;
    t->dim = 3;
    ref_radix(x, &t->radX, &t->logX);
    ref_radix(y, &t->radY, &t->logY);
    ref_radix(z, &t->radZ, &t->logZ);
    PlanImpl *P = nullptr;
    int rc = build_plan(f, 3, x, y, z, PREC_F64, MP_VALU, d, &P, &t->totalPreKernel, &t->totalMemory);
    f->error = rc;
    if (rc) return;
    P->fw->live_plans++;
    t->impl = P;
    t->dataSize = P->bytes;
    t->twdX = P->axis_tw[0].table;
    t->twdY = P->axis_tw[1].table;
    t->twdZ = P->axis_tw[2].table;
    sync_fields(t, P);
}

void tope3DExec(struct topeFFT *f, struct topePlan3D *t, double *d, int dir)
{
    if (!f || !t || !t->impl) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    PlanImpl *P = impl_of(t->impl);
    f->error = exec_host(f, P, d, dir, &t->totalKernel, &t->totalMemory);
    sync_fields(t, P);
}

void tope3DExecDevice(struct topeFFT *f, struct topePlan3D *t, int dir)
{
    if (!f || !t || !t->impl) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    PlanImpl *P = impl_of(t->impl);
    f->error = exec_device(f, P, dir);
    sync_fields(t, P);
}

void tope3DDestroy(struct topeFFT *f, struct topePlan3D *t)
{
    if (!f || !t || !t->impl) return;
    destroy_common(f, impl_of(t->impl));
    t->impl = nullptr;
    t->data = t->scratch = t->twdX = t->twdY = t->twdZ = nullptr;
    f->error = TOPE_SUCCESS;
}

/* ---- 2D/3D plans in a chosen precision ------------------------------------ */

void tope2DPlanInitMode(struct topeFFT *f, struct topePlan2D *t, int x, int y, int type, int precision, int mode,
                        const void *host)
{
    if (!f || !t) return;
    std::memset(t, 0, sizeof(*t));
    t->x = x;
    t->y = y;
    t->type = type;
    t->dim = 2;
    ref_radix(x, &t->radX, &t->logX);
    ref_radix(y, &t->radY, &t->logY);
    PlanImpl *P = nullptr;
    int rc = build_plan(f, 2, x, y, 1, precision == TOPE_PREC_F32 ? PREC_F32 : PREC_F64, mode, host, &P,
                        &t->totalPreKernel, &t->totalMemory);
    if (!rc && precision != TOPE_PREC_F64 && precision != TOPE_PREC_F32) rc = TOPE_ERROR_INVALID_ARG;
    f->error = rc;
    if (rc) {
        if (P) free_plan(P);
        return;
    }
    P->fw->live_plans++;
    t->impl = P;
    t->dataSize = P->bytes;
    t->twdX = P->axis_tw[0].table;
    t->twdY = P->axis_tw[1].table;
    sync_fields(t, P);
}

void tope2DPlanInitEx(struct topeFFT *f, struct topePlan2D *t, int x, int y, int type, int precision,
                      const void *host)
{
    tope2DPlanInitMode(f, t, x, y, type, precision, TOPE_MODE_VALU, host);
}

void tope3DPlanInitMode(struct topeFFT *f, struct topePlan3D *t, int x, int y, int z, int type, int precision,
                        int mode, const void *host)
{
    if (!f || !t) return;
    std::memset(t, 0, sizeof(*t));
    t->x = x;
    t->y = y;
    t->z = z;
    t->type = type;
    t->dim = 3;
    ref_radix(x, &t->radX, &t->logX);
    ref_radix(y, &t->radY, &t->logY);
    ref_radix(z, &t->radZ, &t->logZ);
    PlanImpl *P = nullptr;
    int rc = build_plan(f, 3, x, y, z, precision == TOPE_PREC_F32 ? PREC_F32 : PREC_F64, mode, host, &P,
                        &t->totalPreKernel, &t->totalMemory);
    if (!rc && precision != TOPE_PREC_F64 && precision != TOPE_PREC_F32) rc = TOPE_ERROR_INVALID_ARG;
    f->error = rc;
    if (rc) {
        if (P) free_plan(P);
        return;
    }
    P->fw->live_plans++;
    t->impl = P;
    t->dataSize = P->bytes;
    t->twdX = P->axis_tw[0].table;
    t->twdY = P->axis_tw[1].table;
    t->twdZ = P->axis_tw[2].table;
    sync_fields(t, P);
}

void tope3DPlanInitEx(struct topeFFT *f, struct topePlan3D *t, int x, int y, int z, int type, int precision,
                      const void *host)
{
    tope3DPlanInitMode(f, t, x, y, z, type, precision, TOPE_MODE_VALU, host);
}

void tope2DExecEx(struct topeFFT *f, struct topePlan2D *t, void *host, int dir)
{
    if (!f || !t || !t->impl) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    PlanImpl *P = impl_of(t->impl);
    f->error = exec_host(f, P, host, dir, &t->totalKernel, &t->totalMemory);
    sync_fields(t, P);
}

void tope3DExecEx(struct topeFFT *f, struct topePlan3D *t, void *host, int dir)
{
    if (!f || !t || !t->impl) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    PlanImpl *P = impl_of(t->impl);
    f->error = exec_host(f, P, host, dir, &t->totalKernel, &t->totalMemory);
    sync_fields(t, P);
}

/* ---- batched 1D transforms ---------------------------------------------- */

void topeBatchPlanInit(struct topeFFT *f, struct topePlanBatch *t, int n, int count, int precision, int mode,
                       const void *host)
{
    if (!f || !t) return;
    std::memset(t, 0, sizeof(*t));
    t->n = n;
    t->count = count;
    t->precision = precision;
    t->mode = mode;
    PlanImpl *P = nullptr;
    int rc = build_plan(f, SHAPE_BATCH, n, count, 1, precision, mode, host, &P, &t->totalPreKernel,
                        &t->totalMemory);
    f->error = rc;
    if (rc) return;
    P->fw->live_plans++;
    t->impl = P;
    t->dataSize = P->bytes;
    t->data = P->buf[0];
}

void topeBatchExec(struct topeFFT *f, struct topePlanBatch *t, void *host, int dir)
{
    if (!f || !t || !t->impl) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    PlanImpl *P = impl_of(t->impl);
    f->error = exec_host(f, P, host, dir, &t->totalKernel, &t->totalMemory);
    t->data = P->buf[0];
}

void topeBatchExecDevice(struct topeFFT *f, struct topePlanBatch *t, int dir)
{
    if (!f || !t || !t->impl) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    PlanImpl *P = impl_of(t->impl);
    f->error = exec_device(f, P, dir);
    t->data = P->buf[0];
}

void topeBatchDestroy(struct topeFFT *f, struct topePlanBatch *t)
{
    if (!f || !t || !t->impl) return;
    destroy_common(f, impl_of(t->impl));
    t->impl = nullptr;
    t->data = nullptr;
    f->error = TOPE_SUCCESS;
}

/* ---- extensions -------------------------------------------------------- */

void topePlanUpload(struct topeFFT *f, void *plan_impl, const void *host, size_t bytes)
{
    PlanImpl *P = impl_of(plan_impl);
    FrameworkImpl *fw = fw_of(f);
    if (!P || !fw || !host || bytes > P->bytes) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    hipError_t err = hipMemcpyAsync(P->buf[0], host, bytes, hipMemcpyHostToDevice, fw->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(fw->stream);
    f->error = map_err(err);
}

void topePlanDownload(struct topeFFT *f, void *plan_impl, void *host, size_t bytes)
{
    PlanImpl *P = impl_of(plan_impl);
    FrameworkImpl *fw = fw_of(f);
    if (!P || !fw || !host || bytes > P->bytes) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    hipError_t err = hipMemcpyAsync(host, P->buf[0], bytes, hipMemcpyDeviceToHost, fw->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(fw->stream);
    f->error = map_err(err);
}

void topeFFTSynchronize(struct topeFFT *f)
{
    FrameworkImpl *fw = fw_of(f);
    if (!fw) {
        if (f) f->error = TOPE_ERROR_NOT_INITIALIZED;
        return;
    }
    f->error = map_err(hipStreamSynchronize(fw->stream));
}

int topePlanPassCount(void *plan_impl)
{
    PlanImpl *P = impl_of(plan_impl);
    return P ? (int)P->passes.size() : 0;
}

void topePlanPassTiming(void *plan_impl, int enable)
{
    PlanImpl *P = impl_of(plan_impl);
    if (!P) return;
    P->timing = enable != 0;
    if (!P->timing) {
        P->pending.clear();
        P->pool_used = 0;
    }
}

int topePlanPassTimes(struct topeFFT *f, void *plan_impl, double *ms, int max_passes)
{
    PlanImpl *P = impl_of(plan_impl);
    FrameworkImpl *fw = fw_of(f);
    if (!P || !fw || !ms) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return 0;
    }
    hipError_t err = hipStreamSynchronize(fw->stream);
    if (err != hipSuccess) {
        f->error = map_err(err);
        return 0;
    }
    const int np = (int)P->passes.size();
    std::vector<double> sum(np, 0.0);
    std::vector<long long> cnt(np, 0);
    for (auto &pe : P->pending) {
        float t = 0;
        if (hipEventElapsedTime(&t, pe.second.first, pe.second.second) == hipSuccess) {
            sum[pe.first] += t;
            cnt[pe.first]++;
        }
    }
    P->pending.clear();
    P->pool_used = 0;
    const int n = np < max_passes ? np : max_passes;
    for (int i = 0; i < n; ++i) ms[i] = cnt[i] ? sum[i] / (double)cnt[i] : 0.0;
    f->error = TOPE_SUCCESS;
    return n;
}

double topePlanTimePass(struct topeFFT *f, void *plan_impl, int pass, int dir, int reps)
{
    PlanImpl *P = impl_of(plan_impl);
    FrameworkImpl *fw = fw_of(f);
    if (!P || !fw || pass < 0 || pass >= (int)P->passes.size() || reps < 1 ||
        (dir != FORWARD && dir != INVERSE)) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return -1.0;
    }
    EventPair ev;
    hipError_t err = time_start(ev, fw->stream, false);
    int rc = TOPE_SUCCESS;
    for (int r = 0; r < reps && err == hipSuccess && !rc; ++r) rc = launch_pass(P, pass, fw->stream, dir);
    tope_ulong ns = 0;
    if (err == hipSuccess && !rc) err = time_stop(ev, fw->stream, &ns);
    if (ev.start) (void)hipEventDestroy(ev.start);
    if (ev.stop) (void)hipEventDestroy(ev.stop);
    f->error = rc ? rc : map_err(err);
    return f->error ? -1.0 : (double)ns * 1e-6 / reps;
}

const char *topePlanPassName(void *plan_impl, int pass)
{
    PlanImpl *P = impl_of(plan_impl);
    if (!P || pass < 0 || pass >= (int)P->passes.size()) return "";
    return P->passes[pass].name;
}

/* ---- real-input transforms (no reference counterpart) ------------------- */

void topeRealPlanInitEx(struct topeFFT *f, struct topePlanReal *t, int x, int y, int z, int precision,
                        const void *real)
{
    if (!f || !t) return;
    std::memset(t, 0, sizeof(*t));
    t->x = x;
    t->y = y;
    t->z = z;
    PlanImpl *P = nullptr;
    const int rc = build_real_plan(f, x, y, z, precision, real, &P, &t->totalPreKernel, &t->totalMemory);
    f->error = rc;
    if (rc) return;
    P->fw->live_plans++;
    t->impl = P;
    t->dim = P->ndim;
    t->real = P->buf[0];
    t->data = P->buf[1];
    t->realSize = P->real_bytes;
    t->dataSize = P->bytes;
    t->pitch = (int)P->real_pitch;
    t->precision = precision;
}

void topeRealPlanInit(struct topeFFT *f, struct topePlanReal *t, int x, int y, int z, const double *real)
{
    topeRealPlanInitEx(f, t, x, y, z, PREC_F64, real);
}

void topeRealExecEx(struct topeFFT *f, struct topePlanReal *t, void *host, int dir)
{
    if (!f || !t) return;
    PlanImpl *P = impl_of(t->impl);
    if (!P || !P->real) {
        f->error = TOPE_ERROR_NOT_INITIALIZED;
        return;
    }
    FrameworkImpl *fw = fw_of(f);
    if (!fw || P->fw != fw) {
        f->error = TOPE_ERROR_NOT_INITIALIZED;
        return;
    }
    if (!host) {
        f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    hipStream_t s = fw->stream;
    hipError_t err = time_start(fw->last, s);
    int rc = err == hipSuccess ? enqueue(P, s, dir) : map_err(err);
    if (!rc) rc = map_err(time_stop(fw->last, s, &t->totalKernel));
    if (!rc) {
        EventPair dn;
        err = time_start(dn, s);
        if (err == hipSuccess) {
            if (dir == FORWARD) {
                // rows of m + 1 elements out of the padded device pitch
                const size_t row = (size_t)(P->real_m + 1) * P->elem, dpitch = (size_t)P->real_pitch * P->elem;
                err = hipMemcpy2DAsync(host, row, P->buf[1], dpitch, row, (size_t)P->real_rows, hipMemcpyDeviceToHost,
                                       s);
            } else {
                err = hipMemcpyAsync(host, P->buf[0], P->real_bytes, hipMemcpyDeviceToHost, s);
            }
        }
        if (err == hipSuccess) err = time_stop(dn, s, &t->totalMemory);
        (void)hipEventDestroy(dn.start);
        (void)hipEventDestroy(dn.stop);
        rc = map_err(err);
    }
    f->error = rc;
}

void topeRealExec(struct topeFFT *f, struct topePlanReal *t, double *host, int dir)
{
    topeRealExecEx(f, t, host, dir);
}

void topeRealExecDevice(struct topeFFT *f, struct topePlanReal *t, int dir)
{
    if (!f || !t) return;
    PlanImpl *P = impl_of(t->impl);
    f->error = (P && P->real) ? exec_device(f, P, dir) : TOPE_ERROR_NOT_INITIALIZED;
}

void topeRealExecPtr(struct topeFFT *f, struct topePlanReal *t, void *real, void *spectrum, int dir)
{
    if (!f || !t) return;
    PlanImpl *P = impl_of(t->impl);
    FrameworkImpl *fw = fw_of(f);
    if (!P || !P->real || !fw || P->fw != fw) {
        f->error = TOPE_ERROR_NOT_INITIALIZED;
        return;
    }
    // the real array is read as complex elements: both arrays aligned to the
    // complex element size, the spectrum to 16 bytes where complex64 column
    // pairs move as one element
    const uintptr_t salign = P->pair_tiles ? 16 : (uintptr_t)P->elem;
    if (!real || !spectrum || (dir != FORWARD && dir != INVERSE) ||
        reinterpret_cast<uintptr_t>(real) % P->elem || reinterpret_cast<uintptr_t>(spectrum) % salign) {
        f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    if (P->real_scratch0) {
        // y or z above 4096: the four-step scratch needs dataSize bytes of
        // buffer 0, more than a caller's real array holds
        f->error = TOPE_ERROR_UNSUPPORTED;
        return;
    }
    // the plan's passes run on the caller's arrays in place of its own
    // (buffer 0: real array, buffer 1: half spectrum at the plan's pitch)
    void *own0 = P->buf[0], *own1 = P->buf[1];
    P->buf[0] = real;
    P->buf[1] = spectrum;
    const int rc = enqueue(P, fw->stream, dir);
    P->buf[0] = own0;
    P->buf[1] = own1;
    f->error = rc;
}

void topeRealDestroy(struct topeFFT *f, struct topePlanReal *t)
{
    if (!f || !t) return;
    destroy_common(f, impl_of(t->impl));
    t->impl = nullptr;
    t->data = t->real = nullptr;
    f->error = TOPE_SUCCESS;
}

/* ---- caller streams and caller-owned device arrays ----------------------- */

void topeFFTSetStream(struct topeFFT *f, void *stream)
{
    if (!f) return;
    FrameworkImpl *fw = fw_of(f);
    if (!fw) {
        f->error = TOPE_ERROR_NOT_INITIALIZED;
        return;
    }
    hipStream_t next = stream ? static_cast<hipStream_t>(stream) : fw->own;
    int rc = TOPE_SUCCESS;
    if (next != fw->stream) {
        // work already queued on the old stream completes before anything
        // enqueued on the new one
        hipEvent_t ev = nullptr;
        hipError_t err = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (err == hipSuccess) err = hipEventRecord(ev, fw->stream);
        if (err == hipSuccess) err = hipStreamWaitEvent(next, ev, 0);
        if (ev) (void)hipEventDestroy(ev);
        rc = map_err(err);
        if (!rc) fw->stream = next;
    }
    f->stream = fw->stream;
    f->error = rc;
}

void topePlanExecPtr(struct topeFFT *f, void *plan_impl, void *data, int dir)
{
    if (!f) return;
    PlanImpl *P = impl_of(plan_impl);
    FrameworkImpl *fw = fw_of(f);
    if (!P || !fw || P->fw != fw) {
        f->error = TOPE_ERROR_NOT_INITIALIZED;
        return;
    }
    if (!data || P->real || P->buf_bytes > P->bytes) {
        // real plans and padded (Bluestein) plans work on arrays of their own
        f->error = !data ? TOPE_ERROR_INVALID_ARG : TOPE_ERROR_UNSUPPORTED;
        return;
    }
    // elements are loaded whole: the array aligned to the element size, or
    // to 16 bytes where complex64 column pairs move as one element
    const uintptr_t align = P->pair_tiles ? 16 : (uintptr_t)P->elem;
    if (reinterpret_cast<uintptr_t>(data) % align) {
        f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    // run the plan's passes with the caller's array in place of its own;
    // schedules that end in the scratch buffer copy the result back
    void *own0 = P->buf[0], *own1 = P->buf[1];
    P->buf[0] = data;
    int rc = enqueue(P, fw->stream, dir);
    if (!rc && P->final_buf == 1)
        rc = map_err(hipMemcpyAsync(data, P->buf[0], P->bytes, hipMemcpyDeviceToDevice, fw->stream));
    P->buf[0] = own0;
    P->buf[1] = own1;
    f->error = rc;
}

size_t topeStructSize(int which)
{
    switch (which) {
    case 0: return sizeof(struct topeFFT);
    case 1: return sizeof(struct topePlan1D);
    case 2: return sizeof(struct topePlan2D);
    case 3: return sizeof(struct topePlan3D);
    case 4: return sizeof(struct topePlanBatch);
    case 5: return sizeof(struct topePlanMany);
    case 6: return sizeof(struct topePlanReal);
    default: return 0;
    }
}

const char *topeErrorString(int code)
{
    switch (code) {
    case TOPE_SUCCESS: return "success";
    case TOPE_ERROR_INVALID_SIZE: return "unsupported transform size";
    case TOPE_ERROR_INVALID_ARG: return "invalid argument";
    case TOPE_ERROR_NO_DEVICE: return "no HIP device";
    case TOPE_ERROR_NOT_INITIALIZED: return "framework not initialised (topeFFTInit)";
    case TOPE_ERROR_NO_KERNEL: return "no compiled kernel for this shape";
    case TOPE_ERROR_UNSUPPORTED: return "unsupported request";
    case TOPE_ERROR_SCRATCH_LIMIT: return "transform exceeds the allocation-free entry point's scratch array";
    default: return code > 0 ? hipGetErrorString((hipError_t)code) : "unknown error";
    }
}

void topeExec1DNoAlloc(struct topeFFT *f, void *data, int n, int count, int dir)
{
    FrameworkImpl *fw = fw_of(f);
    if (!fw) {
        if (f) f->error = TOPE_ERROR_NOT_INITIALIZED;
        return;
    }
    if (!data || (dir != FORWARD && dir != INVERSE) || reinterpret_cast<uintptr_t>(data) % 16) {
        f->error = TOPE_ERROR_INVALID_ARG;  // complex128 elements are loaded whole
        return;
    }
    f->error = noalloc_exec_1d(f->device, data, n, count, dir == INVERSE, fw->stream);
}

const char *topeAxisVariantName(int index)
{
    int cnt = 0;
    const AxisVariant *v = axis_variants(&cnt);
    return (index >= 0 && index < cnt) ? v[index].name : nullptr;
}

long long topePlanTwiddles(struct topeFFT *f, void *plan_impl, int axis, double *host, long long max_entries)
{
    PlanImpl *P = impl_of(plan_impl);
    FrameworkImpl *fw = fw_of(f);
    if (!P || !fw || axis < 0 || axis > 2 || !P->axis_tw[axis].table) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return -1;
    }
    const auto &at = P->axis_tw[axis];
    long long count = at.e ? host_sched(at.n, at.e).twsize : at.n;
    if (!host) return count;
    long long n = count < max_entries ? count : max_entries;
    hipError_t err = hipMemcpyAsync(host, at.table, (size_t)n * P->elem, hipMemcpyDeviceToHost, fw->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(fw->stream);
    f->error = map_err(err);
    return err == hipSuccess ? n : -1;
}

}  // extern "C"
