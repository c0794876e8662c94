// C-ABI of the packed uneven batch: many transforms of different power-of-two
// lengths stored back to back (offset table), run by persistent kernels
// (one for the long lengths, one beside it on a side stream for the rest),
// optionally replayed from a captured hipGraph, with a per-length
// radix schedule chosen by an autotuner whose results persist in a text
// "wisdom" file.  No reference counterpart: the reference can express this
// only as a loop of 1D plans, one topeFFTInit each (src/fft1d.c:303-407).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/topefft.h"
#include "internal.hpp"
#include "many_kernel.hpp"

using namespace tope;

namespace {

struct ManyImpl {
    void *fw = nullptr;  // owning framework
    hipStream_t stream = nullptr;
    int prec = PREC_F64;
    size_t elem = 16;
    int count = 0;
    long long elems = 0;
    std::vector<int> sizes;
    std::vector<long long> offs;
    void *data = nullptr;
    std::vector<void *> allocs;
    // persistent-kernel tables: tasks of class L are seg0[L] .. seg1[L];
    // the class's transforms sit at positions pos0[L] .. pos0[L] + cnt[L]
    // of the class-grouped order, d_toffs[pos] the element offset of the
    // transform at position pos, posof[i] the position of transform i
    long long *d_toffs = nullptr;
    void *d_tw[16][2] = {};  // stage tables per class: [L][0] E = 16, [L][1] E = 8
    int seg0[16] = {}, seg1[16] = {};
    int pos0[16] = {}, cnt[16] = {};
    std::vector<int> posof;
    int sched[16] = {};
    int grid = 0;        // blocks of the 1024-thread kernel (classes of P > 256)
    int grid_small = 0;  // blocks of the 256-thread kernel (P <= 256)
    int method = TOPE_MANY_PERSISTENT;
    // graphs per (method, dir)
    hipGraphExec_t gexec[4][2] = {};
    hipEvent_t e0 = nullptr, e1 = nullptr;  // timing only (autotuner): no system fence
    hipEvent_t x0 = nullptr, x1 = nullptr;  // around exec-with-download: fenced
    // the 256-thread kernel runs beside the 1024-thread one on a second
    // stream (fork/join through two events), filling the CUs the long
    // classes leave idle at their tail
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

int map_err(hipError_t e) { return e == hipSuccess ? TOPE_SUCCESS : (int)e; }

int ilog2i(long long n)
{
    int l = 0;
    while ((1LL << l) < n) ++l;
    return l;
}

bool e8_ok(int L) { return (1 << L) >= 8 && (1 << L) / 8 <= MANY_THREADS; }

int many_P(int L, bool e8)
{
    const int n = 1 << L;
    if (e8 && e8_ok(L)) return n / 8;
    return n < 16 ? 1 : n / 16;
}

// Block size of the kernel that runs class L: 256 threads if one transform
// of its schedule fits them, else 1024.
int class_threads(const ManyImpl *M, int L)
{
    return many_P(L, M->sched[L] != 0) <= MANY_SMALL_THREADS ? MANY_SMALL_THREADS : MANY_THREADS;
}

void drop_graphs(ManyImpl *M)
{
    for (auto &g : M->gexec)
        for (auto &e : g)
            if (e) {
                (void)hipGraphExecDestroy(e);
                e = nullptr;
            }
}

// Task list: classes from the largest down, each class's transforms in
// their packed order, TB(class) transforms per task.
hipError_t build_tasks(ManyImpl *M)
{
    std::vector<long long> toffs;
    M->posof.assign(M->count, 0);
    int ntasks = 0;
    for (int L = 0; L < 16; ++L) M->seg0[L] = M->seg1[L] = M->pos0[L] = M->cnt[L] = 0;
    for (int L = MANY_LMAX; L >= MANY_LMIN; --L) {
        const int tb = class_threads(M, L) / many_P(L, M->sched[L] != 0);
        M->pos0[L] = (int)toffs.size();
        for (int i = 0; i < M->count; ++i)
            if (M->sizes[i] == (1 << L)) {
                M->posof[i] = (int)toffs.size();
                toffs.push_back(M->offs[i]);
            }
        M->cnt[L] = (int)toffs.size() - M->pos0[L];
        M->seg0[L] = ntasks;
        ntasks += (M->cnt[L] + tb - 1) / tb;
        M->seg1[L] = ntasks;
    }
    hipError_t err = hipSuccess;
    if (!M->d_toffs) err = hipMalloc(&M->d_toffs, sizeof(long long) * std::max<size_t>(1, toffs.size()));
    if (err == hipSuccess && !toffs.empty())
        err = hipMemcpyAsync(M->d_toffs, toffs.data(), sizeof(long long) * toffs.size(), hipMemcpyHostToDevice,
                             M->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(M->stream);
    drop_graphs(M);
    return err;
}

// Arguments of one of the kernels (by block size): the segment bounds of
// the classes it runs, empty segments for the others.
template <typename R>
ManyArgs<R> many_args(const ManyImpl *M, int threads, int only_class = -1)
{
    ManyArgs<R> a{};
    for (int L = 0; L < 16; ++L) {
        const bool on = (only_class < 0 || only_class == L) && class_threads(M, L) == threads;
        a.seg0[L] = on ? M->seg0[L] : 0;
        a.seg1[L] = on ? M->seg1[L] : 0;
        a.pos0[L] = M->pos0[L];
        a.cnt[L] = M->cnt[L];
        a.sched[L] = M->sched[L];
        a.tw[L] = static_cast<const cpx<R> *>(M->d_tw[L][M->sched[L] ? 1 : 0]);
    }
    a.toffs = M->d_toffs;
    a.elems = M->elems;
    return a;
}

// The classes of each block size on their kernel (launched only if it has
// tasks).  The two kernels touch disjoint transforms: the 1024-thread one
// runs on the plan's stream, the 256-thread one on the side stream beside
// it (fork/join), filling the CUs the long classes leave idle.
// `grid` overrides the 1024-thread kernel's block count (autotuner).
int launch_persistent(ManyImpl *M, int dir, hipStream_t s, int only_class = -1, int grid = 0)
{
    const int sign = dir == FORWARD ? -1 : 1;
    const int sizes[2] = {MANY_THREADS, MANY_SMALL_THREADS};
    bool any[2] = {false, false};
    for (int L = MANY_LMIN; L <= MANY_LMAX; ++L)
        if ((only_class < 0 || only_class == L) && M->seg1[L] > M->seg0[L])
            for (int k = 0; k < 2; ++k)
                if (class_threads(M, L) == sizes[k]) any[k] = true;
    const int first = any[0] ? 0 : 1;
    const bool fork = any[0] && any[1] && M->side;
    hipError_t err = hipSuccess;
    if (fork) {
        err = hipEventRecord(M->fork, s);
        if (err == hipSuccess) err = hipStreamWaitEvent(M->side, M->fork, 0);
        if (err != hipSuccess) return map_err(err);
    }
    for (int k = 0; k < 2; ++k) {
        if (!any[k]) continue;
        const int g = k == 0 ? (grid > 0 ? grid : M->grid) : M->grid_small;
        hipStream_t st = (fork && k != first) ? M->side : s;
        if (M->prec == PREC_F64) {
            ManyArgs<double> a = many_args<double>(M, sizes[k], only_class);
            err = launch_many(PREC_F64, sizes[k], M->data, M->data, &a, g, sign, st);
        } else {
            ManyArgs<float> a = many_args<float>(M, sizes[k], only_class);
            err = launch_many(PREC_F32, sizes[k], M->data, M->data, &a, g, sign, st);
        }
        if (err != hipSuccess) return map_err(err);
    }
    if (fork) {
        err = hipEventRecord(M->join, M->side);
        if (err == hipSuccess) err = hipStreamWaitEvent(s, M->join, 0);
    }
    return map_err(err);
}

// Baseline: one launch per transform -- the same device code on a
// one-transform class range, one block, in place at the transform's offset.
int launch_per_transform(ManyImpl *M, int dir, hipStream_t s)
{
    const int sign = dir == FORWARD ? -1 : 1;
    for (int i = 0; i < M->count; ++i) {
        const int L = ilog2i(M->sizes[i]);
        const int nt = class_threads(M, L);
        hipError_t err;
        if (M->prec == PREC_F64) {
            ManyArgs<double> a = many_args<double>(M, nt, L);
            a.seg0[L] = 0;
            a.seg1[L] = 1;
            a.pos0[L] = M->posof[i];
            a.cnt[L] = 1;
            err = launch_many(PREC_F64, nt, M->data, M->data, &a, 1, sign, s);
        } else {
            ManyArgs<float> a = many_args<float>(M, nt, L);
            a.seg0[L] = 0;
            a.seg1[L] = 1;
            a.pos0[L] = M->posof[i];
            a.cnt[L] = 1;
            err = launch_many(PREC_F32, nt, M->data, M->data, &a, 1, sign, s);
        }
        if (err != hipSuccess) return map_err(err);
    }
    return TOPE_SUCCESS;
}

int launch_method(ManyImpl *M, int method, int dir, hipStream_t s)
{
    if (method == TOPE_MANY_PERSISTENT) return launch_persistent(M, dir, s);
    if (method == TOPE_MANY_PER_TRANSFORM) return launch_per_transform(M, dir, s);
    // graph methods: capture the plain method once per direction, then replay
    const int base = method == TOPE_MANY_PERSISTENT_GRAPH ? TOPE_MANY_PERSISTENT : TOPE_MANY_PER_TRANSFORM;
    const int d = dir == FORWARD ? 0 : 1;
    hipGraphExec_t &ge = M->gexec[method][d];
    if (!ge) {
        hipGraph_t g = nullptr;
        hipError_t err = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (err != hipSuccess) return map_err(err);
        int rc = launch_method(M, base, dir, s);
        err = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (err == hipSuccess) err = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (err != hipSuccess) {
            ge = nullptr;
            return map_err(err);
        }
    }
    return map_err(hipGraphLaunch(ge, s));
}

ManyImpl *impl_of(struct topePlanMany *t) { return t ? static_cast<ManyImpl *>(t->impl) : nullptr; }

void free_many(ManyImpl *M)
{
    if (!M) return;
    if (M->stream) (void)hipStreamSynchronize(M->stream);
    if (M->side) (void)hipStreamSynchronize(M->side);
    drop_graphs(M);
    for (void *p : M->allocs) (void)hipFree(p);
    if (M->data) (void)hipFree(M->data);
    if (M->d_toffs) (void)hipFree(M->d_toffs);
    if (M->e0) (void)hipEventDestroy(M->e0);
    if (M->e1) (void)hipEventDestroy(M->e1);
    if (M->x0) (void)hipEventDestroy(M->x0);
    if (M->x1) (void)hipEventDestroy(M->x1);
    if (M->fork) (void)hipEventDestroy(M->fork);
    if (M->join) (void)hipEventDestroy(M->join);
    if (M->side) (void)hipStreamDestroy(M->side);
    delete M;
}

std::string device_name()
{
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
    return std::string(prop.gcnArchName);
}

// Wisdom file: one line per (device, precision):
//   <device> <f64|f32> grid_per_cu=<g> sched=<L>:<16|8>,<L>:<16|8>,...
void load_wisdom(ManyImpl *M, int cus)
{
    const char *path = std::getenv("TOPEFFT_WISDOM");
    if (!path) return;
    FILE *fp = std::fopen(path, "r");
    if (!fp) return;
    const std::string dev = device_name(), pr = M->prec == PREC_F64 ? "f64" : "f32";
    char line[1024];
    while (std::fgets(line, sizeof line, fp)) {
        char d[256], p[16], rest[700];
        int g = 0;
        if (std::sscanf(line, "%255s %15s grid_per_cu=%d sched=%699s", d, p, &g, rest) != 4) continue;
        if (dev != d || pr != p) continue;
        if (g >= 1 && g <= 8) M->grid = g * cus;
        char *tok = std::strtok(rest, ",");
        while (tok) {
            int L = 0, e = 0;
            if (std::sscanf(tok, "%d:%d", &L, &e) == 2 && L >= MANY_LMIN && L <= MANY_LMAX)
                M->sched[L] = (e == 8 && e8_ok(L)) ? 1 : 0;
            tok = std::strtok(nullptr, ",");
        }
    }
    std::fclose(fp);
}

int save_wisdom(ManyImpl *M, const char *path, int cus)
{
    const std::string dev = device_name(), pr = M->prec == PREC_F64 ? "f64" : "f32";
    std::vector<std::string> keep;
    if (FILE *fp = std::fopen(path, "r")) {
        char line[1024];
        while (std::fgets(line, sizeof line, fp)) {
            char d[256], p[16];
            if (std::sscanf(line, "%255s %15s", d, p) == 2 && dev == d && pr == p) continue;
            keep.push_back(line);
        }
        std::fclose(fp);
    }
    FILE *fp = std::fopen(path, "w");
    if (!fp) return TOPE_ERROR_INVALID_ARG;
    for (auto &l : keep) std::fputs(l.c_str(), fp);
    std::fprintf(fp, "%s %s grid_per_cu=%d sched=", dev.c_str(), pr.c_str(), std::max(1, M->grid / cus));
    bool firstc = true;
    for (int L = MANY_LMIN; L <= MANY_LMAX; ++L) {
        std::fprintf(fp, "%s%d:%d", firstc ? "" : ",", L, M->sched[L] ? 8 : 16);
        firstc = false;
    }
    std::fprintf(fp, "\n");
    std::fclose(fp);
    return TOPE_SUCCESS;
}

int device_cus()
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return 256;
    return cus;
}

// Mean ms of `reps` launches of the persistent kernel (optionally one class).
double time_persistent(ManyImpl *M, int reps, int only_class, int grid)
{
    if (!M->e0 && hipEventCreateWithFlags(&M->e0, hipEventDisableSystemFence) != hipSuccess) return -1;
    if (!M->e1 && hipEventCreateWithFlags(&M->e1, hipEventDisableSystemFence) != hipSuccess) return -1;
    if (launch_persistent(M, FORWARD, M->stream, only_class, grid)) return -1;  // warm-up
    if (hipEventRecord(M->e0, M->stream) != hipSuccess) return -1;
    for (int r = 0; r < reps; ++r)
        if (launch_persistent(M, FORWARD, M->stream, only_class, grid)) return -1;
    if (hipEventRecord(M->e1, M->stream) != hipSuccess || hipEventSynchronize(M->e1) != hipSuccess) return -1;
    float ms = 0;
    if (hipEventElapsedTime(&ms, M->e0, M->e1) != hipSuccess) return -1;
    return ms / reps;
}

}  // namespace

extern "C" {

void topeManyPlanInit(struct topeFFT *f, struct topePlanMany *t, int count, const int *sizes, int precision,
                      const void *host)
{
    if (!f || !t) return;
    std::memset(t, 0, sizeof(*t));
    t->count = count;
    t->precision = precision;
    hipStream_t s = fw_stream(f);
    if (!s) {
        f->error = TOPE_ERROR_NOT_INITIALIZED;
        return;
    }
    if (count < 1 || !sizes || !host || (precision != TOPE_PREC_F64 && precision != TOPE_PREC_F32)) {
        f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    for (int i = 0; i < count; ++i) {
        const int n = sizes[i];
        if (n < (1 << MANY_LMIN) || n > (1 << MANY_LMAX) || (n & (n - 1))) {
            f->error = TOPE_ERROR_INVALID_SIZE;
            return;
        }
    }
    ManyImpl *M = new (std::nothrow) ManyImpl();
    if (!M) {
        f->error = TOPE_ERROR_UNSUPPORTED;
        return;
    }
    M->fw = fw_handle(f);
    M->stream = s;
    M->prec = precision == TOPE_PREC_F64 ? PREC_F64 : PREC_F32;
    M->elem = M->prec == PREC_F64 ? 16 : 8;
    M->count = count;
    M->sizes.assign(sizes, sizes + count);
    M->offs.resize(count);
    long long acc = 0;
    for (int i = 0; i < count; ++i) {
        M->offs[i] = acc;
        acc += sizes[i];
    }
    M->elems = acc;
    const int cus = device_cus();
    M->grid = cus;            // one 1024-thread block per CU (LDS-bound for FP64)
    M->grid_small = 4 * cus;  // four 256-thread blocks per CU (registers / LDS)
    // defaults: radix-16 schedules everywhere
    for (int L = 0; L < 16; ++L) M->sched[L] = 0;
    load_wisdom(M, cus);

    auto fail = [&](int rc) {
        free_many(M);
        f->error = rc;
    };
    if ((long long)M->elems * (long long)M->elem >= 0x7fffffffLL) return fail(TOPE_ERROR_INVALID_SIZE);

    hipEvent_t p0 = nullptr, p1 = nullptr;
    hipError_t err = hipStreamCreateWithFlags(&M->side, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&M->fork, hipEventDisableTiming);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&M->join, hipEventDisableTiming);
    if (err != hipSuccess) return fail(map_err(err));
    err = hipEventCreate(&p0);
    if (err == hipSuccess) err = hipEventCreate(&p1);
    if (err == hipSuccess) err = hipEventRecord(p0, s);
    if (err == hipSuccess) err = hipMalloc(&M->data, (size_t)M->elems * M->elem);
    // twiddles for every class present, both schedules
    bool present[16] = {};
    for (int i = 0; i < count; ++i) present[ilog2i(sizes[i])] = true;
    for (int L = MANY_LMIN; L <= MANY_LMAX && err == hipSuccess; ++L) {
        if (!present[L]) continue;
        const int n = 1 << L;
        for (int k = 0; k < 2 && err == hipSuccess; ++k) {
            if (k == 1 && !e8_ok(L)) continue;
            const int e = k == 0 ? (n < 16 ? n : 16) : 8;
            HostSched h = host_sched(n, e);
            void *tw = nullptr;
            err = hipMalloc(&tw, (size_t)std::max(1, h.twsize) * M->elem);
            if (err != hipSuccess) break;
            M->allocs.push_back(tw);
            M->d_tw[L][k] = tw;
            err = gen_stage_twiddles(tw, M->prec, n, e, s);
        }
    }
    if (err == hipSuccess) err = hipStreamSynchronize(s);
    tope_ulong pre = 0, mem = 0;
    if (err == hipSuccess) err = hipEventRecord(p1, s);
    if (err == hipSuccess) err = hipEventSynchronize(p1);
    float ms = 0;
    if (err == hipSuccess && hipEventElapsedTime(&ms, p0, p1) == hipSuccess) pre = (tope_ulong)(ms * 1e6);
    if (err == hipSuccess) err = hipEventRecord(p0, s);
    if (err == hipSuccess)
        err = hipMemcpyAsync(M->data, host, (size_t)M->elems * M->elem, hipMemcpyHostToDevice, s);
    if (err == hipSuccess) err = hipEventRecord(p1, s);
    if (err == hipSuccess) err = hipEventSynchronize(p1);
    if (err == hipSuccess && hipEventElapsedTime(&ms, p0, p1) == hipSuccess) mem = (tope_ulong)(ms * 1e6);
    if (p0) (void)hipEventDestroy(p0);
    if (p1) (void)hipEventDestroy(p1);
    if (err != hipSuccess) return fail(map_err(err));
    err = build_tasks(M);
    if (err != hipSuccess) return fail(map_err(err));
    t->impl = M;
    t->data = M->data;
    t->dataSize = (size_t)M->elems * M->elem;
    t->elements = M->elems;
    t->totalPreKernel = pre;
    t->totalMemory = mem;
    t->method = M->method;
    fw_retain(f);
    f->error = TOPE_SUCCESS;
}

void topeManySetMethod(struct topeFFT *f, struct topePlanMany *t, int method)
{
    ManyImpl *M = impl_of(t);
    if (!M || method < TOPE_MANY_PERSISTENT || method > TOPE_MANY_PER_TRANSFORM_GRAPH) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    M->method = method;
    t->method = method;
    f->error = TOPE_SUCCESS;
}

void topeManyExecDevice(struct topeFFT *f, struct topePlanMany *t, int dir)
{
    ManyImpl *M = impl_of(t);
    if (!f || !M || M->fw != fw_handle(f)) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    if (dir != FORWARD && dir != INVERSE) {
        f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    f->error = launch_method(M, M->method, dir, M->stream);
}

double topeManyTimeExec(struct topeFFT *f, struct topePlanMany *t, int reps)
{
    ManyImpl *M = impl_of(t);
    if (!f || !M || M->fw != fw_handle(f) || reps < 1) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return -1.0;
    }
    hipError_t err = hipSuccess;
    if (!M->e0) err = hipEventCreateWithFlags(&M->e0, hipEventDisableSystemFence);
    if (err == hipSuccess && !M->e1) err = hipEventCreateWithFlags(&M->e1, hipEventDisableSystemFence);
    int rc = map_err(err);
    // one untimed exec first: a graph method captures its graph there
    if (!rc) rc = launch_method(M, M->method, FORWARD, M->stream);
    if (!rc) rc = map_err(hipEventRecord(M->e0, M->stream));
    for (int r = 0; r < reps && !rc; ++r) rc = launch_method(M, M->method, r % 2 ? INVERSE : FORWARD, M->stream);
    if (!rc) rc = map_err(hipEventRecord(M->e1, M->stream));
    if (!rc) rc = map_err(hipEventSynchronize(M->e1));
    float ms = 0;
    if (!rc) rc = map_err(hipEventElapsedTime(&ms, M->e0, M->e1));
    f->error = rc;
    return rc ? -1.0 : (double)ms / reps;
}

void topeManyExecPtr(struct topeFFT *f, struct topePlanMany *t, void *data, int dir)
{
    ManyImpl *M = impl_of(t);
    if (!f || !M || M->fw != fw_handle(f)) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    if (!data || (dir != FORWARD && dir != INVERSE) || reinterpret_cast<uintptr_t>(data) % M->elem) {
        f->error = TOPE_ERROR_INVALID_ARG;  // elements are loaded whole
        return;
    }
    // the graphs hold the plan's own array: the graph methods run their
    // launches directly here
    int method = M->method;
    if (method == TOPE_MANY_PERSISTENT_GRAPH) method = TOPE_MANY_PERSISTENT;
    if (method == TOPE_MANY_PER_TRANSFORM_GRAPH) method = TOPE_MANY_PER_TRANSFORM;
    void *own = M->data;
    M->data = data;
    const int rc = launch_method(M, method, dir, M->stream);
    M->data = own;
    f->error = rc;
}

void topeManyExec(struct topeFFT *f, struct topePlanMany *t, void *host, int dir)
{
    ManyImpl *M = impl_of(t);
    if (!f || !M || !host || M->fw != fw_handle(f)) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    if (dir != FORWARD && dir != INVERSE) {
        f->error = TOPE_ERROR_INVALID_ARG;
        return;
    }
    // the download below reads what the kernels wrote: fenced events
    if (!M->x0 && hipEventCreate(&M->x0) != hipSuccess) {
        f->error = TOPE_ERROR_UNSUPPORTED;
        return;
    }
    if (!M->x1 && hipEventCreate(&M->x1) != hipSuccess) {
        f->error = TOPE_ERROR_UNSUPPORTED;
        return;
    }
    hipError_t err = hipEventRecord(M->x0, M->stream);
    int rc = err == hipSuccess ? launch_method(M, M->method, dir, M->stream) : map_err(err);
    if (!rc) rc = map_err(hipEventRecord(M->x1, M->stream));
    if (!rc) rc = map_err(hipEventSynchronize(M->x1));
    float ms = 0;
    if (!rc && hipEventElapsedTime(&ms, M->x0, M->x1) == hipSuccess) t->totalKernel += (tope_ulong)(ms * 1e6);
    if (!rc)
        rc = map_err(hipMemcpyAsync(host, M->data, (size_t)M->elems * M->elem, hipMemcpyDeviceToHost, M->stream));
    if (!rc) rc = map_err(hipStreamSynchronize(M->stream));
    f->error = rc;
}

long long topeManyOffsets(struct topePlanMany *t, long long *offsets)
{
    ManyImpl *M = impl_of(t);
    if (!M) return -1;
    if (offsets) std::copy(M->offs.begin(), M->offs.end(), offsets);
    return M->count;
}

void topeManyGetSchedule(struct topePlanMany *t, int *radix_by_log2, int *grid)
{
    ManyImpl *M = impl_of(t);
    if (!M) return;
    if (radix_by_log2)
        for (int L = 0; L < 16; ++L) radix_by_log2[L] = M->sched[L] ? 8 : 16;
    if (grid) *grid = M->grid;
}

int topeManyTune(struct topeFFT *f, struct topePlanMany *t, const char *wisdom_path, int reps)
{
    ManyImpl *M = impl_of(t);
    if (!f || !M || M->fw != fw_handle(f) || reps < 1) {
        if (f) f->error = TOPE_ERROR_INVALID_ARG;
        return f ? f->error : TOPE_ERROR_INVALID_ARG;
    }
    // The device data is scratch during tuning (repeated forward transforms):
    // keep a copy and restore it afterwards.
    void *save = nullptr;
    hipError_t err = hipMalloc(&save, (size_t)M->elems * M->elem);
    if (err == hipSuccess)
        err = hipMemcpyAsync(save, M->data, (size_t)M->elems * M->elem, hipMemcpyDeviceToDevice, M->stream);
    if (err != hipSuccess) {
        f->error = map_err(err);
        return f->error;
    }
    const int cus = device_cus();
    bool present[16] = {};
    for (int n : M->sizes) present[ilog2i(n)] = true;
    int rc = TOPE_SUCCESS;
    // 1) radix schedule per class: time the persistent kernel on that class
    //    alone with each candidate schedule
    for (int L = MANY_LMIN; L <= MANY_LMAX && !rc; ++L) {
        if (!present[L] || !e8_ok(L)) continue;
        double best = -1;
        int best_s = 0;
        for (int cand = 0; cand < 2 && !rc; ++cand) {
            M->sched[L] = cand;
            if (build_tasks(M) != hipSuccess) {
                rc = TOPE_ERROR_UNSUPPORTED;
                break;
            }
            const double ms = time_persistent(M, reps, L, 0);
            if (ms < 0) rc = TOPE_ERROR_UNSUPPORTED;
            else if (best < 0 || ms < best) {
                best = ms;
                best_s = cand;
            }
        }
        M->sched[L] = best_s;
    }
    // 2) resident blocks per CU (LDS allows two 1024-thread blocks for FP32)
    if (!rc && build_tasks(M) == hipSuccess) {
        double best = -1;
        int best_g = cus;
        const int maxg = M->prec == PREC_F64 ? 1 : 2;
        for (int g = 1; g <= maxg && !rc; ++g) {
            const double ms = time_persistent(M, reps, -1, g * cus);
            if (ms < 0) rc = TOPE_ERROR_UNSUPPORTED;
            else if (best < 0 || ms < best) {
                best = ms;
                best_g = g * cus;
            }
        }
        M->grid = best_g;
    }
    if (!rc && build_tasks(M) != hipSuccess) rc = TOPE_ERROR_UNSUPPORTED;
    err = hipMemcpyAsync(M->data, save, (size_t)M->elems * M->elem, hipMemcpyDeviceToDevice, M->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(M->stream);
    (void)hipFree(save);
    if (!rc && err != hipSuccess) rc = map_err(err);
    if (!rc && wisdom_path) rc = save_wisdom(M, wisdom_path, cus);
    f->error = rc;
    return rc;
}

void topeManyDestroy(struct topeFFT *f, struct topePlanMany *t)
{
    ManyImpl *M = impl_of(t);
    if (!f || !M) return;
    void *owner = M->fw;
    free_many(M);
    t->impl = nullptr;
    t->data = nullptr;
    fw_release(f, owner);
    f->error = TOPE_SUCCESS;
}

}  // extern "C"
