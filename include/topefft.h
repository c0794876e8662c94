/*
 * topefft.h -- drop-in C interface of the MI355X (gfx950) topeFFT library.
 *
 * Source-compatible superset of the reference header src/topefft.h:
 * the same struct tags, public field names, constants and the $CHECKERROR
 * macro, with OpenCL handles replaced by opaque pointers and <stdint.h>
 * types, and with the prototypes the reference header never declared
 * (callers of the reference relied on implicit declarations).
 *
 * Data layout (unchanged): interleaved complex128, x fastest,
 *     d[2*((z*y_len + j)*x_len + i)] = re,  d[... + 1] = im.
 * Semantics (unchanged): PlanInit copies d to the device; Exec transforms
 * the device-resident copy in place and writes the result into d; FORWARD
 * uses W = e^{-2 pi i / N} unnormalised, INVERSE the conjugate and divides
 * by the total element count.
 *
 * Documented deviations from the reference (bugs not reproduced):
 *   - tope1DExec returns the transformed data (the reference reads an
 *     unallocated scratch buffer on its radix path, src/fft1d.c:140-142).
 *   - 1D sizes that are not powers of two are transformed in both
 *     directions (the reference's DFT kernel ignores `dir`,
 *     src/kernels1D.cl:549-573): 7-smooth lengths (factors 2, 3, 5, 7) up
 *     to 4096 by a mixed-radix pass, other lengths by a GPU DFT up to 64
 *     and Bluestein's chirp-z convolution above; 7-smooth lengths above
 *     4096 that split into two factors of at most 4096 run a four-step of
 *     two mixed-radix passes.  2D/3D axes that are not powers of two run the
 *     same way (Bluestein for lengths up to 2^29) where the reference exits
 *     (src/fft2d.c:355-362).  Power-of-two axes run one pass up to 4096,
 *     the four-step (two passes) up to 2^24 and three passes up to 2^30.
 *   - the n == 2 1D inverse divides both real and imaginary parts by 2
 *     (src/fft1d.c:23-26 divides only the real parts).
 *   - tope*Destroy frees the plan; the framework (stream) is released when
 *     its last plan is destroyed or by topeFFTDestroy (the reference
 *     releases the whole OpenCL context on every Destroy, src/fft2d.c:453-462).
 *   - errors set f->error and return instead of assert/exit
 *     (src/fft2d.c:355-362, src/util.c:149-151).
 */
#ifndef TOPEFFT_H
#define TOPEFFT_H

#include <assert.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* reference src/topefft.h:4-11 */
#define MAX_BINARY_SIZE (0x100000)
#define MAX_SOURCE_SIZE (0x100000)
#define TOPE_SUCCESS 0
#define $CHECKERROR assert(f->error == TOPE_SUCCESS);

#define C2C 1
#define R2C 0
#define FORWARD 1
#define INVERSE 0

/* status codes stored in f->error (HIP runtime errors are stored as their
 * positive hipError_t value) */
#define TOPE_ERROR_INVALID_SIZE (-1001)
#define TOPE_ERROR_INVALID_ARG (-1002)
#define TOPE_ERROR_NO_DEVICE (-1003)
#define TOPE_ERROR_NOT_INITIALIZED (-1004)
#define TOPE_ERROR_NO_KERNEL (-1005)
#define TOPE_ERROR_UNSUPPORTED (-1006)
#define TOPE_ERROR_SCRATCH_LIMIT (-1007)

typedef uint64_t tope_ulong; /* was cl_ulong */

/* Plan structs: reference src/topefft.h:13-81.  `data` is the device buffer
 * holding the transform (interleaved complex128); `impl` is the opaque
 * schedule.  Counters are nanoseconds, filled from HIP events:
 * totalMemory = host<->device copies, totalPreKernel = plan-time twiddle
 * generation, totalKernel = transform kernels. */
struct topePlan1D {
    int x;           /* length */
    int log;         /* reference stage count of the radix rule */
    int type;        /* C2C (R2C is stored, as in the reference, and runs C2C) */
    int radix;       /* reference radix rule: 8, 4, 2, or -1 (DFT path) */
    void *data;      /* device buffer (complex128, x elements) */
    void *scratch;   /* device scratch for multi-pass schedules, or NULL */
    void *twiddle;   /* device twiddle tables */
    size_t dataSize; /* bytes */
    tope_ulong totalMemory;
    tope_ulong totalPreKernel;
    tope_ulong totalKernel;
    unsigned int dim;
    void *impl;
};

struct topePlan2D {
    int x, y;
    int logX, logY;
    int type;
    int radX, radY;
    void *data;
    void *scratch;
    void *twdX, *twdY;
    size_t dataSize;
    tope_ulong totalMemory;
    tope_ulong totalPreKernel;
    tope_ulong totalKernel;
    unsigned int dim;
    void *impl;
};

struct topePlan3D {
    int x, y, z;
    int logX, logY, logZ;
    int type;
    int radX, radY, radZ;
    void *data;
    void *scratch;
    void *twdX, *twdY, *twdZ;
    size_t dataSize;
    tope_ulong totalMemory;
    tope_ulong totalPreKernel;
    tope_ulong totalKernel;
    unsigned int dim;
    void *impl;
};

/* Framework state: reference src/topefft.h:83-95. */
struct topeFFT {
    int error;                      /* last status, TOPE_SUCCESS == 0 */
    void *event;                    /* event pair of the last exec (profileThis) */
    int device;                     /* HIP device ordinal */
    void *stream;                   /* hipStream_t all plans enqueue on */
    void *impl;                     /* opaque framework state */
    unsigned int ret_num_platforms; /* 1 */
    unsigned int ret_num_devices;   /* HIP devices visible */
};

/* ---- reference entry points ------------------------------------------ */

/* src/util.c:21 */
void topeFFTInit(struct topeFFT *f);

/* src/fft1d.c:303 / :12 / :372 */
void tope1DPlanInit(struct topeFFT *f, struct topePlan1D *t, int x, int type, double *d);
void tope1DExec(struct topeFFT *f, struct topePlan1D *t, double *d, int dir);
void tope1DDestroy(struct topeFFT *f, struct topePlan1D *t);

/* src/fft2d.c:319 / :136 / :442 */
void tope2DPlanInit(struct topeFFT *f, struct topePlan2D *t, int x, int y, int type, double *d);
void tope2DExec(struct topeFFT *f, struct topePlan2D *t, double *d, int dir);
void tope2DDestroy(struct topeFFT *f, struct topePlan2D *t);

/* src/fft3d.c:474 / :216 / :628 */
void tope3DPlanInit(struct topeFFT *f, struct topePlan3D *t, int x, int y, int z, int type, double *d);
void tope3DExec(struct topeFFT *f, struct topePlan3D *t, double *d, int dir);
void tope3DDestroy(struct topeFFT *f, struct topePlan3D *t);

/* src/util.c:12 -- END-START nanoseconds of an event; pass f->event for the
 * device time of the last transform (waits for it to finish). */
tope_ulong profileThis(void *event);

/* ---- additions (no reference counterpart) ------------------------------ */

/* Release the framework's stream and state (the reference folds this into
 * every tope*Destroy). */
void topeFFTDestroy(struct topeFFT *f);

/* Device-resident execution: transform t->data in place on f->stream and
 * return without waiting and without any host copy. */
void tope1DExecDevice(struct topeFFT *f, struct topePlan1D *t, int dir);
void tope2DExecDevice(struct topeFFT *f, struct topePlan2D *t, int dir);
void tope3DExecDevice(struct topeFFT *f, struct topePlan3D *t, int dir);

/* Copy between the host and a plan's device buffer (blocking). */
void topePlanUpload(struct topeFFT *f, void *plan_impl, const void *host, size_t bytes);
void topePlanDownload(struct topeFFT *f, void *plan_impl, void *host, size_t bytes);

/* Wait for all work enqueued on f->stream. */
void topeFFTSynchronize(struct topeFFT *f);

/* Number of kernel launches one Exec performs, and per-launch timing:
 * while enabled, every Exec records HIP events around each launch on
 * f->stream; topePlanPassTimes synchronises and returns the mean duration
 * in milliseconds of each launch position over the recorded execs. */
int topePlanPassCount(void *plan_impl);
void topePlanPassTiming(void *plan_impl, int enable);
int topePlanPassTimes(struct topeFFT *f, void *plan_impl, double *ms, int max_passes);
/* Mean device time in milliseconds of launch position `pass` alone: HIP
 * events around `reps` back-to-back launches of that pass on f->stream (no
 * events between the launches).  Repeating one pass in isolation leaves the
 * plan's device data in an undefined state: upload again before the next
 * transform whose result matters. */
double topePlanTimePass(struct topeFFT *f, void *plan_impl, int pass, int dir, int reps);
/* Kernel name of launch position `pass` (static string). */
const char *topePlanPassName(void *plan_impl, int pass);

/* ---- batched 1D transforms (no reference counterpart) -------------------
 * `count` independent length-`n` C2C transforms stored back to back.  FP64
 * plans hold interleaved complex128 (host pointer: double*), FP32 plans
 * complex64 (host pointer: float*).  FP32 plans choose how the radix-16
 * stages run:
 *   TOPE_MODE_VALU        FP32 butterflies on the vector ALUs
 *   TOPE_MODE_MFMA_BF16X3 16x16 DFT-matrix products on MFMA, operands split
 *                         into 2 bf16 parts, 3 products (~1e-5 rel-L2)
 *   TOPE_MODE_MFMA_BF16X6 3 bf16 parts, 6 products (~FP32 accuracy)
 *   TOPE_MODE_MFMA_FP32   plain FP32 operands on MFMA
 *   TOPE_MODE_MFMA_FP16X3 2 fp16 parts, 3 products (~FP32 accuracy), each
 *                         transform scaled by a power of two into fp16 range
 * MFMA modes take n = 256, 512, 1024, 2048 or 4096.  Twiddles and DFT-matrix constants are
 * generated in FP64 on the GPU and rounded once; INVERSE scales by 1/n.
 * Both precisions take powers of two and 7-smooth n (mixed-radix pass up to
 * 4096, a four-step of two such passes above; VALU mode); FP64 plans also
 * take other n up to 2^29 (Bluestein). */
#define TOPE_PREC_F64 0
#define TOPE_PREC_F32 1
#define TOPE_MODE_VALU 0
#define TOPE_MODE_MFMA_BF16X3 1
#define TOPE_MODE_MFMA_BF16X6 2
#define TOPE_MODE_MFMA_FP32 3
#define TOPE_MODE_MFMA_FP16X3 4

struct topePlanBatch {
    int n;
    int count;
    int precision;
    int mode;
    void *data; /* device buffer, count * n complex elements */
    size_t dataSize;
    tope_ulong totalMemory;
    tope_ulong totalPreKernel;
    tope_ulong totalKernel;
    void *impl;
};

void topeBatchPlanInit(struct topeFFT *f, struct topePlanBatch *t, int n, int count, int precision, int mode,
                       const void *host);
void topeBatchExec(struct topeFFT *f, struct topePlanBatch *t, void *host, int dir);
void topeBatchExecDevice(struct topeFFT *f, struct topePlanBatch *t, int dir);
void topeBatchDestroy(struct topeFFT *f, struct topePlanBatch *t);

/* 2D/3D plans in a chosen precision: TOPE_PREC_F64 (complex128, host
 * double*, identical to tope2DPlanInit/tope3DPlanInit) or TOPE_PREC_F32
 * (complex64 storage, host float*, FP32 butterflies with FP64-generated
 * twiddles; axes powers of two or 7-smooth).  Exec/ExecDevice/Destroy as for the reference plans; the Ex
 * Exec variants take the host buffer as void*. */
void tope2DPlanInitEx(struct topeFFT *f, struct topePlan2D *t, int x, int y, int type, int precision,
                      const void *host);
void tope3DPlanInitEx(struct topeFFT *f, struct topePlan3D *t, int x, int y, int z, int type, int precision,
                      const void *host);
/* Complex64 2D/3D plans whose axis passes run on the matrix cores in the
 * given TOPE_MODE_MFMA_* operand mode where a kernel covers the axis:
 * contiguous rows of 256 to 4096, unit-stride column axes of 256, 512 or
 * 1024 with a multiple of 16 lines; other axes take the FP32 VALU pass
 * (topePlanPassName tells which).  precision must be TOPE_PREC_F32 unless
 * mode is TOPE_MODE_VALU (then identical to the Ex functions). */
void tope2DPlanInitMode(struct topeFFT *f, struct topePlan2D *t, int x, int y, int type, int precision, int mode,
                        const void *host);
void tope3DPlanInitMode(struct topeFFT *f, struct topePlan3D *t, int x, int y, int z, int type, int precision,
                        int mode, const void *host);
void tope2DExecEx(struct topeFFT *f, struct topePlan2D *t, void *host, int dir);
void tope3DExecEx(struct topeFFT *f, struct topePlan3D *t, void *host, int dir);

/* ---- caller streams and caller-owned device arrays (no reference
 * counterpart) ------------------------------------------------------------
 * topeFFTSetStream: every later operation of the framework's plans is
 * enqueued on `stream` (a hipStream_t; NULL restores the framework's own
 * stream).  Work already queued on the previous stream is ordered before
 * it.  Batch plans (topeMany*) keep the stream they were created with.
 * topePlanExecPtr: transforms the caller's device array `data` (the plan's
 * element type and layout, interleaved complex) in place, asynchronously on
 * the framework's stream, without host copies; plan_impl is the plan
 * struct's impl field.  Not for real-input plans (topeRealExecPtr) or
 * packed uneven batches (topeManyExecPtr).  The array must be aligned to
 * its element size (16 bytes complex128, 8 bytes complex64), and to 16
 * bytes for complex64 plans whose strided 4096-point axes move two columns
 * as one 16-byte element; otherwise TOPE_ERROR_INVALID_ARG, before anything
 * is enqueued (device allocations and tensors at offset 0 are aligned).
 * The ExecPtr calls enqueue the
 * plan's passes on the caller's arrays and return; the plan can take the
 * next call at once (its own arrays are left untouched).  One host thread at
 * a time per framework and its plans. */
void topeFFTSetStream(struct topeFFT *f, void *stream);
void topePlanExecPtr(struct topeFFT *f, void *plan_impl, void *data, int dir);

/* ---- real-input transforms (no reference counterpart; the reference
 * declares R2C but runs C2C) ---------------------------------------------
 * x (even) by y by z (1 for unused axes), FP64, x fastest; x/2, y and z
 * each a power of two up to 2^30 or 7-smooth (factors 2, 3, 5, 7) up to 4096,
 * or above when it splits into two 7-smooth factors of at most 4096
 * (others: TOPE_ERROR_UNSUPPORTED, odd x: TOPE_ERROR_INVALID_SIZE).
 * PlanInit uploads the real array (x*y*z doubles).
 * topeRealExec FORWARD (R2C) transforms the device real array into the half
 * spectrum, (x/2+1)*y*z complex128 in the layout of numpy's rfftn (kx
 * fastest), and downloads it into host (2*(x/2+1)*y*z doubles; on the
 * device the rows of x/2+1 values sit at a pitch of `pitch` elements); INVERSE
 * (C2R) transforms the device half spectrum back into the real array,
 * scaled by 1/(x*y*z), and downloads x*y*z doubles.  FORWARD consumes the
 * real array and INVERSE the half spectrum (in-place pipeline). */
struct topePlanReal {
    int x, y, z, dim;
    int pitch;       /* device half-spectrum row pitch (elements) */
    int precision;   /* TOPE_PREC_F64 or TOPE_PREC_F32 (topeRealPlanInitEx) */
    void *data;      /* device half spectrum */
    void *real;      /* device real array */
    size_t dataSize; /* bytes of data */
    size_t realSize; /* bytes of real */
    tope_ulong totalMemory, totalPreKernel, totalKernel;
    void *impl;
};
void topeRealPlanInit(struct topeFFT *f, struct topePlanReal *t, int x, int y, int z, const double *real);
void topeRealExec(struct topeFFT *f, struct topePlanReal *t, double *host, int dir);
void topeRealExecDevice(struct topeFFT *f, struct topePlanReal *t, int dir);
/* Precision variant: TOPE_PREC_F64 (as above) or TOPE_PREC_F32 (float real
 * array, complex64 half spectrum, FP32 arithmetic with FP64-generated
 * tables rounded once; the same sizes).
 * ExecEx downloads into `host` as Exec does, in the plan's precision. */
void topeRealPlanInitEx(struct topeFFT *f, struct topePlanReal *t, int x, int y, int z, int precision,
                        const void *real);
void topeRealExecEx(struct topeFFT *f, struct topePlanReal *t, void *host, int dir);
/* topeRealExecPtr: the same transform on caller-owned device arrays, in
 * the plan's device layout and precision: `real` holds x*y*z values
 * (realSize bytes), `spectrum` dataSize bytes (rows of x/2+1 complex values
 * at a pitch of `pitch` elements).  Asynchronous on the framework's stream, no host copies.
 * FORWARD writes `spectrum` and may overwrite `real`; INVERSE writes `real`
 * and may overwrite `spectrum`.  Both arrays aligned to the complex element
 * size (16 bytes FP64, 8 bytes FP32), the spectrum to 16 bytes for complex64
 * plans whose 4096-point column passes move column pairs as 16-byte
 * elements (TOPE_ERROR_INVALID_ARG otherwise).
 * Plans with y or z above 4096 use the real array's buffer as dataSize
 * bytes of scratch and are refused (TOPE_ERROR_UNSUPPORTED). */
void topeRealExecPtr(struct topeFFT *f, struct topePlanReal *t, void *real, void *spectrum, int dir);
void topeRealDestroy(struct topeFFT *f, struct topePlanReal *t);

/* ---- packed batch of uneven power-of-two lengths (no reference
 * counterpart) ---------------------------------------------------------
 * `count` transforms with lengths sizes[i] (powers of two, 16 .. 16384)
 * stored back to back: transform i starts at element sum_{k<i} sizes[k]
 * (topeManyOffsets).  Methods:
 *   TOPE_MANY_PERSISTENT            persistent kernels for all (one for the
 *                                   lengths above 4096, one beside it for
 *                                   the others)
 *   TOPE_MANY_PERSISTENT_GRAPH      those launches replayed from a hipGraph
 *   TOPE_MANY_PER_TRANSFORM         one kernel launch per transform
 *   TOPE_MANY_PER_TRANSFORM_GRAPH   those launches replayed from one hipGraph
 * topeManyTune times the candidate radix schedules of every length class
 * (radix-16 or radix-8 stages) and the resident blocks per CU on the device,
 * keeps the fastest, and appends them to the wisdom file `wisdom_path`;
 * TOPEFFT_WISDOM=<path> makes PlanInit reuse them.  Tuning transforms the
 * device data repeatedly and restores it afterwards. */
#define TOPE_MANY_PERSISTENT 0
#define TOPE_MANY_PERSISTENT_GRAPH 1
#define TOPE_MANY_PER_TRANSFORM 2
#define TOPE_MANY_PER_TRANSFORM_GRAPH 3

struct topePlanMany {
    int count;
    int precision; /* TOPE_PREC_F64 (complex128) or TOPE_PREC_F32 (complex64) */
    int method;
    int reserved;
    void *data; /* device buffer, packed transforms */
    size_t dataSize;
    long long elements;
    tope_ulong totalMemory;
    tope_ulong totalPreKernel;
    tope_ulong totalKernel;
    void *impl;
};

void topeManyPlanInit(struct topeFFT *f, struct topePlanMany *t, int count, const int *sizes, int precision,
                      const void *host);
void topeManySetMethod(struct topeFFT *f, struct topePlanMany *t, int method);
void topeManyExec(struct topeFFT *f, struct topePlanMany *t, void *host, int dir);
void topeManyExecDevice(struct topeFFT *f, struct topePlanMany *t, int dir);
/* topeManyExecPtr: the plan's transforms on a caller-owned device array of
 * the same packed layout (dataSize bytes), in place, asynchronously on the
 * plan's stream; the graph methods run their launches without the graph
 * (the graphs hold the plan's own array).  The array aligned to the element
 * size (TOPE_ERROR_INVALID_ARG otherwise). */
void topeManyExecPtr(struct topeFFT *f, struct topePlanMany *t, void *data, int dir);
/* topeManyTimeExec: mean device time in milliseconds of one exec of the
 * plan's current method, from HIP events around `reps` back-to-back execs
 * on the plan's stream (directions alternating, FORWARD first; the graph
 * methods replay their graphs), after one untimed exec.  Transforms the
 * plan's array. */
double topeManyTimeExec(struct topeFFT *f, struct topePlanMany *t, int reps);
long long topeManyOffsets(struct topePlanMany *t, long long *offsets);
void topeManyGetSchedule(struct topePlanMany *t, int *radix_by_log2, int *grid);
int topeManyTune(struct topeFFT *f, struct topePlanMany *t, const char *wisdom_path, int reps);
void topeManyDestroy(struct topeFFT *f, struct topePlanMany *t);

/* Size of the public structs, so foreign-language bindings can check their
 * mirrors of this header. which: 0 topeFFT, 1 topePlan1D, 2 topePlan2D,
 * 3 topePlan3D, 4 topePlanBatch, 5 topePlanMany, 6 topePlanReal. */
size_t topeStructSize(int which);

/* Human-readable message for a status code. */
const char *topeErrorString(int code);

/* Copy the stage twiddle table the plan generated for axis `axis` into host
 * memory as interleaved doubles (forward sign), returning the entry count;
 * with host == NULL only the count is returned.  Lets tests pin the GPU's
 * FP64 twiddle generation against exact values. */
long long topePlanTwiddles(struct topeFFT *f, void *plan_impl, int axis, double *host, long long max_entries);

/* ---- allocation-free transforms (no reference counterpart) -------------
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.
 * topeExec1DNoAlloc transforms `count` length-n complex128 transforms stored
 * back to back in the caller's device array `data`, in place, on f->stream,
 * without a plan and without allocating: it runs the kernels of a 1D plan's
 * topePlanExecPtr, with their twiddle tables and (n > 4096) the four-step's
 * intermediate array in a static device array of the library (one per GPU,
 * held until the process ends).  n: power of two 1 .. 2^24; for n > 4096,
 * count * n <= TOPE_NOALLOC_MAX_ELEMENTS, else f->error =
 * TOPE_ERROR_SCRATCH_LIMIT.  The static array is shared by every call on a
 * GPU: calls that overlap on one GPU (other streams, other host threads)
 * race.  FORWARD unnormalised, INVERSE scaled by 1/n.  Asynchronous.
 * `data` 16-byte aligned (TOPE_ERROR_INVALID_ARG otherwise). */
#define TOPE_NOALLOC_MAX_ELEMENTS (1 << 21)
void topeExec1DNoAlloc(struct topeFFT *f, void *data, int n, int count, int dir);

/* Name of the index-th compiled axis-pass kernel variant (the names
 * topePlanPassName reports and TOPEFFT_VARIANT_<n>_<row|col>=<name>
 * selects), or NULL past the last one.  Lets tests run every variant. */
const char *topeAxisVariantName(int index);

#ifdef __cplusplus
}
#endif

#endif /* TOPEFFT_H */
