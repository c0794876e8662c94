// Compiled instantiations of the fused axis kernel and their registry.
//
// Block shapes are chosen for gfx950: wave64, 160 KiB LDS per CU.  Row
// variants (contiguous axis) put P = N/E consecutive lanes on one transform
// so each load instruction covers 64 consecutive complex elements.  Column
// variants (strided axis) put TB transforms on consecutive lanes so a load
// instruction covers 64/TB rows of TB contiguous elements (TB*16 bytes for
// FP64); the split-exchange column variants trade two extra barriers per
// stage for twice as many columns per block.
#include "axis_ct3.hpp"
#include "axis_dit2.hpp"
#include "internal.hpp"

namespace tope {

template <typename R, int N, int E, int TB, bool CFAST, bool SPLIT, bool BIG, bool DIF2>
static void launch_axis_t(const void *in, void *out, const void *args, unsigned grid, hipStream_t s, int sign)
{
    constexpr int NT = TB * (N / E);
    const AxisArgs<R> &a = *static_cast<const AxisArgs<R> *>(args);
    const cpx<R> *i = static_cast<const cpx<R> *>(in);
    cpx<R> *o = static_cast<cpx<R> *>(out);
    if constexpr (CFAST) {
        // four-step first steps are strided tiles: only they carry the post-twiddle
        if (a.pt_n) {
            if (sign < 0)
                hipLaunchKernelGGL((axis_kernel<R, N, E, TB, CFAST, SPLIT, -1, true, BIG>), dim3(grid), dim3(NT), 0,
                                   s, i, o, a);
            else
                hipLaunchKernelGGL((axis_kernel<R, N, E, TB, CFAST, SPLIT, 1, true, BIG>), dim3(grid), dim3(NT), 0,
                                   s, i, o, a);
            return;
        }
    }
    if (sign < 0)
        hipLaunchKernelGGL((axis_kernel<R, N, E, TB, CFAST, SPLIT, -1, false, BIG, DIF2>), dim3(grid), dim3(NT), 0, s,
                           i, o, a);
    else
        hipLaunchKernelGGL((axis_kernel<R, N, E, TB, CFAST, SPLIT, 1, false, BIG, DIF2>), dim3(grid), dim3(NT), 0, s,
                           i, o, a);
}

// Even/odd-split column tiles (dit2_kernel): NT = TB * (N/2) / E threads.
template <typename R, int N, int E, int TB, bool SPLIT, bool BIG>
static void launch_dit2_t(const void *in, void *out, const void *args, unsigned grid, hipStream_t s, int sign)
{
    constexpr int NT = TB * (N / 2 / E);
    const AxisArgs<R> &a = *static_cast<const AxisArgs<R> *>(args);
    const cpx<R> *i = static_cast<const cpx<R> *>(in);
    cpx<R> *o = static_cast<cpx<R> *>(out);
    if (a.pt_n) {
        if (sign < 0)
            hipLaunchKernelGGL((dit2_kernel<R, N, E, TB, SPLIT, -1, true, BIG>), dim3(grid), dim3(NT), 0, s, i, o, a);
        else
            hipLaunchKernelGGL((dit2_kernel<R, N, E, TB, SPLIT, 1, true, BIG>), dim3(grid), dim3(NT), 0, s, i, o, a);
    } else {
        if (sign < 0)
            hipLaunchKernelGGL((dit2_kernel<R, N, E, TB, SPLIT, -1, false, BIG>), dim3(grid), dim3(NT), 0, s, i, o, a);
        else
            hipLaunchKernelGGL((dit2_kernel<R, N, E, TB, SPLIT, 1, false, BIG>), dim3(grid), dim3(NT), 0, s, i, o, a);
    }
}

// Three-step column tiles (ct3_kernel): 4096 points, 4 columns, 1024 threads.
template <typename R, bool BIG>
static void launch_ct3_t(const void *in, void *out, const void *args, unsigned grid, hipStream_t s, int sign)
{
    const AxisArgs<R> &a = *static_cast<const AxisArgs<R> *>(args);
    const cpx<R> *i = static_cast<const cpx<R> *>(in);
    cpx<R> *o = static_cast<cpx<R> *>(out);
    if (a.pt_n) {
        if (sign < 0) hipLaunchKernelGGL((ct3_kernel<R, -1, true, BIG>), dim3(grid), dim3(1024), 0, s, i, o, a);
        else hipLaunchKernelGGL((ct3_kernel<R, 1, true, BIG>), dim3(grid), dim3(1024), 0, s, i, o, a);
    } else {
        if (sign < 0) hipLaunchKernelGGL((ct3_kernel<R, -1, false, BIG>), dim3(grid), dim3(1024), 0, s, i, o, a);
        else hipLaunchKernelGGL((ct3_kernel<R, 1, false, BIG>), dim3(grid), dim3(1024), 0, s, i, o, a);
    }
}

// Complex64 three-step tiles (ct3_pair_kernel): 4096 points, 8 columns in
// pairs, 1024 threads.
template <bool BIG>
static void launch_ct3p_t(const void *in, void *out, const void *args, unsigned grid, hipStream_t s, int sign)
{
    const AxisArgs<float> &a = *static_cast<const AxisArgs<float> *>(args);
    const cpx<float> *i = static_cast<const cpx<float> *>(in);
    cpx<float> *o = static_cast<cpx<float> *>(out);
    if (a.pt_n) {
        if (sign < 0) hipLaunchKernelGGL((ct3_pair_kernel<-1, true, BIG>), dim3(grid), dim3(1024), 0, s, i, o, a);
        else hipLaunchKernelGGL((ct3_pair_kernel<1, true, BIG>), dim3(grid), dim3(1024), 0, s, i, o, a);
    } else {
        if (sign < 0) hipLaunchKernelGGL((ct3_pair_kernel<-1, false, BIG>), dim3(grid), dim3(1024), 0, s, i, o, a);
        else hipLaunchKernelGGL((ct3_pair_kernel<1, false, BIG>), dim3(grid), dim3(1024), 0, s, i, o, a);
    }
}

#ifdef TOPE_AXIS_BIG
#define TOPE_LAUNCH_CT3P &launch_ct3p_t<true>
#define TOPE_LAUNCH_CT3(R) &launch_ct3_t<R, true>
#define TOPE_LAUNCH(R, N, E, TB, CF, SP, DF) &launch_axis_t<R, N, E, TB, CF, SP, true, DF>
#define TOPE_LAUNCH_DIT2(R, N, E, TB, SP) &launch_dit2_t<R, N, E, TB, SP, true>
#else
#define TOPE_LAUNCH_CT3P &launch_ct3p_t<false>
#define TOPE_LAUNCH_CT3(R) &launch_ct3_t<R, false>
#define TOPE_LAUNCH(R, N, E, TB, CF, SP, DF) &launch_axis_t<R, N, E, TB, CF, SP, false, DF>
#define TOPE_LAUNCH_DIT2(R, N, E, TB, SP) &launch_dit2_t<R, N, E, TB, SP, false>
#endif

template <typename R>
constexpr int prec_of() { return sizeof(R) == 8 ? PREC_F64 : PREC_F32; }

#define TOPE_LAYOUT_true "_col"
#define TOPE_LAYOUT_false "_row"
#define TOPE_SPLIT_true "_split"
#define TOPE_SPLIT_false ""
#define TOPE_DIF2_true "_dif2"
#define TOPE_DIF2_false ""
#define TOPE_STR2(x) #x
#define TOPE_STR(x) TOPE_STR2(x)
#define VARD(R, N, E, TB, CF, SP, DF)                                                                        \
    AxisVariant{prec_of<R>(), N, E, TB, CF, SP, DF, false, false, TB * (N / E), Sched<N, E>::TWSIZE,        \
                Exchange<R, N, E, TB, CF, SP>::BYTES, TOPE_LAUNCH(R, N, E, TB, CF, SP, DF),                \
                "axis_" TOPE_STR(R) "_n" TOPE_STR(N) "_e" TOPE_STR(E) "_tb" TOPE_STR(TB) TOPE_LAYOUT_##CF \
                    TOPE_SPLIT_##SP TOPE_DIF2_##DF}
#define VAR(R, N, E, TB, CF, SP) VARD(R, N, E, TB, CF, SP, false)
#define VARH(R, N, E, TB, SP)                                                                                \
    AxisVariant{prec_of<R>(), N, E, TB, true, SP, false, true, false, TB * (N / 2 / E), Sched<N / 2, E>::TWSIZE, \
                Dit2Lds<R, N, E, TB, SP>::BYTES, TOPE_LAUNCH_DIT2(R, N, E, TB, SP),                        \
                "axis_" TOPE_STR(R) "_n" TOPE_STR(N) "_e" TOPE_STR(E) "_tb" TOPE_STR(TB) "_col" TOPE_SPLIT_##SP "_dit2"}

// Three-step column tile (4096 points, 4 columns, 16 elements per thread).
#define VARC(R)                                                                                               \
    AxisVariant{prec_of<R>(), 4096, 16, 4, true, true, false, false, true, 1024, CT3_TW, Ct3Lds<R, true>::BYTES, \
                TOPE_LAUNCH_CT3(R), "axis_" TOPE_STR(R) "_n4096_e16_tb4_col_ct3"}

// Complex64 three-step column tile (4096 points, 8 columns as 4 pairs).
#define VARCP                                                                                                 \
    AxisVariant{PREC_F32, 4096, 16, 8, true, true, false, false, true, 1024, CT3_TW, Ct3PairLds<true>::BYTES,       \
                TOPE_LAUNCH_CT3P, "axis_float_n4096_e16_tb8_col_ct3"}

// Even/odd-split column tiles for the long strided axes.
#define DIT2COLS(R) VARH(R, 4096, 8, 4, true), VARH(R, 4096, 16, 4, true), VARH(R, 4096, 16, 4, false), \
    VARH(R, 4096, 8, 4, false)

// Row-pair variants (2D passes whose next axis takes a fused radix-2 DIF
// step): the row table's tiles, at least two rows, split exchange.
#define DIF2ROWS(R)                                                                                              \
    VARD(R, 2, 2, 256, false, true, true), VARD(R, 4, 4, 256, false, true, true),                               \
        VARD(R, 8, 8, 256, false, true, true), VARD(R, 16, 16, 256, false, true, true),                         \
        VARD(R, 32, 16, 128, false, true, true), VARD(R, 64, 16, 64, false, true, true),                        \
        VARD(R, 128, 16, 32, false, true, true), VARD(R, 256, 16, 16, false, true, true),                       \
        VARD(R, 512, 16, 8, false, true, true), VARD(R, 1024, 16, 4, false, true, true),                        \
        VARD(R, 2048, 16, 2, false, true, true), VARD(R, 4096, 16, 2, false, true, true)

// Row variants: 256 threads per block.
#define ROWS(R)                                                                                     \
    VAR(R, 2, 2, 256, false, false), VAR(R, 4, 4, 256, false, false), VAR(R, 8, 8, 256, false, false), \
        VAR(R, 16, 16, 256, false, false), VAR(R, 32, 16, 128, false, false),                         \
        VAR(R, 64, 16, 64, false, false), VAR(R, 128, 16, 32, false, false),                          \
        VAR(R, 256, 16, 16, false, false), VAR(R, 512, 16, 8, false, false),                          \
        VAR(R, 1024, 16, 4, false, false), VAR(R, 2048, 16, 2, false, false),                         \
        VAR(R, 4096, 16, 1, false, false), VAR(R, 4096, 16, 1, false, true), VAR(R, 4096, 16, 2, false, true), \
        VAR(R, 2048, 16, 2, false, true), VAR(R, 1024, 16, 4, false, true), VAR(R, 512, 16, 8, false, true),   \
        VAR(R, 256, 16, 16, false, true), VAR(R, 256, 16, 4, false, false), VAR(R, 256, 16, 1, false, false)

// Column variants: TB columns per block.
#define COLS(R)                                                                                     \
    VAR(R, 2, 2, 64, true, false), VAR(R, 4, 4, 64, true, false), VAR(R, 8, 8, 64, true, false),    \
        VAR(R, 16, 16, 64, true, false), VAR(R, 32, 16, 32, true, false),                           \
        VAR(R, 64, 16, 16, true, false), VAR(R, 128, 16, 16, true, false),                          \
        VAR(R, 256, 16, 16, true, false), VAR(R, 512, 16, 8, true, false),                          \
        VAR(R, 1024, 16, 4, true, false), VAR(R, 1024, 16, 8, true, true),                          \
        VAR(R, 2048, 16, 2, true, false), VAR(R, 2048, 16, 4, true, true),                          \
        VAR(R, 4096, 16, 1, true, false), VAR(R, 4096, 16, 2, true, true), VAR(R, 4096, 16, 4, true, true), \
        VAR(R, 4096, 32, 2, true, false), VAR(R, 4096, 32, 4, true, true), VAR(R, 2048, 32, 4, true, false), \
        VAR(R, 2048, 32, 8, true, true), VAR(R, 512, 16, 16, true, true), VAR(R, 512, 16, 32, true, true),      \
        VAR(R, 256, 16, 32, true, true), VAR(R, 256, 16, 64, true, true), VAR(R, 256, 16, 4, true, false),     \
        VAR(R, 256, 4, 1, true, false), VAR(R, 512, 8, 1, true, false), VAR(R, 1024, 16, 1, true, false), \
        VAR(R, 2048, 16, 1, true, false)

static const AxisVariant kVariants[] = {ROWS(double), COLS(double), DIF2ROWS(double), ROWS(float), COLS(float),
                                        DIT2COLS(double), VARH(float, 4096, 16, 8, true),
                                        VARH(float, 4096, 16, 8, false), VAR(float, 4096, 32, 8, true, true),
                                        VARC(double), VARCP};

// The same list is compiled twice: normally (32-bit buffer offsets) and with
// TOPE_AXIS_BIG (64-bit pointers) for blocks spanning more than 2^31 bytes.
#ifdef TOPE_AXIS_BIG
const AxisVariant *axis_variants_big(int *count)
#else
const AxisVariant *axis_variants(int *count)
#endif
{
    *count = (int)(sizeof(kVariants) / sizeof(kVariants[0]));
    return kVariants;
}

}  // namespace tope
