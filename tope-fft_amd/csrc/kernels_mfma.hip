// Instantiations of the MFMA row kernel and the FP64 generation of its
// DFT-16 operand constants.
#include <cstdlib>

#include "internal.hpp"
#include "mfma_kernel.hpp"
#include "mfma_tile.hpp"

namespace tope {

// bf16 round-to-nearest-even of a float's bit pattern
__device__ __forceinline__ unsigned short bf16_rne(float f)
{
    const unsigned int u = __builtin_bit_cast(unsigned int, f);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ double bf16_val(unsigned short h)
{
    return (double)__builtin_bit_cast(float, (unsigned int)h << 16);
}

// One thread per (dir, re/im, lane, slot): the DFT-16 matrix entry for that
// operand slot, computed in FP64 (sincospi of an exact multiple of 1/8),
// then split into three bf16 parts and rounded once to FP32.
__global__ void gen_mfma_consts(MfmaConsts *K)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;  // 2 * 2 * 64 * 8
    if (t >= 2 * 2 * 64 * 8) return;
    const int slot = t & 7, lane = (t >> 3) & 63, ri = (t >> 9) & 1, dir = t >> 10;
    const int i = lane & 15, g = lane >> 4;
    const double sign = dir == 0 ? -1.0 : 1.0;
    // bf16 form: slot < 4 -> Re x_{4g+slot}; slot >= 4 -> Im x_{4g+slot-4}
    {
        const int m = 4 * g + (slot & 3);
        double s, c;
        sincospi(2.0 * (double)((i * m) & 15) / 16.0, &s, &c);
        const double wr = c, wi = sign * s;
        double a;
        if (ri == 0) a = slot < 4 ? wr : -wi;  // Yr = Re W Re x - Im W Im x
        else a = slot < 4 ? wi : wr;           // Yi = Im W Re x + Re W Im x
        double rem = a;
        for (int p = 0; p < 3; ++p) {
            const unsigned short h = bf16_rne((float)rem);
            rem -= bf16_val(h);
            unsigned short *dst = reinterpret_cast<unsigned short *>(&K->bf[dir][ri][p][lane][0]);
            dst[slot] = h;
        }
    }
    // f16 form: the bf16 slot layout, two fp16 parts (round to nearest)
    {
        const int m = 4 * g + (slot & 3);
        double s, c;
        sincospi(2.0 * (double)((i * m) & 15) / 16.0, &s, &c);
        const double wr = c, wi = sign * s;
        const double a = ri == 0 ? (slot < 4 ? wr : -wi) : (slot < 4 ? wi : wr);
        const _Float16 h0 = (_Float16)(float)a;
        const _Float16 h1 = (_Float16)(float)(a - (double)(float)h0);
        _Float16 *d0 = reinterpret_cast<_Float16 *>(&K->h[dir][ri][0][lane][0]);
        _Float16 *d1 = reinterpret_cast<_Float16 *>(&K->h[dir][ri][1][lane][0]);
        d0[slot] = h0;
        d1[slot] = h1;
    }
    // f32 form: instruction kk = slot, lane group g supplies K = 4kk + g
    {
        const int kk = slot;
        const int m = 4 * g + (kk & 3);
        double s, c;
        sincospi(2.0 * (double)((i * m) & 15) / 16.0, &s, &c);
        const double wr = c, wi = sign * s;
        double a;
        if (ri == 0) a = kk < 4 ? wr : -wi;
        else a = kk < 4 ? wi : wr;
        K->f32[dir][ri][kk][lane] = (float)a;
    }
}

hipError_t gen_mfma_constants(void *out, hipStream_t s)
{
    hipLaunchKernelGGL(gen_mfma_consts, dim3(8), dim3(256), 0, s, (MfmaConsts *)out);
    return hipGetLastError();
}

size_t mfma_constants_bytes() { return sizeof(MfmaConsts); }

template <int MODE, int N, int W>
static hipError_t launch_nw(const void *in, void *out, const MfmaArgs &a, int sign, hipStream_t s)
{
    using SH = MfmaShape<N, W>;
    const long long blocks = (a.count + SH::TPB - 1) / SH::TPB;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    const cpx<float> *i = static_cast<const cpx<float> *>(in);
    cpx<float> *o = static_cast<cpx<float> *>(out);
    if (sign < 0)
        hipLaunchKernelGGL((mfma_rows_kernel<MODE, N, -1, W>), dim3((unsigned)blocks), dim3(SH::THREADS), 0, s, i,
                           o, a);
    else
        hipLaunchKernelGGL((mfma_rows_kernel<MODE, N, 1, W>), dim3((unsigned)blocks), dim3(SH::THREADS), 0, s, i,
                           o, a);
    return hipGetLastError();
}

// waves per transform: 4 (4 groups per lane, the measured default) or 8
// (2 groups per lane), TOPEFFT_MFMA_WAVES=8 selects the latter
static int mfma_waves()
{
    static const int w = [] {
        const char *e = std::getenv("TOPEFFT_MFMA_WAVES");
        return e && std::atoi(e) == 8 ? 8 : 4;
    }();
    return w;
}

template <int MODE, int N>
static hipError_t launch_n(const void *in, void *out, const MfmaArgs &a, int sign, hipStream_t s)
{
    if (mfma_waves() == 4) return launch_nw<MODE, N, 4>(in, out, a, sign, s);
    return launch_nw<MODE, N, 8>(in, out, a, sign, s);
}

template <int MODE, int N>
static hipError_t launch_rem(const void *in, void *out, const MfmaArgs &a, int sign, hipStream_t s)
{
    using SH = MfmaRemShape<N>;
    const long long blocks = (a.count + SH::TPB - 1) / SH::TPB;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    const cpx<float> *i = static_cast<const cpx<float> *>(in);
    cpx<float> *o = static_cast<cpx<float> *>(out);
    if (sign < 0)
        hipLaunchKernelGGL((mfma_rows_rem_kernel<MODE, N, -1>), dim3((unsigned)blocks), dim3(SH::THREADS), 0, s, i, o,
                           a);
    else
        hipLaunchKernelGGL((mfma_rows_rem_kernel<MODE, N, 1>), dim3((unsigned)blocks), dim3(SH::THREADS), 0, s, i, o,
                           a);
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_mode(int n, const void *in, void *out, const MfmaArgs &a, int sign, hipStream_t s)
{
    switch (n) {
    case 256: return launch_n<MODE, 256>(in, out, a, sign, s);
    case 512: return launch_rem<MODE, 512>(in, out, a, sign, s);
    case 1024: return launch_rem<MODE, 1024>(in, out, a, sign, s);
    case 2048: return launch_rem<MODE, 2048>(in, out, a, sign, s);
    case 4096: return launch_n<MODE, 4096>(in, out, a, sign, s);
    default: return hipErrorInvalidValue;
    }
}

bool mfma_supported(int n) { return n == 256 || n == 512 || n == 1024 || n == 2048 || n == 4096; }

template <int MODE, int N>
static hipError_t launch_tile_n(const void *in, void *out, const MfmaTileArgs &a, unsigned grid, int sign,
                                hipStream_t s)
{
    const cpx<float> *i = static_cast<const cpx<float> *>(in);
    cpx<float> *o = static_cast<cpx<float> *>(out);
    constexpr int T = MfmaTile<N>::THREADS;
    if (sign < 0) hipLaunchKernelGGL((mfma_tile_kernel<MODE, N, -1>), dim3(grid), dim3(T), 0, s, i, o, a);
    else hipLaunchKernelGGL((mfma_tile_kernel<MODE, N, 1>), dim3(grid), dim3(T), 0, s, i, o, a);
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_tile_mode(int n, const void *in, void *out, const MfmaTileArgs &a, unsigned grid, int sign,
                                   hipStream_t s)
{
    switch (n) {
    case 256: return launch_tile_n<MODE, 256>(in, out, a, grid, sign, s);
    case 512: return launch_tile_n<MODE, 512>(in, out, a, grid, sign, s);
    case 1024: return launch_tile_n<MODE, 1024>(in, out, a, grid, sign, s);
    default: return hipErrorInvalidValue;
    }
}

bool mfma_tile_supported(int n) { return n == 256 || n == 512 || n == 1024; }

hipError_t launch_mfma_tile(int mode, int n, const void *in, void *out, const MfmaTileArgs &a, unsigned grid, int sign,
                            hipStream_t s)
{
    switch (mode) {
    case MP_BF16X3: return launch_tile_mode<MP_BF16X3>(n, in, out, a, grid, sign, s);
    case MP_BF16X6: return launch_tile_mode<MP_BF16X6>(n, in, out, a, grid, sign, s);
    case MP_FP32: return launch_tile_mode<MP_FP32>(n, in, out, a, grid, sign, s);
    case MP_FP16X3: return launch_tile_mode<MP_FP16X3>(n, in, out, a, grid, sign, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_mfma_rows(int mode, int n, const void *in, void *out, const MfmaArgs &a, int sign, hipStream_t s)
{
    switch (mode) {
    case MP_BF16X3: return launch_mode<MP_BF16X3>(n, in, out, a, sign, s);
    case MP_BF16X6: return launch_mode<MP_BF16X6>(n, in, out, a, sign, s);
    case MP_FP32: return launch_mode<MP_FP32>(n, in, out, a, sign, s);
    case MP_FP16X3: return launch_mode<MP_FP16X3>(n, in, out, a, sign, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace tope
