// Mixed-precision transforms of strided axes (column tiles) on the matrix
// cores.  No reference counterpart: the reference runs its Y/Z axes as FP64
// radix-2/4/8 passes with 16-byte lanes along the stride (src/kernels2D.cl
// :290-456 type 2, src/kernels3D.cl:300-474 types 2 and 3); the Stockham data
// flow, twiddle table layout (Sched<N, 16>) and sign convention are those of
// the FP64 axis kernel, and each radix-16 stage is the DFT-matrix product of
// the row kernel (mfma_kernel.hpp) with a different lane assignment.
//
// A block transforms 16 adjacent lines (columns, unit line stride) of one
// length-N axis, N = R 16^s (R = 1, 2, 4; N <= 1024).  In every MFMA
// instruction the 16 B-operand columns are the 16 lines at the same
// butterfly j, so lane (c, g) holds inputs m = 4g..4g+3 of butterfly j of
// line c: a load or store instruction of a wave touches 16 adjacent lines at
// four rows -- four 128-byte segments -- with no staging.  The twiddles of a
// butterfly are the same for all 16 lines.
//
//   stage 0 (R > 1): radix-R butterflies on the VALU, 16/R per thread
//   MFMA stages: groups of 16 lines x one butterfly, 4 groups per wave
//   between stages: the tile in LDS, one component at a time (32 KiB for
//   N = 512), element i of line c at word tile_lds(i, c)
//   last stage: outputs stored straight to HBM
//
// MP_FP16X3 scales every line by its own power of two (the row kernel's
// rule, with the headroom of this N: the last stage's input stays below
// 2^15).
#pragma once
#include "mfma_kernel.hpp"

namespace tope {

// Word of element i of tile line c: 16 lines interleaved per element, with
// bits 4, 5 (the 16-word window within 64 banks) XORed with two parities of
// i's bits above bit 1, so that every access of the kernel -- lanes over 16
// lines and four rows whose indices differ in one 2-bit field of i -- hits
// 64 distinct banks (and each half wave 32).  Linear over GF(2) in (i, c).
__device__ __forceinline__ constexpr int tile_lds(int i, int c)
{
    const int k0 = __builtin_popcount(i >> 2) & 1;
    const int k1 = __builtin_popcount((i >> 3) & 0x55555555) & 1;
    return ((i << 4) | c) ^ ((k0 | (k1 << 1)) << 4);
}
// Every element index the kernel forms is lane + k with lane and k in
// disjoint bits (lane: wave, g; k: compile-time q, r, s, u, m), so its LDS
// word is tile_lds(lane, c) ^ tile_lds(k, 0) and its byte offset in the
// line is (c + lane is) 8 (per lane) + k is 8 (uniform).

struct MfmaTileArgs {
    long long O, os;       // outer index o < O at element offset o * os
    long long tiles;       // 16-line tiles per o (lines c at offset c, unit stride)
    long long is;          // element stride along the axis
    const cpx<float> *tw;  // stage twiddles, layout of Sched<N,16> (FP64-generated, rounded)
    const MfmaConsts *k;   // DFT-16 operand constants
    float scale;           // output scale (1/N_total on the inverse's last pass, else 1)
    long long in_elems, out_elems;
    int xcd;
};

template <int N>
struct MfmaTile {
    using SC = Sched<N, 16>;
    static constexpr int R = SC::REM;          // leading radix (1: none)
    static constexpr int S0 = R > 1 ? 1 : 0;   // first MFMA stage of the schedule
    static constexpr int NB = N / 16;          // butterflies per line per MFMA stage
    static constexpr int GPW = 4;              // groups per wave per MFMA stage
    static constexpr int W = NB / GPW;         // waves per block
    static constexpr int THREADS = 64 * W;
    static_assert(R <= 4 && N >= 256 && N <= 1024, "N = R 16^s, R <= 4, 256 <= N <= 1024");
};

template <int MODE, int N, int SIGN>
__global__ void __launch_bounds__(MfmaTile<N>::THREADS) mfma_tile_kernel(const cpx<float> *in, cpx<float> *out,
                                                                          const MfmaTileArgs a)
{
    using T = MfmaTile<N>;
    using SC = typename T::SC;
    constexpr int R = T::R, NB = T::NB, GPW = T::GPW, W = T::W;
    constexpr int CB = 8;
    __shared__ float lds[16 * N];
    __shared__ float line_max[W][16];

    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int c = lane & 15, g = lane >> 4;
    const int dir = SIGN < 0 ? 0 : 1;
    const long long blk = xcd_block(a.xcd);
    const long long o = blk / a.tiles;
    const long long first = o * a.os + (blk % a.tiles) * 16;  // element offset of line 0 of the tile
    // byte offsets fit 31 bits (checked on the host): per-lane part + uniform part
    const int is8 = (int)(a.is * CB);
    auto voff = [&](int lane_i) { return (c + lane_i * (int)a.is) * CB; };

    // lane parts of the element indices
    const int li_mfma = wave * GPW + NB * 4 * g;  // MFMA stage inputs and the final outputs
    const int lw_mfma = tile_lds(li_mfma, c) * 4;  // LDS byte offsets (XOR distributes over * 4)
    const int v_mfma = voff(li_mfma);
    auto word = [&](int byte_off) -> float & { return *reinterpret_cast<float *>(reinterpret_cast<char *>(lds) + byte_off); };

    float x[GPW][8];  // slots 0..3: real parts, 4..7: imaginary parts
    {
        const __amdgpu_buffer_rsrc_t rin = make_rsrc(in + first, (a.in_elems - first) * CB);
        if constexpr (R > 1) {
            // stage-0 butterflies j = wave + W g + NB u (u < 16/R), inputs j + (N/R) m;
            // element e = u R + m sits in x[e / 4][e % 4] (re) and x[e / 4][4 + e % 4] (im)
            const int v0 = voff(wave + W * g);
#pragma unroll
            for (int u = 0; u < 16 / R; ++u)
#pragma unroll
                for (int m = 0; m < R; ++m) {
                    const int e = u * R + m;
                    const cpx<float> v = buf_load<float>(rin, v0, (NB * u + (N / R) * m) * is8);
                    x[e / 4][e % 4] = v.x;
                    x[e / 4][4 + e % 4] = v.y;
                }
        } else {
#pragma unroll
            for (int q = 0; q < GPW; ++q)
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const cpx<float> v = buf_load<float>(rin, v_mfma, (q + NB * s) * is8);
                    x[q][s] = v.x;
                    x[q][4 + s] = v.y;
                }
        }
    }

    float out_scale = a.scale;
    if constexpr (MODE == MP_FP16X3) {
        // every value a lane holds belongs to its line c
        float m = 0.0f;
#pragma unroll
        for (int q = 0; q < GPW; ++q)
#pragma unroll
            for (int s = 0; s < 8; ++s) m = fmaxf(m, fabsf(x[q][s]));
        m = fmaxf(m, __shfl_xor(m, 16));
        m = fmaxf(m, __shfl_xor(m, 32));
        if (g == 0) line_max[wave][c] = m;
        __syncthreads();
#pragma unroll
        for (int w = 0; w < W; ++w) m = fmaxf(m, line_max[w][c]);
        // the last MFMA stage's input is at most N/16 times the largest
        // input: scale the line so that it lies below 2^(15 - log2(N/16))
        constexpr int TOP = 15 - ilog2c(N / 16);
        int e = __builtin_amdgcn_frexp_expf(m);
        e = e < TOP - 126 ? TOP - 126 : (e > TOP + 126 ? TOP + 126 : e);
        const float up = __builtin_ldexpf(1.0f, TOP - e);
        out_scale *= __builtin_ldexpf(1.0f, e - TOP);
#pragma unroll
        for (int q = 0; q < GPW; ++q)
#pragma unroll
            for (int s = 0; s < 8; ++s) x[q][s] *= up;
    }

    if constexpr (R > 1) {
        // radix-R stage 0 (span 1: no twiddles); outputs r of butterfly j go to j R + r
#pragma unroll
        for (int u = 0; u < 16 / R; ++u) {
            cpx<float> v[R];
#pragma unroll
            for (int m = 0; m < R; ++m) {
                const int e = u * R + m;
                v[m] = cpx<float>{x[e / 4][e % 4], x[e / 4][4 + e % 4]};
            }
            dft<R, SIGN>(v);
#pragma unroll
            for (int m = 0; m < R; ++m) {
                const int e = u * R + m;
                x[e / 4][e % 4] = v[m].x;
                x[e / 4][4 + e % 4] = v[m].y;
            }
        }
        const int lw0 = tile_lds((wave + W * g) * R, c) * 4;
#pragma unroll
        for (int part = 0; part < 2; ++part) {
            if (part > 0) __syncthreads();
#pragma unroll
            for (int u = 0; u < 16 / R; ++u)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int e = u * R + r;
                    word(lw0 ^ (tile_lds(NB * R * u + r, 0) * 4)) = x[e / 4][4 * part + e % 4];
                }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < GPW; ++q)
#pragma unroll
                for (int s = 0; s < 4; ++s) x[q][4 * part + s] = word(lw_mfma ^ (tile_lds(q + NB * s, 0) * 4));
        }
    }

#pragma unroll
    for (int S = T::S0; S < SC::NST; ++S) {
        const int Ns = SC::span(S);
#pragma unroll
        for (int q = 0; q < GPW; ++q) {
            const int j = wave * GPW + q;
            if (S > 0) {
                // input m = 4g + s of butterfly j: W_{16 Ns}^{(j mod Ns) m}
                const cpx<float> *t = a.tw + SC::twoff(S) + (j & (Ns - 1));
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int m = 4 * g + s;
                    cpx<float> w = t[(m > 0 ? m - 1 : 0) * Ns];
                    if (m == 0) w = cpx<float>{1.0f, 0.0f};
                    const cpx<float> v = twmul<SIGN>(cpx<float>{x[q][s], x[q][4 + s]}, w);
                    x[q][s] = v.x;
                    x[q][4 + s] = v.y;
                }
            }
            float y[8];
            dft16_mfma<MODE>(x[q], y, a.k, dir, lane);
#pragma unroll
            for (int s = 0; s < 8; ++s) x[q][s] = y[s];
        }
        if (S == SC::NST - 1) {
            // outputs r of butterfly j: element j + NB (4g + r) (natural order)
            const __amdgpu_buffer_rsrc_t rout = make_rsrc(out + first, (a.out_elems - first) * CB);
            constexpr bool SCALED = MODE == MP_FP16X3 || SIGN > 0;
#pragma unroll
            for (int q = 0; q < GPW; ++q)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    cpx<float> v{x[q][r], x[q][4 + r]};
                    if constexpr (SCALED) v = scale(v, out_scale);
                    buf_store<float>(v, rout, v_mfma, (q + NB * r) * is8);
                }
        } else {
            // output r of butterfly j = wave GPW + q: (j / Ns) Ns 16 + j % Ns + (4g + r) Ns
            const int lNs = ilog2c(Ns);
            const int jl = wave * GPW;
            const int lw = tile_lds(((jl >> lNs) << lNs) * 16 + (jl & (Ns - 1)) + 4 * g * Ns, c) * 4;
#pragma unroll
            for (int part = 0; part < 2; ++part) {
                __syncthreads();  // the previous reads of lds are done
#pragma unroll
                for (int q = 0; q < GPW; ++q)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        word(lw ^ (tile_lds(((q >> lNs) << lNs) * 16 + (q & (Ns - 1)) + r * Ns, 0) * 4)) =
                            x[q][4 * part + r];
                __syncthreads();
#pragma unroll
                for (int q = 0; q < GPW; ++q)
#pragma unroll
                    for (int s = 0; s < 4; ++s) x[q][4 * part + s] = word(lw_mfma ^ (tile_lds(q + NB * s, 0) * 4));
            }
        }
    }
}

}  // namespace tope
