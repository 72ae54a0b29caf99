// Line transforms of N <= 512 points, 16 lines per workgroup: the passes of the
// 3-D transforms (the reference's strided 3-D kernels, block_fft_kernel_
// C2C_INCREASE / C2C_NONE_XYZ / C2C_DECREASE / C2R_DECREASE with the XZ/XYZ
// permutes, src/fastfft/FastFFT.cu and include/detail/memory_addressing.h:25-104).
//
// A workgroup of 8 wavefronts takes kTileLines lines (tile modes), each wave
// transforming two of them in registers (one P = N point Stockham transform, no
// sub-transform split), so the twiddle set-up is shared by 16 lines; with
// neither TILE_IN nor TILE_OUT (direct lines) it takes 8 lines, one per wave,
// so that a wave's chain of loads, transform and stores is not run twice.
//   TILE_IN:  the lines lie next to each other and their elements far apart (a
//             z-pass): element rows of the tile (16 consecutive values, one
//             128-byte access) are loaded into LDS, all of a thread's loads
//             issued before its LDS writes; else each wave reads its line
//             directly (unit-stride lines: coalesced).
//   TILE_OUT: likewise for the output: the kept window of each line goes back
//             into its LDS row and the workgroup stores window rows; else each
//             wave stores its own line directly.
//   HERM:     the input line holds X[0 .. N/2] of a Hermitian sequence (X[N-k] =
//             conj X[k] above) and only the real part is stored (C2R).
//   forward (DIR = -1): X[k] = sum_{n < L} x[n] W_N^{n k}, k < M (window origin 0)
//   inverse (DIR = +1): y[j] = sum_k X[k] W_N^{-k (o + j)}, j < M
// A line's LDS row also serves as the exchange buffer of its transform once
// the line's values are in registers.
#pragma once

#include <type_traits>

#include "rows_wave.hpp"

namespace fastfft_amd {

constexpr int kTileLines = 16;

// Lines per workgroup: kTileLines in the tile modes, one per wave for direct lines.
template <bool TILE_IN, bool TILE_OUT>
constexpr int tile_workgroup_lines = TILE_IN || TILE_OUT ? kTileLines : 8;

// Grid: nlines / tile_workgroup_lines workgroups of 512 threads.  TILE_IN / TILE_OUT need
// lines l .. l + 15 of a tile at base(l) + 0 .. 15 (unit stride between lines).
// flags bit 0: tiles in XCD-contiguous ranges (xcd_row; the grid a multiple of 8 for it
// to apply); bit 1: tile rows moved as 16-byte pairs of neighbouring lines (FP32
// complex; the host checks even offsets and 16-byte aligned buffers).
template <int N, int DIR, typename TIn, bool HERM, typename TOut, bool TILE_IN, bool TILE_OUT>
__global__ void __launch_bounds__(512)
        k_tile_lines(const TIn* __restrict__ in, LineMap im, int L, TOut* __restrict__ out, LineMap om, int M, int o,
                     int nlines, float scale, const float2* __restrict__ tw, int flags) {
    constexpr int  T = 64, E = N / T, LT = tile_workgroup_lines<TILE_IN, TILE_OUT>, NW = 8, PITCH = N + 2, H = N / 2;
    constexpr int  RPT = N / (512 / LT);
    constexpr bool REAL_OUT = ! std::is_same<TOut, float2>::value;
    // real input lines (the R2C x-pass of a 3-D forward, as a complex transform of x + 0 i)
    constexpr bool REAL_IN = std::is_same<TIn, float>::value || std::is_same<TIn, __half>::value;
    static_assert(! REAL_IN || (! TILE_IN && ! HERM && DIR < 0), "real input lines are read directly, forward");
    using FFT               = WaveFFT<N>;
    static_assert(FFT::SLOTS <= PITCH, "a tile row must hold the exchange buffer of its transform");
    static_assert(! (REAL_OUT && TILE_OUT), "real outputs are stored by row");
    __shared__ float2 s_tile[LT * PITCH];
    __shared__ float2 s_twp[N];
    const int     w     = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    const int     t     = threadIdx.x % T;
    const long    line0 = (long)((flags & 1) ? xcd_row(blockIdx.x, gridDim.x) : (int)blockIdx.x) * LT;
    const float2* twm   = multi_twiddles(tw);
    // values stored per line: N/2 + 1 of a Hermitian line, the non-zero inputs of a forward, else N
    const int kin = HERM ? H + 1 : (DIR < 0 ? (L < N ? L : N) : N);
    // lines read directly: line l of the workgroup into registers
    auto load_direct = [&](int l, float2(&v)[E]) {
        const auto* lp = as_global(in) + im.base(line0 + l);
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int n = t + T * i;
            if constexpr ( HERM ) {
                const bool   lo = n <= H;
                const float2 z  = load_cpx(lp, (long)(lo ? n : N - n) * im.es);
                v[i]            = lo ? z : make_float2(z.x, -z.y);
            }
            else if constexpr ( REAL_IN ) {
                v[i] = make_float2(n < kin ? load_real(lp, (long)n * im.es) : 0.f, 0.f);
            }
            else {
                v[i] = n < kin ? load_cpx(lp, (long)n * im.es) : make_float2(0.f, 0.f);
            }
        }
    };
    // the wave's first direct line is in flight while the pass table is set up, and each
    // next one while the current one is transformed
    float2 nxt[TILE_IN ? 1 : E];
    if constexpr ( ! TILE_IN )
        load_direct(w, nxt);
    FFT::template fill_pass_table<512>(s_twp, twm, threadIdx.x);
    if constexpr ( TILE_IN ) {
        const auto* ip = as_global(in) + im.base(line0);
        bool wide = false;
        if constexpr ( std::is_same<TIn, float2>::value ) {
            wide = (flags & 2) != 0;
            if ( wide ) {
                // two neighbouring lines per thread: 16-byte loads
                const int     l = threadIdx.x % (LT / 2), r = threadIdx.x / (LT / 2); // r < 64
                constexpr int R2 = 512 / (LT / 2);
                float4        ld[N / R2 + 1];
#pragma unroll
                for ( int i = 0; i < N / R2 + 1; i++ ) {
                    const int k = r + R2 * i;
                    ld[i] = k < kin ? load_cpx_pair(ip, ((long)k * im.es >> 1) + l) : make_float4(0.f, 0.f, 0.f, 0.f);
                }
#pragma unroll
                for ( int i = 0; i < N / R2 + 1; i++ ) {
                    const int k = r + R2 * i;
                    if ( k < kin ) {
                        s_tile[2 * l * PITCH + k]       = make_float2(ld[i].x, ld[i].y);
                        s_tile[(2 * l + 1) * PITCH + k] = make_float2(ld[i].z, ld[i].w);
                    }
                }
            }
        }
        if ( ! wide ) {
            const int l = threadIdx.x % LT, r = threadIdx.x / LT; // r < 32
            float2    ld[RPT + 1];                                // RPT + 1 rows: N / 2 + 1 values of a Hermitian line
#pragma unroll
            for ( int i = 0; i < RPT + 1; i++ ) {
                const int k = r + (512 / LT) * i;
                ld[i]       = k < kin ? load_cpx(ip, (long)k * im.es + l) : make_float2(0.f, 0.f);
            }
#pragma unroll
            for ( int i = 0; i < RPT + 1; i++ ) {
                const int k = r + (512 / LT) * i;
                if ( k < kin )
                    s_tile[l * PITCH + k] = ld[i];
            }
        }
    }
    __syncthreads( );
    FFT fft;
    fft.init_pt(t, s_twp);
    const int oN = ((o % N) + N) % N;
#pragma unroll 1
    for ( int h = 0; h < LT / NW; h++ ) {
        const int l   = w + NW * h;
        float2*   row = s_tile + l * PITCH;
        float2    v[E];
        if constexpr ( TILE_IN ) {
#pragma unroll
            for ( int i = 0; i < E; i++ ) {
                const int n = t + T * i;
                if constexpr ( HERM ) {
                    const bool   lo = n <= H;
                    const float2 z  = row[lo ? n : N - n];
                    v[i]            = lo ? z : make_float2(z.x, -z.y);
                }
                else {
                    v[i] = n < kin ? row[n] : make_float2(0.f, 0.f);
                }
            }
            wave_lds_fence( ); // every lane has read the row before it becomes the exchange buffer
        }
        else {
#pragma unroll
            for ( int i = 0; i < E; i++ )
                v[i] = nxt[i];
            if ( h + 1 < LT / NW )
                load_direct(l + NW, nxt);
        }
        fft.template run<DIR>(v, row, t);
        // value n of the transform is output j = (n - o) mod N, kept when j < M
        if constexpr ( TILE_OUT ) {
#pragma unroll
            for ( int i = 0; i < E; i++ ) {
                const int j = (t + T * i - oN + N) & (N - 1);
                if ( j < M )
                    row[j] = make_float2(v[i].x * scale, v[i].y * scale);
            }
        }
        else {
            auto* op = as_global(out) + om.base(line0 + l);
#pragma unroll
            for ( int i = 0; i < E; i++ ) {
                const int j = (t + T * i - oN + N) & (N - 1);
                if ( j < M ) {
                    if constexpr ( REAL_OUT )
                        store_real(op, (long)j * om.es, v[i].x * scale);
                    else
                        store_cpx(op, (long)j * om.es, make_float2(v[i].x * scale, v[i].y * scale));
                }
            }
        }
    }
    if constexpr ( TILE_OUT ) {
        __syncthreads( );
        auto* op = as_global(out) + om.base(line0);
        if ( flags & 2 ) {
            const int l = threadIdx.x % (LT / 2), r = threadIdx.x / (LT / 2);
            for ( int j = r; j < M; j += 512 / (LT / 2) ) {
                const float2 a = s_tile[2 * l * PITCH + j], b = s_tile[(2 * l + 1) * PITCH + j];
                store_cpx_pair(op, ((long)j * om.es >> 1) + l, a, b);
            }
        }
        else {
            const int l = threadIdx.x % LT, r = threadIdx.x / LT;
            for ( int j = r; j < M; j += 512 / LT )
                store_cpx(op, (long)j * om.es + l, s_tile[l * PITCH + j]);
        }
    }
}

} // namespace fastfft_amd
