// Complex line transforms of N <= 512 points whose lines lie next to each other
// in memory and whose elements are far apart (the z-passes of the 3-D
// transforms: C[kz][kx][ky] lines along kz), the reference's strided 3-D
// kernels (block_fft_kernel_C2C_INCREASE / C2C_NONE_XYZ / C2C_DECREASE with the
// XZ/XYZ permutes, src/fastfft/FastFFT.cu and include/detail/memory_addressing.h:25-104).
//
// A workgroup of 8 wavefronts takes kTileLines adjacent lines: it loads element
// rows of the tile (kTileLines consecutive values, one 128-byte access per row;
// all of a thread's loads are issued before its LDS writes) into LDS, each wave
// transforms two lines from LDS in registers (one P = N point Stockham
// transform, so no sub-transform split; the line's own LDS row is its exchange
// buffer), writes the kept window of each line back into that row, and the
// workgroup stores the window rows the same coalesced way.
//   forward (DIR = -1): X[k] = sum_{n < L} x[n] W_N^{n k}, k < M (window origin 0)
//   inverse (DIR = +1): y[j] = sum_k X[k] W_N^{-k (o + j)}, j < M
#pragma once

#include "rows_wave.hpp"

namespace fastfft_amd {

constexpr int kTileLines = 16;

// Grid: nlines / kTileLines workgroups of 512 threads.  im.s1 = om.s1 = 1 (line
// l + 1 follows line l), element strides im.es and om.es.
template <int N, int DIR, typename TIn>
__global__ void __launch_bounds__(512)
        k_tile_lines(const TIn* __restrict__ in, LineMap im, int L, float2* __restrict__ out, LineMap om, int M, int o,
                     int nlines, float scale, const float2* __restrict__ tw) {
    constexpr int T = 64, E = N / T, LT = kTileLines, NW = 8, PITCH = N + 2, RPT = N / (512 / LT);
    using FFT       = WaveFFT<N>;
    static_assert(FFT::SLOTS <= PITCH, "a tile row must hold the exchange buffer of its transform");
    __shared__ float2 s_tile[LT * PITCH];
    __shared__ float2 s_twp[N];
    const int     w     = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    const int     t     = threadIdx.x % T;
    const long    line0 = (long)blockIdx.x * LT;
    const float2* twm   = multi_twiddles(tw);
    FFT::fill_pass_table(s_twp, twm, threadIdx.x, 512);
    // ---- element rows of the tile: thread (r, l) loads element r + 32 i of line l
    const int  kin = DIR < 0 ? (L < N ? L : N) : N;
    const auto* ip = as_global(in) + im.base(line0);
    {
        const int l = threadIdx.x % LT, r = threadIdx.x / LT; // r < 32
        float2    ld[RPT];
#pragma unroll
        for ( int i = 0; i < RPT; i++ ) {
            const int k = r + (512 / LT) * i;
            ld[i]       = k < kin ? load_cpx(ip, (long)k * im.es + l) : make_float2(0.f, 0.f);
        }
#pragma unroll
        for ( int i = 0; i < RPT; i++ ) {
            const int k = r + (512 / LT) * i;
            if ( k < kin )
                s_tile[l * PITCH + k] = ld[i];
        }
    }
    __syncthreads( );
    FFT fft;
    fft.init_pt(t, s_twp);
    const int oN = ((o % N) + N) % N;
#pragma unroll 1
    for ( int h = 0; h < LT / NW; h++ ) {
        const int l  = w + NW * h;
        float2*   row = s_tile + l * PITCH;
        float2    v[E];
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int n = t + T * i;
            v[i]        = n < kin ? row[n] : make_float2(0.f, 0.f);
        }
        fft.template run<DIR>(v, row, t); // the row's values are in registers: it is free
        // value n of the transform is output j = (n - o) mod N, kept when j < M
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int j = (t + T * i - oN + N) & (N - 1);
            if ( j < M )
                row[j] = make_float2(v[i].x * scale, v[i].y * scale);
        }
    }
    __syncthreads( );
    // ---- window rows of the tile
    auto* op = as_global(out) + om.base(line0);
    {
        const int l = threadIdx.x % LT, r = threadIdx.x / LT;
        for ( int j = r; j < M; j += 512 / LT )
            store_cpx(op, (long)j * om.es + l, s_tile[l * PITCH + j]);
    }
}

} // namespace fastfft_amd
