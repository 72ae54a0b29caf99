// Output-pruned inverse line transforms with one workgroup of Q wavefronts per
// line, for windows the tuned C2R row kernel does not take (any origin, and up
// to M = N outputs): the inverse y-pass of a trimmed InvFFT and its C2R x-pass
// (the reference's block_fft_kernel_C2C_DECREASE inverse use and
// C2R_DECREASE_XY, src/fastfft/FastFFT.cu:1405-1431, :1662-1692).
//
// With N = P*Q and U_q = IFFT_P( X[Q k' + q] ) (wave q, register Stockham),
//   y[n] = sum_q W_N^{-n q} U_q[n mod P],   n = o + j, j < M,
// evaluated per output by Horner's rule in w = W_N^{-n}:
//   y[n] = U_0 + w (U_1 + w (U_2 + ... + w U_{Q-1})).
// U_q is kept in LDS in natural order, so an output reads it at any rotation
// (consecutive outputs read consecutive words: no bank conflicts), and the
// twiddle W_N^{o + j} of output j comes from the table.  The workgroup's LDS is
// the exchange buffers and the pass table only (36 KB at N = 4096: four
// workgroups per CU).
// HERM: the line holds X[0 .. N/2] of a Hermitian sequence (X[N-k] = conj X[k]
// above) and only the real part of y is stored (REAL_OUT).
#pragma once

#include "rows_wave.hpp"

namespace fastfft_amd {

template <int P, int Q, typename TIn, bool HERM, typename TOut, bool REAL_OUT>
__global__ void __launch_bounds__(64 * Q)
        k_inv_window_rows(const TIn* __restrict__ in, LineMap im, TOut* __restrict__ out, LineMap om, int M, int o,
                          int nlines, float post_scale, const float2* __restrict__ tw) {
    constexpr int T = 64, E = P / T, N = P * Q, NT = T * Q, H = N / 2;
    using FFT       = WaveFFT<P>;
    static_assert(Q * FFT::SLOTS >= N, "a line must fit the exchange buffers");
    __shared__ __align__(16) float2 lds[Q * FFT::SLOTS];
    __shared__ float2 s_twp[P];
    const int     q    = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    const int     t    = threadIdx.x % T;
    const int     line = xcd_row(blockIdx.x, nlines);
    const float2* twm  = multi_twiddles(tw);
    // ---- the input line into LDS, decimated by Q, two values per lane, every
    // load of the thread in flight before its LDS writes
    {
        float2 ptv[FFT::template pass_table_rows<NT>];
        FFT::template load_pass_table<NT>(ptv, twm, threadIdx.x);
        const auto*   lp = as_global(in) + im.base(line);
        constexpr int KN = HERM ? H + 1 : N;
        constexpr int MR = (KN + 2 * NT - 1) / (2 * NT);
        float2        u[MR][2];
        // unit-stride FP32 lines on 16-byte boundaries: each lane's pair in one load
        bool pairs = false;
        if constexpr ( std::is_same_v<TIn, float2> )
            pairs = im.es == 1 && (im.base(line) & 1) == 0 && (reinterpret_cast<unsigned long>(in) & 15) == 0;
        if ( pairs ) {
            if constexpr ( std::is_same_v<TIn, float2> ) {
#pragma unroll
                for ( int r = 0; r < MR; r++ ) {
                    const int m = threadIdx.x + NT * r;
                    if ( 2 * m + 1 < KN ) {
                        const float4 a = load_cpx_pair(lp, m);
                        u[r][0]        = make_float2(a.x, a.y);
                        u[r][1]        = make_float2(a.z, a.w);
                    }
                    else {
                        u[r][0] = 2 * m < KN ? load_cpx(lp, (long)(2 * m)) : make_float2(0.f, 0.f);
                        u[r][1] = make_float2(0.f, 0.f);
                    }
                }
            }
        }
        else {
#pragma unroll
            for ( int r = 0; r < MR; r++ ) {
                const int m = threadIdx.x + NT * r;
                u[r][0]     = 2 * m < KN ? load_cpx(lp, (long)(2 * m) * im.es) : make_float2(0.f, 0.f);
                u[r][1]     = 2 * m + 1 < KN ? load_cpx(lp, (long)(2 * m + 1) * im.es) : make_float2(0.f, 0.f);
            }
        }
        FFT::template store_pass_table<NT>(s_twp, ptv, threadIdx.x);
#pragma unroll
        for ( int r = 0; r < MR; r++ ) {
            const int m = threadIdx.x + NT * r;
            if ( 2 * m < KN )
                lds[dec_slot<P, Q>(2 * m)] = u[r][0];
            if ( 2 * m + 1 < KN )
                lds[dec_slot<P, Q>(2 * m + 1)] = u[r][1];
        }
    }
    __syncthreads( );
    FFT fft;
    fft.init_pt(t, s_twp);
    float2 v[E];
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const int k = Q * (t + T * i) + q;
        if constexpr ( HERM ) {
            const bool   lo = k <= H;
            const float2 z  = lds[dec_slot<P, Q>(lo ? k : N - k)];
            v[i]            = lo ? z : make_float2(z.x, -z.y);
        }
        else {
            v[i] = lds[dec_slot<P, Q>(k)];
        }
    }
    __syncthreads( );
    float2* wl = lds + q * FFT::SLOTS;
    fft.template run<+1>(v, wl, t);
    // U_q[m], m = t + 64 i, in natural order in this wave's own region
#pragma unroll
    for ( int i = 0; i < E; i++ )
        wl[t + T * i] = v[i];
    __syncthreads( );
    // ---- the window: Horner over q for each output
    const int oN = ((o % N) + N) % N;
    const int oP = oN & (P - 1);
    TOut*     op = out + om.base(line);
    // complex unit-stride outputs from an even origin on 16-byte boundaries: two outputs per
    // thread, their twiddles, shares and stores as 16-byte accesses
    if constexpr ( ! REAL_OUT ) {
        if ( om.es == 1 && (oN & 1) == 0 && (M & 1) == 0 && (om.base(line) & 1) == 0 &&
             (reinterpret_cast<unsigned long>(out) & 15) == 0 ) {
            const auto* lds4 = reinterpret_cast<const float4*>(lds);
            for ( int r = 0; 2 * r * NT < M; r++ ) {
                const int j = 2 * ((int)threadIdx.x + NT * r);
                if ( j >= M )
                    break;
                const float4 w2 = load_cpx_pair(as_global(twm) + N, ((oN + j) & (N - 1)) >> 1); // W_N^{o + j}, W_N^{o + j + 1}
                const int    m  = (oP + j) & (P - 1);
                float4       a  = lds4[((Q - 1) * FFT::SLOTS + m) >> 1];
#pragma unroll
                for ( int qq = Q - 2; qq >= 0; qq-- ) {
                    const float4 u  = lds4[(qq * FFT::SLOTS + m) >> 1];
                    const float2 a0 = cmulc(make_float2(a.x, a.y), make_float2(w2.x, w2.y));
                    const float2 a1 = cmulc(make_float2(a.z, a.w), make_float2(w2.z, w2.w));
                    a               = make_float4(a0.x + u.x, a0.y + u.y, a1.x + u.z, a1.y + u.w);
                }
                store_cpx_pair(as_global(op), j >> 1, make_float2(a.x * post_scale, a.y * post_scale),
                               make_float2(a.z * post_scale, a.w * post_scale));
            }
            return;
        }
    }
    for ( int r = 0; r * NT < M; r++ ) {
        const int j = (int)threadIdx.x + NT * r;
        if ( j >= M )
            break;
        const float2 w  = load_cpx(as_global(twm), N + ((oN + j) & (N - 1))); // W_N^{o + j}
        const int    m  = (oP + j) & (P - 1);
        float2       acc = lds[(Q - 1) * FFT::SLOTS + m];
#pragma unroll
        for ( int qq = Q - 2; qq >= 0; qq-- ) {
            const float2 u = lds[qq * FFT::SLOTS + m];
            acc            = cmulc(acc, w); // acc * conj(w) = acc * W_N^{-(o + j)}
            acc.x += u.x;
            acc.y += u.y;
        }
        if constexpr ( REAL_OUT )
            store_real(op, (long)j * om.es, acc.x * post_scale);
        else
            store_cpx(op, (long)j * om.es, make_float2(acc.x * post_scale, acc.y * post_scale));
    }
}

} // namespace fastfft_amd
