// Full-length forward line transforms (no size change) with one workgroup of
// Q wavefronts per line: the R2C x-pass and the C2C y-pass of an unpadded
// FwdFFT (the reference's block_fft_kernel_R2C_NONE_XY and C2C_NONE,
// src/fastfft/FastFFT.cu:1055-1077, :1348-1367), used for the micrograph
// spectrum of the correlation path.
//
// N = P*Q is split decimation-in-frequency style:
//   y_q[m]      = sum_s x[s P + m] W_Q^{s q}          (radix-Q pass over the line in LDS)
//   X[Q k' + q] = FFT_P( y_q[m] W_N^{m q} )[k']         (wave q, register Stockham)
// The line is read with coalesced loads into LDS.  A unit-stride output line is
// written through LDS with 16-byte stores; a strided one (the transposed x-pass
// output) one value per output line, with each XCD owning a contiguous range
// of lines so that the partial lines meet in one L2.
#pragma once

#include "rows_wave.hpp"

namespace fastfft_amd {

template <int P, int Q, typename TIn, bool REAL>
__global__ void __launch_bounds__(64 * Q)
        k_fwd_full_wave(const TIn* __restrict__ in, LineMap im, float2* __restrict__ out, LineMap om, int nout, int nlines,
                        float pre_scale, const float2* __restrict__ tw) {
    constexpr int T = 64, E = P / T, N = P * Q, NT = T * Q;
    using FFT       = WaveFFT<P>;
    static_assert(Q * FFT::SLOTS >= N, "a line must fit the exchange buffers");
    static_assert(P % NT == 0 || NT % P == 0, "whole groups of the radix-Q pass per thread");
    __shared__ float2 lds[Q * FFT::SLOTS];
    __shared__ float2 s_twp[P];
    __shared__ float2 s_twn[NT];
    __shared__ float  s_real[REAL ? N : 1];
    const int     q    = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    const int     t    = threadIdx.x % T;
    const int     line = xcd_row(blockIdx.x, nlines);
    const float2* twm  = multi_twiddles(tw);
    // ---- the line into LDS (natural order), every load of the thread in
    // flight before its LDS writes
    {
        RowTwiddleStage<P, Q> tws;
        tws.load(twm);
        const auto* lp = as_global(in) + im.base(line);
        float2      z[N / NT];
#pragma unroll
        for ( int r = 0; r < N / NT; r++ ) {
            const int n = threadIdx.x + NT * r;
            if constexpr ( REAL )
                z[r] = make_float2(load_real(lp, n * im.es), 0.f);
            else
                z[r] = load_cpx(lp, n * im.es);
        }
        tws.store(s_twp, s_twn);
#pragma unroll
        for ( int r = 0; r < N / NT; r++ ) {
            const int n = threadIdx.x + NT * r;
            if constexpr ( REAL )
                s_real[n] = z[r].x * pre_scale;
            else
                lds[n] = make_float2(z[r].x * pre_scale, z[r].y * pre_scale);
        }
    }
    __syncthreads( );
    // ---- radix-Q pass: thread m holds x[s P + m], s < Q, and leaves y_q[m] at q P + m
    for ( int m = threadIdx.x; m < P; m += NT ) {
        float2 u[Q];
#pragma unroll
        for ( int s = 0; s < Q; s++ ) {
            if constexpr ( REAL )
                u[s] = make_float2(s_real[s * P + m], 0.f);
            else
                u[s] = lds[s * P + m];
        }
        SmallDFT<Q, -1>::run(u);
#pragma unroll
        for ( int s = 0; s < Q; s++ )
            lds[s * P + m] = u[s];
    }
    __syncthreads( );
    FFT fft;
    fft.init_pt(t, s_twp);
    const float2 lw = lane_twiddle(s_twn, t, q); // W_N^{t q}
    float2       v[E];
#pragma unroll
    for ( int i = 0; i < E; i++ )
        v[i] = cmul(cmul(lds[q * P + t + T * i], twm[N + ((T * i * q) & (N - 1))]), lw);
    __syncthreads( );
    fft.template run<-1>(v, lds + q * FFT::SLOTS, t);
    float2* op = out + om.base(line);
    if ( om.es == 1 ) {
        // natural order through LDS, then 16-byte stores (op is 16-byte aligned)
        __syncthreads( );
#pragma unroll
        for ( int i = 0; i < E; i++ )
            lds[dec_slot<P, Q>(Q * (t + T * i) + q)] = v[i];
        __syncthreads( );
        for ( int m = threadIdx.x; 2 * m < nout; m += NT ) {
            const float2 u0 = lds[dec_slot<P, Q>(2 * m)];
            if ( 2 * m + 1 < nout ) {
                const float2 u1                  = lds[dec_slot<P, Q>(2 * m + 1)];
                reinterpret_cast<float4*>(op)[m] = make_float4(u0.x, u0.y, u1.x, u1.y);
            }
            else {
                op[2 * m] = u0;
            }
        }
    }
    else {
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int k = Q * (t + T * i) + q;
            if ( k < nout )
                op[k * om.es] = v[i];
        }
    }
}

// Two real lines as one complex line z = a + i b (the R2C x-pass of a FwdFFT
// without size change, the reference's block_fft_kernel_R2C_NONE_XY,
// FastFFT.cu:1055-1077, which computes one real line as a complex one):
//   Z = FFT_N(z),  A[k] = (Z[k] + conj Z[N-k]) / 2,  B[k] = -i (Z[k] - conj Z[N-k]) / 2,
// so a pair costs one N-point transform.  The output is the transposed spectrum
// (om.s1 = 1: lines 2 p and 2 p + 1 adjacent), and A[k], B[k] leave as one 16-byte
// store per k (om.es even, out 16-byte aligned).  A workgroup takes PPW
// neighbouring pairs, Q waves each, and stores the 2 PPW values of each k as one
// 16 PPW-byte segment.  Grid: npairs / PPW workgroups.
template <int P, int Q, typename TIn, int PPW = 1>
__global__ void __launch_bounds__(64 * Q * PPW)
        k_fwd_full_wave_pair(const TIn* __restrict__ in, LineMap im, float2* __restrict__ out, LineMap om, int nout,
                             int npairs, float pre_scale, const float2* __restrict__ tw) {
    constexpr int T = 64, E = P / T, N = P * Q, NT = T * Q;
    using FFT       = WaveFFT<P>;
    static_assert(Q * FFT::SLOTS >= N, "a line must fit the exchange buffers");
    __shared__ float2 lds_all[PPW][Q * FFT::SLOTS];
    __shared__ float2 s_twp[P];
    __shared__ float2 s_twn[NT];
    const int     g     = __builtin_amdgcn_readfirstlane(threadIdx.x / NT); // pair of the workgroup
    const int     tid   = threadIdx.x % NT;
    const int     q     = __builtin_amdgcn_readfirstlane(tid / T);
    const int     t     = tid % T;
    const int     pair0 = xcd_row(blockIdx.x, npairs / PPW) * PPW;
    const int     pair  = pair0 + g;
    const float2* twm   = multi_twiddles(tw);
    float2*       lds   = lds_all[g];
    // ---- the pair into LDS as one complex line (natural order), every load of the
    // thread in flight before its LDS writes; the first pair's waves stage the tables
    {
        RowTwiddleStage<P, Q> tws;
        if ( g == 0 )
            tws.load(twm);
        const auto* la = as_global(in) + im.base(2 * pair);
        const auto* lb = as_global(in) + im.base(2 * pair + 1);
        float2      z[N / NT];
#pragma unroll
        for ( int r = 0; r < N / NT; r++ ) {
            const int n = tid + NT * r;
            z[r]        = make_float2(load_real(la, n * im.es), load_real(lb, n * im.es));
        }
        if ( g == 0 )
            tws.store(s_twp, s_twn);
#pragma unroll
        for ( int r = 0; r < N / NT; r++ )
            lds[tid + NT * r] = make_float2(z[r].x * pre_scale, z[r].y * pre_scale);
    }
    __syncthreads( );
    // ---- radix-Q pass: thread m holds x[s P + m], s < Q, and leaves y_q[m] at q P + m
    for ( int m = tid; m < P; m += NT ) {
        float2 u[Q];
#pragma unroll
        for ( int s = 0; s < Q; s++ )
            u[s] = lds[s * P + m];
        SmallDFT<Q, -1>::run(u);
#pragma unroll
        for ( int s = 0; s < Q; s++ )
            lds[s * P + m] = u[s];
    }
    __syncthreads( );
    FFT fft;
    fft.init_pt(t, s_twp);
    const float2 lw = lane_twiddle(s_twn, t, q); // W_N^{t q}
    float2       v[E];
#pragma unroll
    for ( int i = 0; i < E; i++ )
        v[i] = cmul(cmul(lds[q * P + t + T * i], twm[N + ((T * i * q) & (N - 1))]), lw);
    __syncthreads( );
    fft.template run<-1>(v, lds + q * FFT::SLOTS, t);
    // ---- Z in natural order through LDS, then A[k], B[k] of the 2 PPW lines
    __syncthreads( );
#pragma unroll
    for ( int i = 0; i < E; i++ )
        lds[dec_slot<P, Q>(Q * (t + T * i) + q)] = v[i];
    __syncthreads( );
    auto* op = as_global(out) + om.base(2 * pair0);
    for ( int j = threadIdx.x; j < nout * PPW; j += NT * PPW ) {
        const int    k  = j / PPW, h = j % PPW;
        const float2 zk = lds_all[h][dec_slot<P, Q>(k)];
        const float2 zn = lds_all[h][dec_slot<P, Q>((N - k) & (N - 1))];
        store_cpx_pair(op + (long)k * om.es, h, make_float2(0.5f * (zk.x + zn.x), 0.5f * (zk.y - zn.y)),
                       make_float2(0.5f * (zk.y + zn.y), -0.5f * (zk.x - zn.x)));
    }
}

} // namespace fastfft_amd

namespace fastfft_amd {

// Full-length inverse line transforms (M = N outputs from 0): the C2C inverse
// passes and the C2R x-pass of an InvFFT without size change (the reference's
// block_fft_kernel_C2C_NONE inverse and C2R_NONE_XY, FastFFT.cu:1369-1403,
// :1641-1660).  With N = P*Q, decimation in time:
//   U_q[m]      = IFFT_P( X[Q k' + q] )[m]                       (wave q)
//   y[m + P s]  = sum_q W_Q^{-s q} ( W_N^{-m q} U_q[m] )         (radix-Q pass in LDS)
// HERM: the input line holds X[0 .. N/2] of a Hermitian sequence (X[N-k] =
// conj X[k] above) and only the real part of y is stored (REAL_OUT).
template <int P, int Q, typename TIn, bool HERM, typename TOut, bool REAL_OUT>
__global__ void __launch_bounds__(64 * Q)
        k_inv_full_wave(const TIn* __restrict__ in, LineMap im, TOut* __restrict__ out, LineMap om, int nlines,
                        float post_scale, const float2* __restrict__ tw) {
    constexpr int T = 64, E = P / T, N = P * Q, NT = T * Q, H = N / 2;
    using FFT       = WaveFFT<P>;
    static_assert(Q * FFT::SLOTS >= N, "a line must fit the exchange buffers");
    __shared__ float2 lds[Q * FFT::SLOTS];
    __shared__ float2 s_twp[P];
    __shared__ float2 s_twn[NT];
    const int     q    = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    const int     t    = threadIdx.x % T;
    const int     line = xcd_row(blockIdx.x, nlines);
    const float2* twm  = multi_twiddles(tw);
    // ---- the input line into LDS, decimated by Q, two values per lane (the
    // decimated layout is conflict-free for pairs), every load of the thread
    // in flight before its LDS writes
    {
        RowTwiddleStage<P, Q> tws;
        tws.load(twm);
        const auto*   lp = as_global(in) + im.base(line);
        constexpr int KN = HERM ? H + 1 : N;
        constexpr int MR = (KN + 2 * NT - 1) / (2 * NT);
        float2        u[MR][2];
#pragma unroll
        for ( int r = 0; r < MR; r++ ) {
            const int m = threadIdx.x + NT * r;
            u[r][0]     = 2 * m < KN ? load_cpx(lp, (long)(2 * m) * im.es) : make_float2(0.f, 0.f);
            u[r][1]     = 2 * m + 1 < KN ? load_cpx(lp, (long)(2 * m + 1) * im.es) : make_float2(0.f, 0.f);
        }
        tws.store(s_twp, s_twn);
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
            const bool lo = k <= H;
            float2     z  = lds[dec_slot<P, Q>(lo ? k : N - k)];
            v[i]          = lo ? z : make_float2(z.x, -z.y);
        }
        else {
            v[i] = lds[dec_slot<P, Q>(k)];
        }
    }
    __syncthreads( );
    float2* wl = lds + q * FFT::SLOTS;
    fft.template run<+1>(v, wl, t);
    // conj(W_N^{m q}) U_q[m], m = t + 64 i, into LDS at q P + m
    const float2 lw = lane_twiddle(s_twn, t, q);
    __syncthreads( );
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const float2 w = cmul(lw, twm[N + ((T * i * q) & (N - 1))]);
        lds[q * P + t + T * i] = cmulc(v[i], w);
    }
    __syncthreads( );
    TOut* op = out + om.base(line);
    for ( int m = threadIdx.x; m < P; m += NT ) {
        float2 u[Q];
#pragma unroll
        for ( int s = 0; s < Q; s++ )
            u[s] = lds[s * P + m];
        SmallDFT<Q, +1>::run(u);
#pragma unroll
        for ( int s = 0; s < Q; s++ ) {
            const long j = (long)(m + P * s) * om.es;
            if constexpr ( REAL_OUT )
                store_real(op, j, u[s].x * post_scale);
            else
                store_cpx(op, j, make_float2(u[s].x * post_scale, u[s].y * post_scale));
        }
    }
}

} // namespace fastfft_amd
