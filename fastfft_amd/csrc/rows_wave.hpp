// Row kernels of the 2-D path with one wavefront per sub-transform.
//
//   k_r2c_rows_wave  R2C x-pass of a zero-padded real row (the reference's
//                    block_fft_kernel_R2C_INCREASE_XY / _NONE_XY,
//                    src/fastfft/FastFFT.cu:1055-1166): for each q the wave q of the
//                    row's block computes X[Q k' + q] = FFT_P(x[n] W_N^{n q})[k'] and
//                    stores the kx <= N/2 values transposed, A[kx][row].
//   k_c2r_rows_wave  trimmed C2R x-pass (block_fft_kernel_C2R_DECREASE_XY / _NONE_XY,
//                    :1641-1692, without the overlapping-row writes): wave q computes
//                    U_q = IFFT_P of the Hermitian-extended X[Q s + q] gathered from
//                    B[kx][row]; the real part of
//                       y[o + j] = sum_q e^{2 pi i (o + j) q / N} U_q[(o + j) mod P]
//                    is reduced over the waves through LDS and stored as one row.
//
// One block per row: Q waves (Q <= 16), each holding a P-point sub-transform
// with E = P/64 values per lane.  Blocks are mapped to rows so that the blocks
// one XCD runs (block b runs on XCD b mod 8) own a contiguous range of rows:
// the transposed accesses A[kx][row] / B[kx][row] of neighbouring rows share
// cache lines, which then meet in that XCD's L2.
#pragma once

#include "line_kernels.hpp"
#include "wave_fft.hpp"

namespace fastfft_amd {

constexpr int kXcds = 8;

// Block -> row bijection giving each XCD a contiguous range of rows (identity
// when the row count is not a multiple of the XCD count).
__device__ __forceinline__ int xcd_row(int b, int nrows) {
    if ( nrows % kXcds != 0 )
        return b;
    return (b % kXcds) * (nrows / kXcds) + b / kXcds;
}

template <int P, int Q, typename TIn>
__global__ void __launch_bounds__(64 * Q)
        k_r2c_rows_wave(const TIn* __restrict__ in, LineMap im, int L, float2* __restrict__ out, LineMap om, int nout,
                        int nrows, float pre_scale, const float2* __restrict__ tw) {
    constexpr int T = 64, E = P / T, N = P * Q, TWS = kTwiddleN / N;
    __shared__ float2 lds[Q * WaveFFT<P>::SLOTS];
    const int q   = threadIdx.x / T;
    const int t   = threadIdx.x % T;
    const int row = xcd_row(blockIdx.x, nrows);
    WaveFFT<P> fft;
    fft.init(t, tw);
    const TIn* rp = in + im.base(row);
    float2     v[E];
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const int   n = t + T * i;
        const float x = n < L ? load_real(rp, n * im.es) * pre_scale : 0.f;
        const float2 w = tw[((n * q) & (N - 1)) * TWS];
        v[i]           = make_float2(x * w.x, x * w.y);
    }
    fft.template run<-1>(v, lds + q * WaveFFT<P>::SLOTS, t);
    float2* op = out + om.base(row);
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const int k = Q * (t + T * i) + q;
        if ( k < nout )
            op[k * om.es] = v[i];
    }
}

template <int P, int Q, int RS, typename TOut>
__global__ void __launch_bounds__(64 * Q)
        k_c2r_rows_wave(const float2* __restrict__ in, LineMap im, TOut* __restrict__ out, LineMap om, int M, int o, int nrows,
                        float post_scale, const float2* __restrict__ tw) {
    constexpr int T = 64, E = P / T, N = P * Q, TWS = kTwiddleN / N;
    __shared__ float2 lds[Q * WaveFFT<P>::SLOTS];
    const int q   = threadIdx.x / T;
    const int t   = threadIdx.x % T;
    const int row = xcd_row(blockIdx.x, nrows);
    WaveFFT<P> fft;
    fft.init(t, tw);
    const float2* cp = in + im.base(row);
    float2        v[E];
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const int k = Q * (t + T * i) + q;
        float2    z;
        if ( k <= N / 2 ) {
            z = cp[k * im.es];
        }
        else {
            z   = cp[(N - k) * im.es];
            z.y = -z.y;
        }
        v[i] = z;
    }
    float2* wl = lds + q * WaveFFT<P>::SLOTS;
    fft.template run<+1>(v, wl, t);
    // this wave's real contribution to every output sample, into its own LDS region
    float*    red   = reinterpret_cast<float*>(wl);
    const int omask = ((o % N) + N) % N;
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const int    j = t + T * i;
        const float2 u = v[(i + RS) % E];
        // e^{+2 pi i (o + j) q / N} = conj(W_N^{(o + j) q})
        const float2 w = tw[(((omask + j) * q) & (N - 1)) * TWS];
        red[j]         = u.x * w.x + u.y * w.y;
    }
    __syncthreads( );
    TOut* op = out + om.base(row);
    for ( int j = threadIdx.x; j < M; j += blockDim.x ) {
        float s = 0.f;
#pragma unroll
        for ( int w = 0; w < Q; w++ )
            s += reinterpret_cast<const float*>(lds + w * WaveFFT<P>::SLOTS)[j];
        store_real(op, j * om.es, s * post_scale);
    }
}

} // namespace fastfft_amd
