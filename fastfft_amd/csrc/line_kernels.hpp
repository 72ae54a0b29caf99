// Line kernels: one 1-D size-changing transform per "line" (a row, a column
// or a z-line of the array), built on the Stockham engine of fft_core.hpp.
//
//   k_fwd    input-pruned forward transform (L nonzero inputs, all N or N/2+1
//            outputs).  Real input = the reference's R2C_*_XY / R2C_*_XZ
//            x-pass (src/fastfft/FastFFT.cu:1055-1230); complex input = the
//            C2C_NONE / C2C_INCREASE y- and z-passes (:1348-1367, :1433-1475,
//            :1559-1620).
//   k_inv    output-pruned inverse (all N inputs, M outputs starting at
//            wrapped index o).  Complex input = C2C inverse passes
//            (:1405-1431, :1369-1403); Hermitian half-spectrum input with a
//            real output = the C2R_NONE_XY / C2R_DECREASE_XY x-pass
//            (:1641-1692) without its overlapping-row write race.
//   k_xcorr  fused second-axis pass of FwdImageInvFFT: input-pruned forward,
//            CONJ_MUL with the image spectrum, output-pruned inverse
//            (:20-56 and :1295-1344, include/detail/functors.h:31-60).
//
// A LineMap addresses element e of line l at
//   (l / n2) * s1 + (l % n2) * s2 + e * es        (in elements of the buffer type),
// which covers contiguous rows, transposed columns and 3-D z-lines alike.
#pragma once

#include "fft_core.hpp"

namespace fastfft_amd {

struct LineMap {
    long s1, s2, es;
    int  n2;
    __device__ __forceinline__ long base(int line) const { return (long)(line / n2) * s1 + (long)(line % n2) * s2; }
};

template <int P>
struct LineShape {
    static constexpr int E     = FFTShape<P>::E;
    static constexpr int T     = FFTShape<P>::T;
    static constexpr int BLOCK = T >= 64 ? T : 64;
    static constexpr int G     = BLOCK / T; // lines per workgroup
    static constexpr int LDS   = G * lds_slots(P); // complex values of LDS per workgroup
};

// ------------------------------------------------------------------ forward
// REAL_IN: TIn is float/__half (x-pass of real data), else float2/__half2.
template <int P, typename TIn, bool REAL_IN, typename TOut>
__global__ void __launch_bounds__(LineShape<P>::BLOCK)
        k_fwd(const TIn* __restrict__ in, LineMap im, int L, TOut* __restrict__ out, LineMap om, int N, int Q, int nout,
              int nlines, float pre_scale, const float2* __restrict__ tw, int tw_stride) {
    using S                = LineShape<P>;
    constexpr int E        = S::E;
    constexpr int T        = S::T;
    extern __shared__ float2 smem[];
    const int g    = threadIdx.x / T;
    const int t    = threadIdx.x % T;
    const int line = blockIdx.x * S::G + g;
    const bool active = line < nlines;
    float2* lds = smem + g * lds_slots(P);

    float2 x[E];
    const long ib = active ? im.base(line) : 0;
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const int n = t + T * i;
        if ( active && n < L ) {
            if constexpr ( REAL_IN )
                x[i] = make_float2(load_real(in, ib + n * im.es) * pre_scale, 0.f);
            else
                x[i] = cscale(load_cpx(in, ib + n * im.es), pre_scale);
        }
        else {
            x[i] = make_float2(0.f, 0.f);
        }
    }
    const long ob   = active ? om.base(line) : 0;
    const int  mask = N - 1;
    for ( int q = 0; q < Q; q++ ) {
        float2 v[E];
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int n = t + T * i;
            if ( q == 0 ) {
                v[i] = x[i];
            }
            else {
                const float2 w = twiddle<-1>(tw, ((n * q) & mask) * tw_stride);
                if constexpr ( REAL_IN )
                    v[i] = make_float2(x[i].x * w.x, x[i].x * w.y);
                else
                    v[i] = cmul(x[i], w);
            }
        }
        block_fft<P, E, -1>(v, lds, t, tw);
        if ( active ) {
#pragma unroll
            for ( int i = 0; i < E; i++ ) {
                const int k = Q * (t + T * i) + q;
                if ( k < nout )
                    store_cpx(out, ob + k * om.es, v[i]);
            }
        }
    }
}

// ------------------------------------------------------------------ inverse
// HERM_IN: the line stores k = 0..N/2 of a Hermitian spectrum; k > N/2 is
// conj(Z[N-k]).  REAL_OUT: keep the real part (C2R), TOut float/__half;
// otherwise TOut is float2/__half2.
template <int P, typename TIn, bool HERM_IN, typename TOut, bool REAL_OUT>
__global__ void __launch_bounds__(LineShape<P>::BLOCK)
        k_inv(const TIn* __restrict__ in, LineMap im, TOut* __restrict__ out, LineMap om, int N, int M, int o, int Q,
              int nlines, float post_scale, const float2* __restrict__ tw, int tw_stride) {
    using S         = LineShape<P>;
    constexpr int E = S::E;
    constexpr int T = S::T;
    extern __shared__ float2 smem[];
    const int g    = threadIdx.x / T;
    const int t    = threadIdx.x % T;
    const int line = blockIdx.x * S::G + g;
    const bool active = line < nlines;
    float2* lds = smem + g * lds_slots(P);

    const long ib   = active ? im.base(line) : 0;
    const int  mask = N - 1;
    const int  om_  = o & mask; // o mod N (two's complement)
    const int  half = N / 2;
    float2 acc[E];
#pragma unroll
    for ( int i = 0; i < E; i++ )
        acc[i] = make_float2(0.f, 0.f);

    for ( int r = 0; r < Q; r++ ) {
        float2 v[E];
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int k = Q * (t + T * i) + r;
            float2    z = make_float2(0.f, 0.f);
            if ( active ) {
                if constexpr ( HERM_IN ) {
                    if ( k <= half )
                        z = load_cpx(in, ib + k * im.es);
                    else {
                        z   = load_cpx(in, ib + (N - k) * im.es);
                        z.y = -z.y;
                    }
                }
                else {
                    z = load_cpx(in, ib + k * im.es);
                }
            }
            if ( om_ != 0 )
                z = cmul(z, twiddle<+1>(tw, ((om_ * k) & mask) * tw_stride));
            v[i] = z;
        }
        block_fft<P, E, +1>(v, lds, t, tw);
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int j = t + T * i;
            if ( r == 0 )
                acc[i] = cadd(acc[i], v[i]);
            else
                acc[i] = cadd(acc[i], cmul(v[i], twiddle<+1>(tw, (j * r) * tw_stride)));
        }
    }
    if ( active ) {
        const long ob = om.base(line);
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int j = t + T * i;
            if ( j < M ) {
                if constexpr ( REAL_OUT )
                    store_real(out, ob + j * om.es, acc[i].x * post_scale);
                else
                    store_cpx(out, ob + j * om.es, cscale(acc[i], post_scale));
            }
        }
    }
}

// ------------------------------------------------------------------ fused correlation pass
// For each line: X = FFT_N(pad_L(a)), Z = img * conj(X) * intra_scale,
// out[j] = IFFT_N(Z)[o + j] for j < M.  a and out are internal float2 buffers.
template <int P, typename TImg>
__global__ void __launch_bounds__(LineShape<P>::BLOCK)
        k_xcorr(const float2* __restrict__ a, LineMap am, int L, const TImg* __restrict__ img, LineMap gm,
                float2* __restrict__ out, LineMap om, int N, int M, int o, int Q, int nlines, float intra_scale,
                const float2* __restrict__ tw, int tw_stride) {
    using S         = LineShape<P>;
    constexpr int E = S::E;
    constexpr int T = S::T;
    extern __shared__ float2 smem[];
    const int g    = threadIdx.x / T;
    const int t    = threadIdx.x % T;
    const int line = blockIdx.x * S::G + g;
    const bool active = line < nlines;
    float2* lds = smem + g * lds_slots(P);

    const long ab   = active ? am.base(line) : 0;
    const long gb   = active ? gm.base(line) : 0;
    const int  mask = N - 1;
    const int  om_  = o & mask;
    float2 x[E], acc[E];
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const int n = t + T * i;
        x[i]        = (active && n < L) ? a[ab + n * am.es] : make_float2(0.f, 0.f);
        acc[i]      = make_float2(0.f, 0.f);
    }
    for ( int q = 0; q < Q; q++ ) {
        float2 v[E];
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int n = t + T * i;
            v[i] = q == 0 ? x[i] : cmul(x[i], twiddle<-1>(tw, ((n * q) & mask) * tw_stride));
        }
        block_fft<P, E, -1>(v, lds, t, tw);
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int k = Q * (t + T * i) + q;
            float2    gk = active ? load_cpx(img, gb + k * gm.es) : make_float2(0.f, 0.f);
            float2    z  = cscale(cmulc(gk, v[i]), intra_scale);
            if ( om_ != 0 )
                z = cmul(z, twiddle<+1>(tw, ((om_ * k) & mask) * tw_stride));
            v[i] = z;
        }
        block_fft<P, E, +1>(v, lds, t, tw);
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int j = t + T * i;
            acc[i] = (q == 0) ? cadd(acc[i], v[i]) : cadd(acc[i], cmul(v[i], twiddle<+1>(tw, (j * q) * tw_stride)));
        }
    }
    if ( active ) {
        const long ob = om.base(line);
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int j = t + T * i;
            if ( j < M )
                out[ob + j * om.es] = acc[i];
        }
    }
}

} // namespace fastfft_amd
