// Fused correlation pass, one wavefront per line (the hot kernel of
// FwdImageInvFFT's second axis; replaces the reference's
// block_fft_kernel_C2C_FWD_INCREASE_OP_INV_NONE / _FWD_NONE_INV_DECREASE_ConjMul,
// src/fastfft/FastFFT.cu:20-56, :1295-1344, for the padded-and-trimmed case).
//
// For one line (a kx column of the transposed spectrum) of full length N = P*Q:
//   X[Q k' + q] = FFT_P( a[n] W_N^{n q} )[k'],          n < L  (input pruned)
//   Z[k]        = img[k] * conj(X[k]) * s                         (CONJ_MUL[_THEN_SCALE])
//   out[j]      = sum_q W_N^{-(o+j) q} IFFT_P( Z[Q s + q] )[(o + j) mod P],  j < M
// The wave holds all Q sub-transforms of the line in registers (lane t owns
// k' = t + 64 i for every q), so the image is read as QB consecutive complex
// values per (lane, i), the q-sum stays in registers and no workgroup barrier
// is ever needed.  The window origin o must be a multiple of 64; its register
// rotation RS = (o/64) mod E is a template parameter.
#pragma once

#include <utility>

#include "wave_fft.hpp"

namespace fastfft_amd {

template <int... Is, class F>
__device__ __forceinline__ void static_for_impl(std::integer_sequence<int, Is...>, F&& f) {
    (f(std::integral_constant<int, Is>{ }), ...);
}

// f(std::integral_constant<int, 0>) ... f(std::integral_constant<int, N-1>)
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(std::make_integer_sequence<int, N>{ }, f);
}

template <int QB, typename TImg>
struct ImgLoad;

template <int QB>
struct ImgLoad<QB, float2> {
    __device__ __forceinline__ static void run(const float2* __restrict__ p, float2 (&g)[QB]) {
        if constexpr ( QB % 2 == 0 ) {
            const float4* p4 = reinterpret_cast<const float4*>(p);
#pragma unroll
            for ( int b = 0; b < QB / 2; b++ ) {
                float4 f     = p4[b];
                g[2 * b]     = make_float2(f.x, f.y);
                g[2 * b + 1] = make_float2(f.z, f.w);
            }
        }
        else {
#pragma unroll
            for ( int b = 0; b < QB; b++ )
                g[b] = p[b];
        }
    }
};

template <int QB>
struct ImgLoad<QB, __half2> {
    __device__ __forceinline__ static void run(const __half2* __restrict__ p, float2 (&g)[QB]) {
#pragma unroll
        for ( int b = 0; b < QB; b++ )
            g[b] = __half22float2(p[b]);
    }
};

template <int P, int Q, int QB, int RS, typename TImg>
__global__ void __launch_bounds__(64)
        k_xcorr_wave(const float2* __restrict__ a, long lda, int L, const TImg* __restrict__ img, long ldg,
                     float2* __restrict__ out, long ldo, int M, int o, float intra_scale, const float2* __restrict__ tw) {
    constexpr int T = 64, E = P / T, N = P * Q, TWS = kTwiddleN / N;
    static_assert(Q % QB == 0, "q batches must tile Q");
    static_assert(RS >= 0 && RS < E, "register rotation out of range");
    __shared__ float2 lds[QB * WaveFFT<P>::SLOTS];
    const int  t   = threadIdx.x;
    const long col = blockIdx.x;

    WaveFFT<P> fft;
    fft.init(t, tw);

    float2 x[E], acc[E];
    const float2* ac = a + col * lda;
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const int n = t + T * i;
        x[i]        = n < L ? ac[n] : make_float2(0.f, 0.f);
        acc[i]      = make_float2(0.f, 0.f);
    }
    const TImg* gc    = img + col * ldg;
    const int   omask = ((o % N) + N) % N; // o mod N

    static_for<Q / QB>([&](auto QI) {
        constexpr int q0 = QI * QB;
        // keep batches apart: interleaving them only raises register pressure
        __builtin_amdgcn_sched_barrier(0);
        // issue this batch's global loads first, so that they are in flight
        // during the forward transform; the empty asm pins the address (and so
        // the loads) to this batch instead of letting them all move to the top
        int goff = Q * t + q0;
        asm volatile("" : "+v"(goff));
        float2 g[E][QB];
#pragma unroll
        for ( int i = 0; i < E; i++ )
            ImgLoad<QB, TImg>::run(gc + goff + (long)Q * T * i, g[i]);
        float2 lw[QB], cq[QB]; // W_N^{t q}, W_N^{o q}
#pragma unroll
        for ( int b = 0; b < QB; b++ ) {
            lw[b] = tw[((t * (q0 + b)) & (N - 1)) * TWS];
            cq[b] = tw[((omask * (q0 + b)) & (N - 1)) * TWS];
        }
        float2 v[QB][E];
        // input twiddle W_N^{(t + 64 i) q} = W_N^{t q} * W_N^{64 i q} (the second factor a constant)
        static_for<QB>([&](auto B) {
            constexpr int q = q0 + B;
            static_for<E>([&](auto I) {
                float2 u = mul_const_tw<(long)T * I * q, N, -1>(x[I]);
                v[B][I]  = q == 0 ? u : cmul(u, lw[B]);
            });
        });
        fft.template run_batch<-1, QB>(v, lds, t);
        // conj-mul with the image: Z = g * conj(X)
#pragma unroll
        for ( int i = 0; i < E; i++ )
#pragma unroll
            for ( int b = 0; b < QB; b++ )
                v[b][i] = cmulc(g[i][b], v[b][i]);
        fft.template run_batch<+1, QB>(v, lds, t);
        // accumulate: output register i' takes the IFFT value of register (i' + RS) mod E
        // times W_N^{-(o + t + 64 i') q} = conj(W_N^{o q} W_N^{t q}) * W_N^{-64 i' q}
        static_for<QB>([&](auto B) {
            constexpr int q = q0 + B;
            float2        s = cmul(lw[B], cq[B]);
            s               = make_float2(s.x * intra_scale, -s.y * intra_scale);
            static_for<E>([&](auto I) {
                constexpr int src = (I + RS) % E;
                float2        u   = mul_const_tw<(long)T * I * q, N, +1>(v[B][src]);
                acc[I].x          = fmaf(u.x, s.x, fmaf(-u.y, s.y, acc[I].x));
                acc[I].y          = fmaf(u.x, s.y, fmaf(u.y, s.x, acc[I].y));
            });
        });
    });
    float2* oc = out + col * ldo;
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const int j = t + T * i;
        if ( j < M )
            oc[j] = acc[i];
    }
}

} // namespace fastfft_amd
