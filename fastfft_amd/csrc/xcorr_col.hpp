// Fused correlation pass with one workgroup of Q wavefronts per column (the
// default hot kernel of FwdImageInvFFT's second axis; replaces the reference's
// block_fft_kernel_C2C_FWD_INCREASE_OP_INV_NONE / _FWD_NONE_INV_DECREASE_ConjMul,
// src/fastfft/FastFFT.cu:20-56, :1295-1344, for the padded-and-trimmed case).
//
// For a column of full length N = P*Q, wave q of the workgroup computes
//   X[Q k' + q] = FFT_P( a[n] W_N^{n q} )[k'],                       n < L
//   Z[Q k' + q] = img[Q k' + q] * conj(X[Q k' + q]) * s
//   U_q         = IFFT_P( Z[Q k' + q] )
// and leaves its share W_N^{-(o+j) q} U_q[(o + j) mod P] of the M outputs in
// LDS; the workgroup sums the Q shares.  The image column is first copied into
// LDS with coalesced 16-byte loads, so that the decimated reads of the waves
// never go to memory.
//
// DC and Nyquist columns.  Over the columns kx = 0 .. NX/2 of a real 2-D
// transform with even NX, the template values of columns 0 and NX/2 are real
// and the C2R x-pass uses only the real part of their outputs.  Workgroup 0
// computes both columns with one transform of z = a0 + i a1 (a0, a1 the real
// template columns): with Z = FFT(z) and H_c[k] = (I_c[k] + conj I_c[-k]) / 2
// the Hermitian part of image column c, the real outputs are
//   Re IFFT(I_0 conj FFT(a0)) + i Re IFFT(I_1 conj FFT(a1))
//       = IFFT( (conj Z[k] (H_0 - H_1)[k] + Z[-k] (H_0 + H_1)[k]) / 2 ),
// so NX/2 workgroups cover the NX/2 + 1 columns and every workgroup has the
// same amount of work.  Those two columns are stored with a zero imaginary part.
#pragma once

#include <type_traits>

#include "rows_wave.hpp"

namespace fastfft_amd {

// 16-byte loads of an image column: PER complex values per load.
template <typename TImg>
struct ImgStage;

template <>
struct ImgStage<float2> {
    static constexpr int PER = 2;
    __device__ __forceinline__ static void load(const float2* __restrict__ p, int m, float2 (&v)[PER]) {
        const float4 f = reinterpret_cast<const float4*>(p)[m];
        v[0]           = make_float2(f.x, f.y);
        v[1]           = make_float2(f.z, f.w);
    }
};

template <>
struct ImgStage<__half2> {
    static constexpr int PER = 4;
    __device__ __forceinline__ static void load(const __half2* __restrict__ p, int m, float2 (&v)[PER]) {
        const uint4    u = reinterpret_cast<const uint4*>(p)[m];
        const unsigned w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for ( int e = 0; e < 4; e++ ) {
            __half2 h;
            __builtin_memcpy(&h, &w[e], sizeof(h));
            v[e] = __half22float2(h);
        }
    }
};

// Grid: nlines - 1 workgroups when pack_ends (columns 0 and last_line share
// workgroup 0), else nlines, times gridDim.y template groups.  img[c*ldg + n]
// is element n of image column c (img and img + ldg 16-byte aligned).  The
// workgroup correlates its column with the templates of the launch: `one`, or
// its group's share of items[0 .. nitems) when items is not null (group b of
// G takes ceil(nitems / G) consecutive items; gridDim.x a multiple of 8 keeps
// the column-to-XCD mapping of the one-group grid).  The image column is read from memory once per launch and kept in
// registers (in LDS for the packed end columns) for all templates.  With
// a_es, o_es > 1 (row-major template spectra and outputs, as the x-pass kernels
// prefer) the template gather and the output scatter are the only strided
// accesses.
//
// Every table a wave needs is copied into LDS once per workgroup with
// coalesced loads (W_P^k for the Stockham passes, W_N^k for k < 64 Q for the
// per-lane input and output twiddles), and the wave-uniform twiddles
// W_N^{64 i q} are scalar loads, so the only vector memory traffic is the
// column data.
//
// ONE (one template per launch, items == nullptr): the image values and the
// per-register twiddles are formed where they are used instead of being held in
// registers for a loop over templates, and the image column is read from memory
// at its product (the Q waves read neighbouring words of the same lines) instead
// of being staged in LDS first (the one-correlation y-pass of a 4096^2 image:
// 35.4 -> 33.0 us).
template <int P, int Q, int RS, typename TImg, bool ONE = false>
__global__ void __launch_bounds__(64 * Q) __attribute__((amdgpu_waves_per_eu(P <= 512 ? 4 : 1)))
        k_xcorr_col(XcorrItem one, const XcorrItem* __restrict__ items, int nitems, long lda, long a_es,
                    const TImg* __restrict__ img, long ldg, long ldo, long o_es, int M, int o, float intra_scale,
                    const float2* __restrict__ tw, int last_line, int pack_ends) {
    constexpr int T = 64, E = P / T, N = P * Q, NT = T * Q, NW = T * Q;
    constexpr int PER = ImgStage<TImg>::PER;
    using FFT         = WaveFFT<P>;
    static_assert(Q * FFT::SLOTS >= N, "the staged column must fit the exchange buffers");
    static_assert(E % PER == 0, "whole 16-byte loads per thread");
    static_assert(RS >= 0 && RS < E, "register rotation out of range");
    __shared__ float2 lds[Q * FFT::SLOTS];
    __shared__ float2 s_twp[P];  // Stockham pass table of WaveFFT<P>
    __shared__ float2 s_twn[NW]; // W_N^{t q} by wave (lane_twiddle)
    __shared__ float2 s_tpl[P];  // template column, first L values
    const int      q    = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    const int      t    = threadIdx.x % T;
    // each XCD takes a contiguous range of columns, so that the strided template
    // gather and output scatter of neighbouring columns share lines in one L2
    const long     col  = xcd_row(blockIdx.x, gridDim.x);
    const bool     ends = pack_ends && col == 0;
    const float2*  twm  = multi_twiddles(tw);
    if ( ! items ) {
        nitems = 1;
    }
    else if ( gridDim.y > 1 ) {
        const int per   = (nitems + (int)gridDim.y - 1) / (int)gridDim.y;
        const int first = (int)blockIdx.y * per;
        items += first;
        nitems = min(per, nitems - first);
        if ( nitems <= 0 ) // the whole workgroup leaves before any barrier
            return;
    }

    // ---- stage: twiddle tables, image column
    stage_row_twiddles<P, Q>(s_twp, s_twn, twm);
    if ( ! ends && ! ONE ) {
        const TImg* gc = img + col * ldg;
#pragma unroll
        for ( int r = 0; r < E / PER; r++ ) {
            const int m = threadIdx.x + NT * r;
            float2    f[PER];
            ImgStage<TImg>::load(gc, m, f);
#pragma unroll
            for ( int e = 0; e < PER; e++ )
                lds[dec_slot<P, Q>(PER * m + e)] = f[e];
        }
    }
    __syncthreads( );
    FFT fft;
    fft.init_pt(t, s_twp);
    const float2 lw = lane_twiddle(s_twn, t, q); // W_N^{t q}
    // The output twiddle of output j = t + 64 i is
    //   conj(W_N^{(o + j) q}) = conj(W_N^{o q}) conj(W_N^{(t + 64 i) q}),
    // the conjugate of the input twiddle times a wave-uniform factor; with the
    // intra scale, sigma = conj(W_N^{o q}) * s is folded into the image values
    // (the inverse transform is linear), so that input and output each take one
    // complex product per value.
    const int    omask = ((o % N) + N) % N;
    const float2 wo    = twm[N + ((omask * q) & (N - 1))];
    const float2 sigma = make_float2(wo.x * intra_scale, -wo.y * intra_scale);
    float2       g[ONE ? 1 : E];
    if ( ! ends && ! ONE ) {
#pragma unroll
        for ( int i = 0; i < E; i++ )
            g[i] = cmul(lds[dec_slot<P, Q>(Q * (t + T * i) + q)], sigma);
    }
    // end columns: conj(W_N^{(o + t) q}) * s, the output twiddle's lane part
    const float2 s  = cmul(make_float2(lw.x, -lw.y), sigma);
    float2*      wl = lds + q * FFT::SLOTS;

    // The packed end columns take their own copy of the template loop, so that
    // neither copy holds the other's values across iterations.
    auto templates = [&](auto ends_c) {
        constexpr bool ENDS = decltype(ends_c)::value;
        // W_N^{(t + 64 i) q}, held for all templates on the common path (formed at each
        // use for one template)
        float2 win[(ENDS || ONE) ? 1 : E];
        auto   win_at = [&](int i) {
            if constexpr ( ONE || ENDS )
                return cmul(twm[N + ((T * i * q) & (N - 1))], lw);
            else
                return win[i];
        };
        if constexpr ( ! ENDS && ! ONE ) {
#pragma unroll
            for ( int i = 0; i < E; i++ )
                win[i] = cmul(twm[N + ((T * i * q) & (N - 1))], lw);
        }
        for ( int it = 0; it < nitems; it++ ) {
            const XcorrItem item = items ? items[it] : one;
            const auto*     ta   = as_global(static_cast<const float2*>(item.a));
            auto*           to   = as_global(static_cast<float2*>(item.out));
            // the previous template's reduction (or the image reads) are done with LDS
            __syncthreads( );
            {
                const auto* ac = ta + col * lda;
                const auto* a1 = ta + (long)last_line * lda;
                for ( int n = threadIdx.x; n < item.L; n += NT )
                    s_tpl[n] = ENDS ? make_float2(load_cpx(ac, n * a_es).x, load_cpx(a1, n * a_es).x) : load_cpx(ac, n * a_es);
            }
            __syncthreads( );
            // input a[n] W_N^{n q}, n = t + 64 i, with W_N^{n q} = W_N^{t q} W_N^{64 i q} (one for
            // sub-transform 0, whose wave skips the products: the branch is wave-uniform)
            float2 v[E];
#pragma unroll
            for ( int i = 0; i < E; i++ ) {
                const int    n = t + T * i;
                const float2 x = n < item.L ? s_tpl[n] : make_float2(0.f, 0.f);
                if constexpr ( ! ENDS )
                    v[i] = x;
                else
                    v[i] = cmul(cmul(x, twm[N + ((T * i * q) & (N - 1))]), lw);
            }
            if constexpr ( ! ENDS ) {
                if ( q != 0 ) {
#pragma unroll
                    for ( int i = 0; i < E; i++ )
                        v[i] = cmul(v[i], win_at(i));
                }
            }
            fft.template run<-1>(v, wl, t);
            if constexpr ( ! ENDS && ONE ) {
                const auto* gc = as_global(img) + col * ldg;
#pragma unroll
                for ( int i = 0; i < E; i++ )
                    v[i] = cmulc(cmul(load_cpx(gc, (long)(Q * (t + T * i) + q)), sigma), v[i]);
            }
            else if constexpr ( ! ENDS ) {
#pragma unroll
                for ( int i = 0; i < E; i++ )
                    v[i] = cmulc(g[i], v[i]); // img * sigma * conj(X)
            }
            else {
                // Z[-k] for k = Q k' + q: -k = Q k'' + q' with q' = (Q - q) mod Q and
                // k'' = (P - k') mod P when q = 0, P - 1 - k' otherwise (the partner
                // wave's region).  The image values are reloaded for every template;
                // the opaque pointers and lane index keep these template-invariant
                // loads and indices from being hoisted out of the loop into registers.
#pragma unroll
                for ( int i = 0; i < E; i++ )
                    wl[FFT::Layout::slot(t + T * i)] = v[i];
                __syncthreads( );
                const int   qp = (Q - q) % Q;
                const auto* g0 = as_global(img);
                const auto* g1 = as_global(img) + (long)last_line * ldg;
                int         tt = t;
                asm volatile("" : "+s"(g0), "+s"(g1), "+v"(tt));
#pragma unroll
                for ( int i = 0; i < E; i++ ) {
                    const int    kp  = tt + T * i;
                    const int    kpp = q == 0 ? (P - kp) & (P - 1) : P - 1 - kp;
                    const float2 zm  = lds[qp * FFT::SLOTS + FFT::Layout::slot(kpp)];
                    const int    k   = Q * kp + q;
                    const int    km  = (N - k) & (N - 1);
                    const float2 i0 = load_cpx(g0, k), i0m = load_cpx(g0, km);
                    const float2 i1 = load_cpx(g1, k), i1m = load_cpx(g1, km);
                    const float2 h0 = make_float2(i0.x + i0m.x, i0.y - i0m.y); // 2 H_0[k]
                    const float2 h1 = make_float2(i1.x + i1m.x, i1.y - i1m.y); // 2 H_1[k]
                    const float2 d  = csub(h0, h1), sm = cadd(h0, h1);
                    const float2 r  = cadd(cmul(make_float2(v[i].x, -v[i].y), d), cmul(zm, sm));
                    v[i]            = make_float2(0.25f * r.x, 0.25f * r.y);
                }
                __syncthreads( );
            }
            fft.template run<+1>(v, wl, t);
            // share e^{2 pi i (o + j) q / N} U_q[(o + j) mod P] of output j = t + 64 i
            // (sigma already applied on the common path; no twiddle for sub-transform 0)
            if constexpr ( ! ENDS ) {
                if ( q != 0 ) {
#pragma unroll
                    for ( int i = 0; i < E; i++ )
                        v[(i + RS) % E] = cmulc(v[(i + RS) % E], win_at(i));
                }
            }
#pragma unroll
            for ( int i = 0; i < E; i++ ) {
                if constexpr ( ! ENDS ) {
                    wl[FFT::Layout::slot(t + T * i)] = v[(i + RS) % E];
                }
                else {
                    float2 c = twm[N + ((T * i * q) & (N - 1))];
                    c.y      = -c.y;
                    wl[FFT::Layout::slot(t + T * i)] = cmul(cmul(v[(i + RS) % E], c), s);
                }
            }
            __syncthreads( );
            for ( int j = threadIdx.x; j < M; j += NT ) {
                float2    acc = make_float2(0.f, 0.f);
                const int sj  = FFT::Layout::slot(j);
#pragma unroll
                for ( int w = 0; w < Q; w++ )
                    acc = cadd(acc, lds[w * FFT::SLOTS + sj]);
                if constexpr ( ! ENDS ) {
                    store_cpx(to, col * ldo + j * o_es, acc);
                }
                else {
                    store_cpx(to, j * o_es, make_float2(acc.x, 0.f));
                    store_cpx(to, (long)last_line * ldo + j * o_es, make_float2(acc.y, 0.f));
                }
            }
        }
    };
    if ( ends )
        templates(std::true_type{ });
    else
        templates(std::false_type{ });
}

} // namespace fastfft_amd
