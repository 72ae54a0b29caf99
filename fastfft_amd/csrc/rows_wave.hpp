// Row kernels of the 2-D path with one wavefront per sub-transform.
//
//   k_r2c_rows_wave  R2C x-pass of a zero-padded real row (the reference's
//                    block_fft_kernel_R2C_INCREASE_XY / _NONE_XY,
//                    src/fastfft/FastFFT.cu:1055-1166): wave q of the row's
//                    workgroup computes X[Q k' + q] = FFT_P(x[n] W_N^{n q})[k'] and
//                    the kx < nout values are stored, either as one contiguous
//                    row (unit output stride, written through LDS with 16-byte
//                    stores) or transposed, one value per output line.
//   k_c2r_rows_wave  trimmed C2R x-pass (block_fft_kernel_C2R_DECREASE_XY / _NONE_XY,
//                    :1641-1692, without the overlapping-row writes): wave q computes
//                    U_q = IFFT_P of the Hermitian-extended X[Q s + q]; the real part of
//                       y[o + j] = sum_q e^{2 pi i (o + j) q / N} U_q[(o + j) mod P]
//                    is reduced over the waves through LDS and stored as one row.
//                    A contiguous input row is first copied into LDS with
//                    coalesced loads; a strided one is gathered directly.
//
// One workgroup per row: Q waves (Q <= 16), each holding a P-point sub-transform
// with E = P/64 values per lane.  The Stockham twiddles (W_P^k) and the per-lane
// twiddles W_N^{t q} come from LDS copies made once per workgroup, the
// wave-uniform W_N^{64 i q} from scalar loads.  Workgroups are mapped to rows so
// that the workgroups one XCD runs (workgroup b runs on XCD b mod 8) own a
// contiguous range of rows, which keeps strided accesses of neighbouring rows
// in one L2.
#pragma once

#include "items.hpp"
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

// Descriptor i of a batched launch (i wave-uniform), by scalar loads.
__device__ __forceinline__ RowBatchItem load_item(const RowBatchItem* items, int i) {
    const auto*  p = as_const(items + i);
    RowBatchItem r;
    r.in    = p->in;
    r.out   = p->out;
    r.keys  = p->keys;
    r.in_ld = p->in_ld;
    r.L     = p->L;
    r.rows  = p->rows;
    return r;
}

__device__ __forceinline__ XcorrItem load_item(const XcorrItem* items, int i) {
    const auto* p = as_const(items + i);
    XcorrItem   r;
    r.a      = p->a;
    r.out    = p->out;
    r.L      = p->L;
    r.unused = 0;
    return r;
}

// LDS slot of element k = Q k' + q of a row: the Q decimated sequences one after
// another, each XOR-swizzled inside blocks of 32 slots, so that wave q's reads
// or writes of consecutive k' and the row-order accesses of two consecutive k
// per lane are free of bank conflicts (64 banks of 4 bytes serve a 64-bit
// access 32 lanes at a time; the 32 lanes of such a group then hold 64/Q
// values of k' for each of the Q/2 even-odd pairs of q, which the swizzle
// moves to distinct groups of 64/Q slots).
template <int P, int Q>
__device__ __forceinline__ int dec_slot(int k) {
    const int q = k % Q, kp = k / Q;
    return q * P + (kp ^ (((q >> 1) * (64 / Q)) & 31));
}

// LDS tables of the twiddles a workgroup of Q waves needs, laid out so that
// every wave reads consecutive slots: s_twp = the Stockham pass table of
// WaveFFT<P> (init_pt), s_twn[64 q + t] = W_N^{t q} (lane_twiddle).
template <int P, int Q>
struct RowTwiddleStage {
    static constexpr int NT = 64 * Q, NR = WaveFFT<P>::template pass_table_rows<NT>;
    float2               pt[NR];
    float2               ln;
    // all loads of a thread (a workgroup of NT threads), so that the caller's
    // own loads can be issued before the stores
    __device__ __forceinline__ void load(const float2* __restrict__ twm) {
        WaveFFT<P>::template load_pass_table<NT>(pt, twm, threadIdx.x);
        ln = load_cpx(as_global(twm), P * Q + (threadIdx.x & 63) * (threadIdx.x >> 6));
    }
    __device__ __forceinline__ void store(float2* s_twp, float2* s_twn) const {
        WaveFFT<P>::template store_pass_table<NT>(s_twp, pt, threadIdx.x);
        s_twn[threadIdx.x] = ln;
    }
};

template <int P, int Q>
__device__ __forceinline__ void stage_row_twiddles(float2* s_twp, float2* s_twn, const float2* __restrict__ twm) {
    RowTwiddleStage<P, Q> st;
    st.load(twm);
    st.store(s_twp, s_twn);
}

// W_N^{t q} of lane t of wave q from the table of stage_row_twiddles
__device__ __forceinline__ float2 lane_twiddle(const float2* s_twn, int t, int q) { return s_twn[64 * q + t]; }

// items == nullptr: one row per workgroup of in / out; else row blockIdx.x of
// template items[blockIdx.y] (rows past its count return at once).
template <int P, int Q, typename TIn>
__global__ void __launch_bounds__(64 * Q)
        k_r2c_rows_wave(const TIn* __restrict__ in, LineMap im, int L, float2* __restrict__ out, LineMap om, int nout,
                        int nrows, float pre_scale, const float2* __restrict__ tw, const RowBatchItem* __restrict__ items) {
    constexpr int T = 64, E = P / T, N = P * Q, NT = T * Q;
    using FFT       = WaveFFT<P>;
    static_assert(Q * FFT::SLOTS >= N, "a row must fit the exchange buffers");
    __shared__ float2 lds[Q * FFT::SLOTS];
    __shared__ float2 s_twp[P];
    __shared__ float2 s_twn[NT];
    __shared__ float  s_row[P];
    const int     q   = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    const int     t   = threadIdx.x % T;
    const int     row = xcd_row(blockIdx.x, nrows);
    const float2* twm = multi_twiddles(tw);
    if ( items ) {
        const RowBatchItem it = load_item(items, blockIdx.y);
        if ( row >= it.rows )
            return;
        in    = static_cast<const TIn*>(it.in);
        out   = static_cast<float2*>(it.out);
        im.s1 = it.in_ld;
        L     = it.L;
    }
    RowTwiddleStage<P, Q> tws;
    tws.load(twm);
    {
        // L <= P: every load of the thread in flight before the LDS writes
        constexpr int RR = (P + NT - 1) / NT;
        const auto*   rp = as_global(in) + im.base(row);
        float         xr[RR];
#pragma unroll
        for ( int r = 0; r < RR; r++ ) {
            const int n = threadIdx.x + NT * r;
            xr[r]       = n < L ? load_real(rp, n * im.es) : 0.f;
        }
        tws.store(s_twp, s_twn);
#pragma unroll
        for ( int r = 0; r < RR; r++ ) {
            const int n = threadIdx.x + NT * r;
            if ( n < L )
                s_row[n] = xr[r] * pre_scale;
        }
    }
    __syncthreads( );
    FFT fft;
    fft.init_pt(t, s_twp);
    const float2 lw = lane_twiddle(s_twn, t, q); // W_N^{t q}
    float2       v[E];
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const int    n = t + T * i;
        const float  x = n < L ? s_row[n] : 0.f;
        const float2 c = twm[N + ((T * i * q) & (N - 1))]; // W_N^{64 i q}, wave-uniform
        v[i]           = cmul(make_float2(x * c.x, x * c.y), lw);
    }
    fft.template run<-1>(v, lds + q * FFT::SLOTS, t);
    auto* op = as_global(out) + om.base(row);
    if ( om.es == 1 ) {
        // row order through LDS, then 16-byte stores (op is 16-byte aligned)
        __syncthreads( );
#pragma unroll
        for ( int i = 0; i < E; i++ )
            lds[dec_slot<P, Q>(Q * (t + T * i) + q)] = v[i];
        __syncthreads( );
        for ( int m = threadIdx.x; 2 * m < nout; m += NT ) {
            const float2 u0 = lds[dec_slot<P, Q>(2 * m)];
            if ( 2 * m + 1 < nout ) {
                const float2 u1                      = lds[dec_slot<P, Q>(2 * m + 1)];
                store_cpx_pair(op, m, u0, u1);
            }
            else {
                store_cpx(op, 2 * m, u0);
            }
        }
    }
    else {
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int k = Q * (t + T * i) + q;
            if ( k < nout )
                store_cpx(op, (long)k * om.es, v[i]);
        }
    }
}

// Largest key over the workgroup, returned to thread 0 (wave reduction, then LDS).
__device__ __forceinline__ unsigned long long block_max_key(unsigned long long key) {
    __shared__ unsigned long long s_key;
#pragma unroll
    for ( int off = 32; off > 0; off >>= 1 ) {
        const unsigned long long other = __shfl_xor(key, off);
        key                            = other > key ? other : key;
    }
    if ( threadIdx.x == 0 )
        s_key = 0;
    __syncthreads( );
    if ( (threadIdx.x & 63) == 0 )
        atomicMax(&s_key, key);
    __syncthreads( );
    return s_key;
}

// row_keys == nullptr: no argmax; else workgroup b stores the largest argmax key
// (fft_core.hpp) of its row, index row * M + j, to row_keys[b] for a one-workgroup
// reduction after the launch.  items != nullptr: template items[blockIdx.y]
// supplies in, out and row_keys.
template <int P, int Q, int RS, typename TOut>
__global__ void __launch_bounds__(64 * Q)
        k_c2r_rows_wave(const float2* __restrict__ in, LineMap im, TOut* __restrict__ out, LineMap om, int M, int o, int nrows,
                        float post_scale, const float2* __restrict__ tw, unsigned long long* __restrict__ row_keys,
                        const RowBatchItem* __restrict__ items) {
    constexpr int T = 64, E = P / T, N = P * Q, NT = T * Q, H = N / 2;
    using FFT       = WaveFFT<P>;
    static_assert(Q * FFT::SLOTS >= N, "a row must fit the exchange buffers");
    __shared__ float2 lds[Q * FFT::SLOTS];
    __shared__ float2 s_twp[P];
    __shared__ float2 s_twn[NT];
    const int     q   = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    const int     t   = threadIdx.x % T;
    const int     row = xcd_row(blockIdx.x, nrows);
    const float2* twm = multi_twiddles(tw);
    if ( items ) {
        const RowBatchItem it = load_item(items, blockIdx.y);
        in                    = static_cast<const float2*>(it.in);
        out                   = static_cast<TOut*>(it.out);
        row_keys              = it.keys;
    }
    RowTwiddleStage<P, Q> tws;
    tws.load(twm);
    const auto* cp     = as_global(in) + im.base(row);
    const bool  staged = im.es == 1;
    {
        // two values per lane (the decimated layout is conflict-free for pairs),
        // 16-byte loads when the row start is 16-byte aligned; all of a thread's
        // loads in flight before its LDS writes
        constexpr int MR  = (H / 2 + NT) / NT; // pairs m = 0 .. H/2
        const bool    a16 = (reinterpret_cast<unsigned long>(cp) & 15) == 0;
        float2        u0[MR], u1[MR];
        if ( staged ) {
#pragma unroll
            for ( int r = 0; r < MR; r++ ) {
                const int m = threadIdx.x + NT * r;
                u0[r] = u1[r] = make_float2(0.f, 0.f);
                if ( a16 && 2 * m + 1 <= H ) {
                    const float4 f = load_cpx_pair(cp, m);
                    u0[r]          = make_float2(f.x, f.y);
                    u1[r]          = make_float2(f.z, f.w);
                }
                else if ( 2 * m <= H ) {
                    u0[r] = load_cpx(cp, 2 * m);
                    u1[r] = 2 * m + 1 <= H ? load_cpx(cp, 2 * m + 1) : make_float2(0.f, 0.f);
                }
            }
        }
        tws.store(s_twp, s_twn);
        if ( staged ) {
#pragma unroll
            for ( int r = 0; r < MR; r++ ) {
                const int m = threadIdx.x + NT * r;
                if ( 2 * m <= H )
                    lds[dec_slot<P, Q>(2 * m)] = u0[r];
                if ( 2 * m + 1 <= H )
                    lds[dec_slot<P, Q>(2 * m + 1)] = u1[r];
            }
        }
    }
    __syncthreads( );
    FFT fft;
    fft.init_pt(t, s_twp);
    // X[k] for k <= N/2, conj(X[N - k]) above
    float2 v[E];
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const int  k  = Q * (t + T * i) + q;
        const bool lo = k <= H;
        const int  kk = lo ? k : N - k;
        float2     z  = staged ? lds[dec_slot<P, Q>(kk)] : load_cpx(cp, (long)kk * im.es);
        if ( ! lo )
            z.y = -z.y;
        v[i] = z;
    }
    __syncthreads( );
    float2* wl = lds + q * FFT::SLOTS;
    fft.template run<+1>(v, wl, t);
    // this wave's real share of every output sample, into its own LDS region:
    // Re(U_q[(o + j) mod P] conj(W_N^{(o + j) q})), W_N^{(o+t+64i) q} = W_N^{o q} W_N^{t q} W_N^{64 i q}
    float*       red   = reinterpret_cast<float*>(wl);
    const int    omask = ((o % N) + N) % N;
    const float2 s     = cmul(lane_twiddle(s_twn, t, q), twm[N + ((omask * q) & (N - 1))]);
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const float2 w = cmul(s, twm[N + ((T * i * q) & (N - 1))]);
        const float2 u = v[(i + RS) % E];
        red[t + T * i] = u.x * w.x + u.y * w.y;
    }
    __syncthreads( );
    auto*              op   = as_global(out) + om.base(row);
    unsigned long long best = 0;
    for ( int j = threadIdx.x; j < M; j += NT ) {
        float s2 = 0.f;
#pragma unroll
        for ( int w = 0; w < Q; w++ )
            s2 += reinterpret_cast<const float*>(lds + w * FFT::SLOTS)[j];
        s2 *= post_scale;
        store_real(op, j * om.es, s2);
        if ( row_keys ) {
            // the value as stored (rounded when the output is half precision)
            const float stored = convert_real<TOut>(s2);
            const auto  k      = argmax_key(stored, (unsigned)(row * M + j));
            best               = k > best ? k : best;
        }
    }
    if ( row_keys ) {
        best = block_max_key(best);
        if ( threadIdx.x == 0 )
            as_global(row_keys)[blockIdx.x] = best;
    }
}

} // namespace fastfft_amd
