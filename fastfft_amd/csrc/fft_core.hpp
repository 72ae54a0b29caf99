// Device FFT engine for gfx950 (CDNA4, wave64).
//
// Block FFTs are Stockham auto-sort transforms held in registers:
//   - a group of T threads computes one P-point transform, each thread holding
//     E = P/T complex values in the "strided" layout: register i of thread t
//     holds element t + T*i;
//   - every pass is a radix-R DFT on values a thread already owns (R <= E),
//     preceded by the Stockham twiddle W_{Ns*R}^{(b mod Ns)*r};
//   - between passes the values are exchanged through LDS (write to the
//     Stockham destination index, read back in the strided layout); the last
//     pass lands directly in natural order in registers, so a P = E^k * r
//     transform needs (number of passes - 1) exchanges.
// Size changes use the Sorensen-Burrus transform decomposition the reference
// uses (N = P*Q; include/detail/fft_properties.h:6-14):
//   input pruned  X[Q*k' + q] = FFT_P( x[n] * W_N^{n q} )[k'],       n < P
//   output pruned y[o + j]    = sum_q W_N^{-j q} IFFT_P( Z[Q*s+q] * W_N^{-o(Qs+q)} )[j]
// Twiddles come from one table of W_NT^k = exp(-2 pi i k / NT) computed in FP64
// and rounded to FP32 (twiddle.hip).  All arithmetic is FP32.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

namespace fastfft_amd {

// Twiddle table length: the largest supported transform size.
constexpr int kTwiddleN = 8192;

// ---------------------------------------------------------------- complex math
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -a.x * b.y));
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (DIR = -1) or +i (DIR = +1)
template <int DIR>
__device__ __forceinline__ float2 mul_dir_i(float2 a) {
    if constexpr ( DIR < 0 )
        return make_float2(a.y, -a.x);
    else
        return make_float2(-a.y, a.x);
}

// W_n^k for the requested direction from the FP64-generated table:
// DIR=-1 -> exp(-2 pi i k/n), DIR=+1 -> exp(+2 pi i k/n).  idx = k * (kTwiddleN/n) < kTwiddleN.
template <int DIR>
__device__ __forceinline__ float2 twiddle(const float2* __restrict__ tw, int idx) {
    float2 w = tw[idx];
    if constexpr ( DIR > 0 )
        w.y = -w.y;
    return w;
}

// Multi-resolution part of the table, after the kTwiddleN base entries:
// multi_twiddles(tw)[n + m] = W_n^m = exp(-2 pi i m / n) for every power of two
// n <= kTwiddleN and m < n, so that the twiddles of any one transform size are
// contiguous (coalesced loads).
__device__ __forceinline__ const float2* multi_twiddles(const float2* tw) { return tw + kTwiddleN; }

// ---------------------------------------------------------------- argmax keys
// 64-bit key ordered like (value, -index): the high word is the IEEE value with
// its order made unsigned (sign bit flipped for non-negatives, all bits for
// negatives), the low word 0xffffffff - index, so that the largest key is the
// first maximum in index order.
__device__ __forceinline__ unsigned long long argmax_key(float v, unsigned idx) {
    const unsigned u  = __float_as_uint(v);
    const unsigned hi = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return ((unsigned long long)hi << 32) | (0xffffffffu - idx);
}

__device__ __forceinline__ float argmax_key_value(unsigned long long k) {
    const unsigned hi = (unsigned)(k >> 32);
    return __uint_as_float((hi & 0x80000000u) ? (hi & 0x7fffffffu) : ~hi);
}

__device__ __forceinline__ unsigned argmax_key_index(unsigned long long k) { return 0xffffffffu - (unsigned)k; }

// ---------------------------------------------------------------- type conversion
template <typename T>
__device__ __forceinline__ float load_real(const T* p, long i);
template <>
__device__ __forceinline__ float load_real<float>(const float* p, long i) { return p[i]; }
template <>
__device__ __forceinline__ float load_real<__half>(const __half* p, long i) { return __half2float(p[i]); }

// v rounded to T and back, as a store of v to T and a load would give
template <typename T>
__device__ __forceinline__ float convert_real(float v);
template <>
__device__ __forceinline__ float convert_real<float>(float v) { return v; }
template <>
__device__ __forceinline__ float convert_real<__half>(float v) { return __half2float(__float2half_rn(v)); }

template <typename T>
__device__ __forceinline__ void store_real(T* p, long i, float v);
template <>
__device__ __forceinline__ void store_real<float>(float* p, long i, float v) { p[i] = v; }
template <>
__device__ __forceinline__ void store_real<__half>(__half* p, long i, float v) { p[i] = __float2half_rn(v); }

template <typename T>
__device__ __forceinline__ float2 load_cpx(const T* p, long i);
template <>
__device__ __forceinline__ float2 load_cpx<float2>(const float2* p, long i) { return p[i]; }
template <>
__device__ __forceinline__ float2 load_cpx<__half2>(const __half2* p, long i) { return __half22float2(p[i]); }

template <typename T>
__device__ __forceinline__ void store_cpx(T* p, long i, float2 v);
template <>
__device__ __forceinline__ void store_cpx<float2>(float2* p, long i, float2 v) { p[i] = v; }
template <>
__device__ __forceinline__ void store_cpx<__half2>(__half2* p, long i, float2 v) { p[i] = __float22half2_rn(v); }

// ---------------------------------------------------------------- global-memory pointers
// A pointer that a kernel reads from memory (a launch descriptor) is generic to
// the compiler, which then accesses it with flat instructions.  Those count on
// the LDS counter too, so every LDS wait of the kernel would also wait for the
// outstanding flat loads and stores.  Accesses through global_t pointers are
// global loads and stores.
template <typename T>
using global_t = __attribute__((address_space(1))) T;

template <typename T>
__device__ __forceinline__ global_t<T>* as_global(T* p) {
    return (global_t<T>*)p;
}

// Launch descriptors, read-only while a kernel runs and read at wave-uniform
// addresses, are read through the constant address space: scalar loads into
// scalar registers.
template <typename T>
using const_t = const __attribute__((address_space(4))) T;

template <typename T>
__device__ __forceinline__ const_t<T>* as_const(const T* p) {
    return (const_t<T>*)p;
}

// (half precision values move as their bit patterns: the half types do not
// copy from an address-space-qualified object)
__device__ __forceinline__ float load_real(const global_t<float>* p, long i) { return p[i]; }
__device__ __forceinline__ float load_real(const global_t<__half>* p, long i) {
    return __half2float(__ushort_as_half(reinterpret_cast<const global_t<unsigned short>*>(p)[i]));
}
__device__ __forceinline__ void store_real(global_t<float>* p, long i, float v) { p[i] = v; }
__device__ __forceinline__ void store_real(global_t<__half>* p, long i, float v) {
    reinterpret_cast<global_t<unsigned short>*>(p)[i] = __half_as_ushort(__float2half_rn(v));
}
// float2 / float4 move as native vectors (the HIP vector classes do not copy
// from an address-space-qualified object either)
typedef float gvec2 __attribute__((ext_vector_type(2)));
typedef float gvec4 __attribute__((ext_vector_type(4)));
// values i and i + 1 of a real line as one load (i even, the line 2-element aligned)
__device__ __forceinline__ float2 load_real_pair(const global_t<float>* p, long i) {
    const gvec2 v = *reinterpret_cast<const global_t<gvec2>*>(p + i);
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float2 load_real_pair(const global_t<__half>* p, long i) {
    const unsigned u = *reinterpret_cast<const global_t<unsigned>*>(p + i);
    return make_float2(__half2float(__ushort_as_half((unsigned short)(u & 0xffffu))),
                       __half2float(__ushort_as_half((unsigned short)(u >> 16))));
}
// values i and i + 1 of a real line as one store (i even, the line 2-element aligned)
__device__ __forceinline__ void store_real_pair(global_t<float>* p, long i, float a, float b) {
    *reinterpret_cast<global_t<gvec2>*>(p + i) = gvec2{a, b};
}
__device__ __forceinline__ void store_real_pair(global_t<__half>* p, long i, float a, float b) {
    const unsigned lo = __half_as_ushort(__float2half_rn(a)), hi = __half_as_ushort(__float2half_rn(b));
    *reinterpret_cast<global_t<unsigned>*>(p + i) = lo | (hi << 16);
}
__device__ __forceinline__ float2 load_cpx(const global_t<float2>* p, long i) {
    const gvec2 v = reinterpret_cast<const global_t<gvec2>*>(p)[i];
    return make_float2(v.x, v.y);
}
// elements 2 m and 2 m + 1 of a 16-byte aligned line, as one 16-byte access
__device__ __forceinline__ float4 load_cpx_pair(const global_t<float2>* p, long m) {
    const gvec4 v = reinterpret_cast<const global_t<gvec4>*>(p)[m];
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void store_cpx_pair(global_t<float2>* p, long m, float2 a, float2 b) {
    reinterpret_cast<global_t<gvec4>*>(p)[m] = gvec4{a.x, a.y, b.x, b.y};
}
__device__ __forceinline__ float2 load_cpx(const global_t<__half2>* p, long i) {
    const unsigned u = reinterpret_cast<const global_t<unsigned>*>(p)[i];
    return make_float2(__half2float(__ushort_as_half((unsigned short)(u & 0xffffu))),
                       __half2float(__ushort_as_half((unsigned short)(u >> 16))));
}
__device__ __forceinline__ void store_cpx(global_t<float2>* p, long i, float2 v) {
    reinterpret_cast<global_t<gvec2>*>(p)[i] = gvec2{v.x, v.y};
}
__device__ __forceinline__ void store_cpx(global_t<__half2>* p, long i, float2 v) {
    const unsigned lo = __half_as_ushort(__float2half_rn(v.x)), hi = __half_as_ushort(__float2half_rn(v.y));
    reinterpret_cast<global_t<unsigned>*>(p)[i] = lo | (hi << 16);
}

// ---------------------------------------------------------------- small DFTs in registers
// In-place radix-R DFT (R = 2,4,8,16) by compile-time radix-2 decimation in time.
// Constant twiddles W_R^k are exact decimal expansions; multiplications by 1 and
// -/+i are elided.
template <int R, int DIR>
struct SmallDFT;

template <int DIR>
struct SmallDFT<1, DIR> {
    __device__ __forceinline__ static void run(float2*) {}
};

template <int DIR>
struct SmallDFT<2, DIR> {
    __device__ __forceinline__ static void run(float2* u) {
        float2 a = u[0], b = u[1];
        u[0]     = cadd(a, b);
        u[1]     = csub(a, b);
    }
};

// multiply by W_R^k (direction DIR) with constant k, R <= 16
template <int R, int K, int DIR>
__device__ __forceinline__ float2 const_twiddle(float2 a) {
    constexpr int k8 = (K * 16) / R; // k in units of 2*pi/16
    static_assert((K * 16) % R == 0, "radix too large");
    if constexpr ( k8 == 0 ) {
        return a;
    }
    else if constexpr ( k8 == 4 ) {
        return mul_dir_i<DIR>(a);
    }
    else if constexpr ( k8 == 2 || k8 == 6 ) {
        // (1 -/+ i)/sqrt2 or (-1 -/+ i)/sqrt2 times a
        constexpr float h = 0.70710678118654752440f;
        float2          r;
        if constexpr ( k8 == 2 ) {
            // W^(2/16) = exp(DIR i pi/4) = h + DIR i h
            if constexpr ( DIR < 0 )
                r = make_float2((a.x + a.y) * h, (a.y - a.x) * h);
            else
                r = make_float2((a.x - a.y) * h, (a.y + a.x) * h);
        }
        else {
            // exp(DIR i 3pi/4) = -h + DIR i h
            if constexpr ( DIR < 0 )
                r = make_float2((a.y - a.x) * h, -(a.x + a.y) * h);
            else
                r = make_float2(-(a.x + a.y) * h, (a.x - a.y) * h);
        }
        return r;
    }
    else {
        // general 16th roots: k8 in {1,3,5,7}
        constexpr float c1 = 0.92387953251128675613f; // cos(pi/8)
        constexpr float s1 = 0.38268343236508977173f; // sin(pi/8)
        constexpr float cr = (k8 == 1) ? c1 : (k8 == 3) ? s1 : (k8 == 5) ? -s1 : -c1;
        constexpr float sr = (k8 == 1) ? s1 : (k8 == 3) ? c1 : (k8 == 5) ? c1 : s1;
        // exp(DIR i k8 pi/8) = cr + DIR i sr
        constexpr float wy = DIR < 0 ? -sr : sr;
        return make_float2(fmaf(a.x, cr, -a.y * wy), fmaf(a.x, wy, a.y * cr));
    }
}

template <int R, int DIR>
struct SmallDFT {
    template <int K>
    __device__ __forceinline__ static void combine(float2* e, float2* o, float2* u) {
        if constexpr ( K < R / 2 ) {
            float2 t     = const_twiddle<R, K, DIR>(o[K]);
            u[K]         = cadd(e[K], t);
            u[K + R / 2] = csub(e[K], t);
            combine<K + 1>(e, o, u);
        }
    }
    __device__ __forceinline__ static void run(float2* u) {
        float2 e[R / 2], o[R / 2];
#pragma unroll
        for ( int k = 0; k < R / 2; k++ ) {
            e[k] = u[2 * k];
            o[k] = u[2 * k + 1];
        }
        SmallDFT<R / 2, DIR>::run(e);
        SmallDFT<R / 2, DIR>::run(o);
        combine<0>(e, o, u);
    }
};

// ---------------------------------------------------------------- Stockham block FFT
// Group-wide synchronisation for the LDS exchange.  Groups never straddle a
// workgroup, and every thread of the workgroup runs the same pass sequence.
__device__ __forceinline__ void group_sync( ) { __syncthreads( ); }

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }
constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

// LDS slot of element index k: one pad slot every 32 complex values breaks the
// power-of-two strides of the Stockham scatter across the LDS banks.
__device__ __forceinline__ int lds_slot(int k) { return k + (k >> 5); }
constexpr int lds_slots(int P) { return P + P / 32; }

// One Stockham pass of radix R with stride Ns over a P-point transform.
template <int P, int E, int R, int Ns, int DIR>
__device__ __forceinline__ void stockham_pass(float2 (&v)[E], float2* __restrict__ lds, int t,
                                              const float2* __restrict__ tw) {
    constexpr int T    = P / E;
    constexpr int B    = E / R; // butterflies per thread
    constexpr bool LAST = (Ns * R == P);
#pragma unroll
    for ( int m = 0; m < B; m++ ) {
        const int b = t + T * m;
        const int j = b & (Ns - 1);
        float2    u[R];
#pragma unroll
        for ( int r = 0; r < R; r++ )
            u[r] = v[m + B * r];
        if constexpr ( Ns > 1 ) {
            const int step = j * (kTwiddleN / (Ns * R));
#pragma unroll
            for ( int r = 1; r < R; r++ )
                u[r] = cmul(u[r], twiddle<DIR>(tw, step * r));
        }
        SmallDFT<R, DIR>::run(u);
        if constexpr ( LAST ) {
#pragma unroll
            for ( int r = 0; r < R; r++ )
                v[m + B * r] = u[r];
        }
        else {
            const int base = (b - j) * R + j;
#pragma unroll
            for ( int r = 0; r < R; r++ )
                lds[lds_slot(base + r * Ns)] = u[r];
        }
    }
    if constexpr ( ! LAST ) {
        group_sync( );
#pragma unroll
        for ( int i = 0; i < E; i++ )
            v[i] = lds[lds_slot(t + T * i)];
        group_sync( );
    }
}

template <int P, int E, int Ns, int DIR>
__device__ __forceinline__ void stockham_passes(float2 (&v)[E], float2* __restrict__ lds, int t,
                                                const float2* __restrict__ tw) {
    if constexpr ( Ns < P ) {
        constexpr int R = (P / Ns) < E ? (P / Ns) : E;
        stockham_pass<P, E, R, Ns, DIR>(v, lds, t, tw);
        stockham_passes<P, E, Ns * R, DIR>(v, lds, t, tw);
    }
}

// Full P-point transform of the values in v (strided layout in, natural-order
// strided layout out).  lds must hold lds_slots(P) complex values for this group.
template <int P, int E, int DIR>
__device__ __forceinline__ void block_fft(float2 (&v)[E], float2* __restrict__ lds, int t,
                                          const float2* __restrict__ tw) {
    static_assert((P & (P - 1)) == 0 && P >= E && (E & (E - 1)) == 0, "power-of-two sizes only");
    stockham_passes<P, E, 1, DIR>(v, lds, t, tw);
}

// Threads-per-transform choice: E values per thread for a P-point transform.
template <int P>
struct FFTShape {
    static constexpr int E = P <= 8 ? P : (P <= 64 ? 8 : (P == 128 ? 8 : (P == 512 ? 8 : 16)));
    static constexpr int T = P / E;
};

} // namespace fastfft_amd
