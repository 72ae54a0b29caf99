// Wave-level FFT building blocks for the hot kernels.
//
// One wavefront (64 lanes) computes a P-point transform with E = P/64 values
// per lane in the strided layout (register i of lane t holds element t + 64*i).
// Compared with the generic engine of fft_core.hpp:
//   - the Stockham twiddles of every pass are computed once per wave from the
//     FP64-generated table and kept in registers;
//   - LDS exchanges are ordered within the wave by wavefront-scope fences (LDS
//     operations of one wave execute in order), so no workgroup barrier and no
//     wait are needed between the scatter and the gather;
//   - fixed twiddles W_n^k whose index is known at compile time are evaluated
//     by constexpr double-precision arithmetic and applied as constants, with
//     the multiplications by 1, -/+i and (+-1 +-i)/sqrt2 reduced to adds.
#pragma once

#include "fft_core.hpp"

namespace fastfft_amd {

// ---------------------------------------------------------------- compile-time trig
namespace ct {
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double sin_series(double x) {
    double x2 = x * x, term = x, sum = x;
    for ( int n = 1; n < 14; n++ ) {
        term *= -x2 / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}

constexpr double cos_series(double x) {
    double x2 = x * x, term = 1.0, sum = 1.0;
    for ( int n = 1; n < 14; n++ ) {
        term *= -x2 / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}

// cos and sin of 2*pi*k/n with exact reduction to [0, pi/4] (n a multiple of 8 or any n).
struct CS {
    double c = 0.0, s = 0.0;
};

constexpr CS cis(long k, long n) {
    k %= n;
    if ( k < 0 )
        k += n;
    // quadrant by 4k/n using exact integer arithmetic on 8k vs n multiples
    long  q8 = (8 * k) / n; // octant 0..7
    // reduce to first quadrant: angle = 2 pi k / n = quad * pi/2 + rem
    long  quad = (4 * k) / n;
    // remainder angle in radians, computed from the exact rational remainder
    double rem = 2.0 * kPi * (double)(4 * k - quad * n) / (4.0 * (double)n);
    double c = 0.0, s = 0.0;
    if ( q8 % 2 == 0 ) {
        c = cos_series(rem);
        s = sin_series(rem);
    }
    else {
        double comp = kPi / 2.0 - rem;
        c           = sin_series(comp);
        s           = cos_series(comp);
    }
    // rotate by quad * 90 degrees
    for ( long r = 0; r < quad; r++ ) {
        double nc = -s, ns = c;
        c         = nc;
        s         = ns;
    }
    return CS{c, s};
}
} // namespace ct

// a * W_n^k with W_n^k = exp(DIR * 2 pi i k / n) (DIR = -1 forward), k and n compile-time.
template <long K, long NN, int DIR>
__device__ __forceinline__ float2 mul_const_tw(float2 a) {
    constexpr long k = ((K % NN) + NN) % NN;
    if constexpr ( k == 0 ) {
        return a;
    }
    else if constexpr ( (8 * k) % NN == 0 ) {
        // multiples of pi/4: reuse the exact radix-8 constant paths of const_twiddle
        constexpr int k8 = (int)(8 * k / NN);
        if constexpr ( k8 == 4 ) {
            return make_float2(-a.x, -a.y);
        }
        else if constexpr ( k8 < 4 ) {
            return const_twiddle<8, k8, DIR>(a);
        }
        else {
            float2 r = const_twiddle<8, k8 - 4, DIR>(a);
            return make_float2(-r.x, -r.y);
        }
    }
    else {
        constexpr ct::CS cs = ct::cis(k, NN);
        constexpr float  c  = (float)cs.c;
        constexpr float  s  = DIR < 0 ? (float)-cs.s : (float)cs.s;
        return make_float2(fmaf(a.x, c, -a.y * s), fmaf(a.x, s, a.y * c));
    }
}

// ---------------------------------------------------------------- wave synchronisation
// Orders this wave's LDS accesses without waiting: all 64 lanes issue every
// LDS instruction together and a wave's LDS instructions complete in order.
__device__ __forceinline__ void wave_lds_fence( ) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier( );
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------- LDS layout of the exchanges
// Slot of element k in a wave's exchange buffer of P complex values: an XOR
// swizzle within each group of 32 slots (256 bytes) that makes every Stockham
// scatter and strided gather of the wave transforms free of LDS bank conflicts
// (checked for 32 and 64 banks of 4 bytes); sizes without a tuned swizzle use
// one pad slot per 32 values.
template <int P, int E>
struct ExchangeLayout {
    static constexpr int SLOTS = P + P / 32;
    __device__ __forceinline__ static int slot(int k) { return k + (k >> 5); }
};
template <>
struct ExchangeLayout<256, 4> {
    static constexpr int SLOTS = 256;
    __device__ __forceinline__ static int slot(int k) { return k ^ ((k >> 3) & 3) ^ (((k >> 4) & 7) << 2); }
};
template <>
struct ExchangeLayout<512, 8> {
    static constexpr int SLOTS = 512;
    __device__ __forceinline__ static int slot(int k) { return k ^ ((k >> 3) & 3) ^ (((k >> 5) & 7) << 2); }
};
template <>
struct ExchangeLayout<1024, 16> {
    static constexpr int SLOTS = 1024;
    __device__ __forceinline__ static int slot(int k) { return k ^ ((k >> 4) & 3) ^ (((k >> 6) & 7) << 2); }
};

// ---------------------------------------------------------------- wave Stockham
// Radix of pass p (stride Ns = E^p ...): E for all passes but a smaller last one.
template <int P, int E>
struct WavePasses {
    static constexpr int count(int ns) { return ns >= P ? 0 : 1 + count(ns * ((P / ns) < E ? (P / ns) : E)); }
    static constexpr int N_PASS = count(1);
};

template <int P>
struct WaveFFT {
    static constexpr int T      = 64;
    static constexpr int E      = P / T;
    static constexpr int NPASS  = WavePasses<P, E>::N_PASS;
    using Layout                = ExchangeLayout<P, E>;
    static constexpr int SLOTS  = Layout::SLOTS; // LDS complex values per transform
    static_assert(P % 64 == 0 && E >= 2 && E <= 16, "wave FFT needs 2..16 values per lane");
    // twiddles of passes 1..NPASS-1: per pass at most B*(R-1) < E values
    float2 w[NPASS > 1 ? NPASS - 1 : 1][E];

    // tab[k] = W_TN^k for k < TN (TN a multiple of P)
    template <int Ns, int PI, int TN>
    __device__ __forceinline__ void init_pass(int t, const float2* __restrict__ tab) {
        if constexpr ( Ns < P ) {
            constexpr int R = (P / Ns) < E ? (P / Ns) : E;
            constexpr int B = E / R;
            if constexpr ( Ns > 1 ) {
#pragma unroll
                for ( int m = 0; m < B; m++ ) {
                    const int j = (t + T * m) & (Ns - 1);
#pragma unroll
                    for ( int r = 1; r < R; r++ )
                        w[PI - 1][m * (R - 1) + r - 1] = tab[j * r * (TN / (Ns * R))];
                }
            }
            init_pass<Ns * R, PI + 1, TN>(t, tab);
        }
    }

    // Stockham twiddles from the global table (W_kTwiddleN^k)
    __device__ __forceinline__ void init(int t, const float2* __restrict__ tw) { init_pass<1, 0, kTwiddleN>(t, tw); }

    // Stockham twiddles from a table of W_P^k, k < P (for example a copy in LDS)
    __device__ __forceinline__ void init_p(int t, const float2* tab) { init_pass<1, 0, P>(t, tab); }

    // ---- pass table: the Stockham twiddles of passes 1.. laid out pass by pass,
    // [r-1][j] within a pass (j = the lane's index mod Ns), so that the reads of
    // one (pass, r) by the 64 lanes are consecutive LDS slots (no bank conflicts).
    static constexpr int pass_table_size( ) {
        int n = 0;
        for ( int ns = 1; ns < P; ) {
            const int R = (P / ns) < E ? (P / ns) : E;
            if ( ns > 1 )
                n += ns * (R - 1);
            ns *= R;
        }
        return n;
    }
    static constexpr int PT = pass_table_size( );
    static_assert(PT <= P, "the pass table fits a P-entry buffer");

    // Index into the multi-resolution twiddle table (twm[n + m] = W_n^m) of
    // pass-table entry g < PT: entry k of the pass with span Ns and radix R is
    // W_{Ns R}^{(k mod Ns)(1 + k / Ns)}.
    template <int Ns = 1, int OFF = 0>
    __device__ __forceinline__ static int pass_table_index(int g) {
        if constexpr ( Ns < P ) {
            constexpr int R  = (P / Ns) < E ? (P / Ns) : E;
            constexpr int SZ = Ns > 1 ? Ns * (R - 1) : 0;
            if ( SZ > 0 && g < OFF + SZ ) {
                const int k = g - OFF;
                return Ns * R + (k & (Ns - 1)) * (1 + k / Ns);
            }
            return pass_table_index<Ns * R, OFF + SZ>(g);
        }
        else {
            return 0;
        }
    }

    // The pass table pt[0 .. PT) is filled by a workgroup of STRIDE threads,
    // thread `first` taking entries first + STRIDE r: load_pass_table issues
    // all of a thread's loads, store_pass_table writes them, so that callers
    // can issue their other loads in between.
    template <int STRIDE>
    static constexpr int pass_table_rows = PT > 0 ? (PT + STRIDE - 1) / STRIDE : 1;

    template <int STRIDE>
    __device__ __forceinline__ static void load_pass_table(float2 (&v)[pass_table_rows<STRIDE>],
                                                           const float2* __restrict__ twm, int first) {
#pragma unroll
        for ( int r = 0; r < pass_table_rows<STRIDE>; r++ ) {
            const int g = first + STRIDE * r;
            v[r]        = g < PT ? load_cpx(as_global(twm), pass_table_index(g)) : make_float2(0.f, 0.f);
        }
    }

    template <int STRIDE>
    __device__ __forceinline__ static void store_pass_table(float2* pt, const float2 (&v)[pass_table_rows<STRIDE>], int first) {
#pragma unroll
        for ( int r = 0; r < pass_table_rows<STRIDE>; r++ ) {
            const int g = first + STRIDE * r;
            if ( g < PT )
                pt[g] = v[r];
        }
    }

    template <int STRIDE>
    __device__ __forceinline__ static void fill_pass_table(float2* pt, const float2* __restrict__ twm, int first) {
        float2 v[pass_table_rows<STRIDE>];
        load_pass_table<STRIDE>(v, twm, first);
        store_pass_table<STRIDE>(pt, v, first);
    }

    template <int Ns, int PI, int OFF>
    __device__ __forceinline__ void init_pass_pt(int t, const float2* pt) {
        if constexpr ( Ns < P ) {
            constexpr int R = (P / Ns) < E ? (P / Ns) : E;
            constexpr int B = E / R;
            if constexpr ( Ns > 1 ) {
#pragma unroll
                for ( int m = 0; m < B; m++ ) {
                    const int j = (t + T * m) & (Ns - 1);
#pragma unroll
                    for ( int r = 1; r < R; r++ )
                        w[PI - 1][m * (R - 1) + r - 1] = pt[OFF + (r - 1) * Ns + j];
                }
            }
            init_pass_pt<Ns * R, PI + 1, OFF + (Ns > 1 ? Ns * (R - 1) : 0)>(t, pt);
        }
    }

    // Stockham twiddles from the pass table (fill_pass_table), e.g. in LDS
    __device__ __forceinline__ void init_pt(int t, const float2* pt) { init_pass_pt<1, 0, 0>(t, pt); }

    // One Stockham pass applied to NB independent transforms held by the wave;
    // each transform exchanges through its own LDS region of SLOTS values, so the
    // NB exchanges of a pass share one LDS round trip.
    // NZ: registers NZ .. E - 1 are known to hold zeros on entry (the first pass then
    // drops their terms at compile time).
    template <int Ns, int PI, int DIR, int NB, int NZ = E>
    __device__ __forceinline__ void pass(float2 (&v)[NB][E], float2* __restrict__ lds, int t) const {
        if constexpr ( Ns < P ) {
            constexpr int  R    = (P / Ns) < E ? (P / Ns) : E;
            constexpr int  B    = E / R;
            constexpr bool LAST = (Ns * R == P);
#pragma unroll
            for ( int nb = 0; nb < NB; nb++ ) {
#pragma unroll
                for ( int m = 0; m < B; m++ ) {
                    float2 u[R];
#pragma unroll
                    for ( int r = 0; r < R; r++ )
                        u[r] = (Ns == 1 && m + B * r >= NZ) ? make_float2(0.f, 0.f) : v[nb][m + B * r];
                    if constexpr ( Ns > 1 ) {
#pragma unroll
                        for ( int r = 1; r < R; r++ ) {
                            float2 tw = w[PI - 1][m * (R - 1) + r - 1];
                            if constexpr ( DIR > 0 )
                                tw.y = -tw.y;
                            u[r] = cmul(u[r], tw);
                        }
                    }
                    SmallDFT<R, DIR>::run(u);
                    if constexpr ( LAST ) {
#pragma unroll
                        for ( int r = 0; r < R; r++ )
                            v[nb][m + B * r] = u[r];
                    }
                    else {
                        const int b    = t + T * m;
                        const int j    = b & (Ns - 1);
                        const int base = (b - j) * R + j;
#pragma unroll
                        for ( int r = 0; r < R; r++ )
                            lds[nb * SLOTS + Layout::slot(base + r * Ns)] = u[r];
                    }
                }
            }
            if constexpr ( ! LAST ) {
                wave_lds_fence( );
#pragma unroll
                for ( int nb = 0; nb < NB; nb++ )
#pragma unroll
                    for ( int i = 0; i < E; i++ )
                        v[nb][i] = lds[nb * SLOTS + Layout::slot(t + T * i)];
                wave_lds_fence( );
            }
            pass<Ns * R, PI + 1, DIR, NB>(v, lds, t);
        }
    }

    // NB transforms in lockstep; lds holds NB * SLOTS values.
    template <int DIR, int NB>
    __device__ __forceinline__ void run_batch(float2 (&v)[NB][E], float2* __restrict__ lds, int t) const {
        pass<1, 0, DIR, NB>(v, lds, t);
    }

    // In-place transform of one wave-held P-point sequence.  lds: SLOTS values.  NZ < E:
    // registers NZ .. E - 1 hold zeros (an input of at most 64 NZ non-zero values).
    template <int DIR, int NZ = E>
    __device__ __forceinline__ void run(float2 (&v)[E], float2* __restrict__ lds, int t) const {
        pass<1, 0, DIR, 1, NZ>(reinterpret_cast<float2(&)[1][E]>(v), lds, t);
    }
};

} // namespace fastfft_amd
