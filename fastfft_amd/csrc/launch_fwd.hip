// Forward line stages: R2C x-pass (real input) and C2C y/z passes.
#include <cstdint>

#include "full_wave.hpp"
#include "launch_common.hpp"
#include "rows_packed.hpp"
#include "rows_wave.hpp"

namespace fastfft_amd {

template <int P, typename TIn, bool REAL, typename TOut>
static hipError_t go_fwd(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    dim3       grid, block;
    size_t     lds;
    hipError_t e = prepare_launch<P>(k_fwd<P, TIn, REAL, TOut>, s.nlines, grid, block, lds);
    if ( e != hipSuccess )
        return e;
    launch_kernel<k_fwd<P, TIn, REAL, TOut>>(grid, block, lds, st, static_cast<const TIn*>(io.src), to_device(s.im), s.L,
                                                        static_cast<TOut*>(io.dst), to_device(s.om), s.N, s.Q, s.nout,
                                                        s.nlines, io.scale, tw, twiddle_stride(s.N));
    return hipGetLastError( );
}

// R2C rows with P in {256, 512, 1024}: one block per row, one wave per sub-transform.
template <int P, int Q, typename TIn>
static hipError_t go_r2c_wave(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    launch_kernel<k_r2c_rows_wave<P, Q, TIn>>(dim3((unsigned)s.nlines), dim3(64 * Q), 0, st,
            static_cast<const TIn*>(io.src), to_device(s.im), s.L, static_cast<float2*>(io.dst), to_device(s.om), s.nout,
            s.nlines, io.scale, tw, nullptr);
    return hipGetLastError( );
}

template <int P, int Q, typename TIn>
static hipError_t go_r2c_wave_batch(const Stage& s, const RowBatchItem* items, int n, int max_rows, float pre,
                                    const float2* tw, hipStream_t st) {
    launch_kernel<k_r2c_rows_wave<P, Q, TIn>>(dim3((unsigned)max_rows, (unsigned)n), dim3(64 * Q), 0, st,
            nullptr, to_device(s.im), s.L, nullptr, to_device(s.om), s.nout, max_rows, pre, tw, items);
    return hipGetLastError( );
}

// R2C rows of 512 or 1024 real points, one wave per row (rows_packed.hpp); the
// batched form takes its per-template in, pitch, L and row count from items.
static bool r2c_packed_covers(const Stage& s) {
    return s.kind == StageKind::fwd_real && (s.N == 512 || s.N == 1024) && s.nout >= 1 && s.nout <= s.N / 2 + 1 &&
           s.L >= 0 && s.L <= s.N && ! env_int("FASTFFT_NO_PACKED_ROWS", 0);
}

template <int P, typename TIn>
static hipError_t go_r2c_packed(const Stage& s, const void* src, void* dst, int nrows, float pre, const float2* tw,
                                const RowBatchItem* items, int n, hipStream_t st) {
    const dim3 grid((unsigned)((nrows + kPackedRows - 1) / kPackedRows), (unsigned)n);
    launch_kernel<k_r2c_rows_packed<P, TIn>>(grid, dim3(64 * kPackedRows), 0, st, static_cast<const TIn*>(src), to_device(s.im), s.L,
                                                                      static_cast<float2*>(dst), to_device(s.om), s.nout,
                                                                      nrows, pre, tw, items);
    return hipGetLastError( );
}

static hipError_t launch_r2c_packed(const Stage& s, const void* src, void* dst, RealT rt, int nrows, float pre,
                                    const float2* tw, const RowBatchItem* items, int n, hipStream_t st) {
    if ( ! r2c_packed_covers(s) || nrows < 1 || n < 1 )
        return hipErrorNotSupported;
    if ( s.N == 512 )
        return rt == RealT::f32 ? go_r2c_packed<256, float>(s, src, dst, nrows, pre, tw, items, n, st)
                                : go_r2c_packed<256, __half>(s, src, dst, nrows, pre, tw, items, n, st);
    return rt == RealT::f32 ? go_r2c_packed<512, float>(s, src, dst, nrows, pre, tw, items, n, st)
                            : go_r2c_packed<512, __half>(s, src, dst, nrows, pre, tw, items, n, st);
}

// (P, Q) of an R2C row stage that k_r2c_rows_wave covers
static bool r2c_rows_shape(const Stage& s) {
    return (s.P == 256 || s.P == 512 || s.P == 1024) && (s.Q == 1 || s.Q == 2 || s.Q == 4 || s.Q == 8 || s.Q == 16) &&
           s.P * s.Q <= kTwiddleN;
}

// Column-major batch output: row r at r, element k at k * es (es even, every column
// holding whole groups of kCmRows rows).
static bool r2c_cm_output(const Stage& s) {
    return s.om.n2 == 1 && s.om.s1 == 1 && s.om.es > 1 && s.om.es % 2 == 0 &&
           s.om.es >= (s.nlines + kCmRows - 1) / kCmRows * kCmRows;
}

bool r2c_rows_batch_covers(const Stage& s) {
    if ( r2c_cm_output(s) )
        return r2c_packed_covers(s) && s.im.n2 == 1 && s.im.es == 1;
    // one template per grid row: single-index lines, unit-stride row-major output with even leading dimension
    return (r2c_rows_shape(s) || r2c_packed_covers(s)) && s.im.n2 == 1 && s.im.es == 1 && s.om.n2 == 1 && s.om.es == 1 &&
           s.om.s1 % 2 == 0;
}

template <int P, typename TIn>
static hipError_t go_r2c_packed_cm(const Stage& s, const RowBatchItem* items, int n, int max_rows, float pre,
                                   const float2* tw, hipStream_t st) {
    const dim3 grid((unsigned)((max_rows + kCmRows - 1) / kCmRows), (unsigned)n);
    // non-zero registers of a row's packed input: ceil(L / 128) for the longest template row
    constexpr int E  = P / 64;
    const int     nz = (s.L + 127) / 128;
    auto          go = [&](auto zc) {
        launch_kernel<k_r2c_rows_packed_cm<P, TIn, decltype(zc)::value>>(grid, dim3(64 * kCmWaves), 0, st, s.om.es, s.nout, pre, tw,
                                                                                               items);
        return hipGetLastError( );
    };
    if ( nz <= 1 )
        return go(std::integral_constant<int, 1>{ });
    if ( nz <= 2 )
        return go(std::integral_constant<int, 2>{ });
    if ( nz <= 3 && E > 3 )
        return go(std::integral_constant<int, (E > 3 ? 3 : E)>{ });
    if ( nz <= 4 && E > 4 )
        return go(std::integral_constant<int, (E > 4 ? 4 : E)>{ });
    return go(std::integral_constant<int, E>{ });
}

hipError_t launch_r2c_rows_batch(const Stage& s, const void* items, int n, int max_rows, RealT in_rt, float pre,
                                 const float2* tw, hipStream_t st) {
    if ( ! r2c_rows_batch_covers(s) )
        return hipErrorNotSupported;
    const auto* it = static_cast<const RowBatchItem*>(items);
    if ( r2c_cm_output(s) ) {
        if ( max_rows < 1 || n < 1 )
            return hipErrorNotSupported;
        if ( s.N == 512 )
            return in_rt == RealT::f32 ? go_r2c_packed_cm<256, float>(s, it, n, max_rows, pre, tw, st)
                                       : go_r2c_packed_cm<256, __half>(s, it, n, max_rows, pre, tw, st);
        return in_rt == RealT::f32 ? go_r2c_packed_cm<512, float>(s, it, n, max_rows, pre, tw, st)
                                   : go_r2c_packed_cm<512, __half>(s, it, n, max_rows, pre, tw, st);
    }
    {
        hipError_t e = launch_r2c_packed(s, nullptr, nullptr, in_rt, max_rows, pre, tw, it, n, st);
        if ( e != hipErrorNotSupported )
            return e;
    }
    auto by_p = [&](auto pc) -> hipError_t {
        constexpr int P  = decltype(pc)::value;
        auto          go = [&](auto qc) -> hipError_t {
            constexpr int Q = decltype(qc)::value;
            if constexpr ( P * Q > kTwiddleN )
                return hipErrorNotSupported;
            else
                return in_rt == RealT::f32 ? go_r2c_wave_batch<P, Q, float>(s, it, n, max_rows, pre, tw, st)
                                           : go_r2c_wave_batch<P, Q, __half>(s, it, n, max_rows, pre, tw, st);
        };
        switch ( s.Q ) {
            case 1: return go(std::integral_constant<int, 1>{ });
            case 2: return go(std::integral_constant<int, 2>{ });
            case 4: return go(std::integral_constant<int, 4>{ });
            case 8: return go(std::integral_constant<int, 8>{ });
            case 16: return go(std::integral_constant<int, 16>{ });
            default: return hipErrorNotSupported;
        }
    };
    if ( s.P == 256 )
        return by_p(std::integral_constant<int, 256>{ });
    if ( s.P == 512 )
        return by_p(std::integral_constant<int, 512>{ });
    return by_p(std::integral_constant<int, 1024>{ });
}

template <typename TIn>
static hipError_t try_r2c_wave(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    auto by_p = [&](auto pc) -> hipError_t {
        constexpr int P = decltype(pc)::value;
        switch ( s.Q ) {
            case 1: return go_r2c_wave<P, 1, TIn>(s, io, tw, st);
            case 2: return go_r2c_wave<P, 2, TIn>(s, io, tw, st);
            case 4: return go_r2c_wave<P, 4, TIn>(s, io, tw, st);
            case 8: return go_r2c_wave<P, 8, TIn>(s, io, tw, st);
            case 16:
                if constexpr ( P * 16 <= kTwiddleN )
                    return go_r2c_wave<P, 16, TIn>(s, io, tw, st);
                return hipErrorNotSupported;
            default: return hipErrorNotSupported;
        }
    };
    if ( s.P == 256 )
        return by_p(std::integral_constant<int, 256>{ });
    if ( s.P == 512 )
        return by_p(std::integral_constant<int, 512>{ });
    if ( s.P == 1024 )
        return by_p(std::integral_constant<int, 1024>{ });
    return hipErrorNotSupported;
}

// Full-length lines (L = N, 256 <= N <= 8192) with a float2 output: one
// workgroup per line, P = 512 (256 for N = 256) and Q = N / P.
template <int P, int Q>
static hipError_t go_full_wave(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    const dim3 grid((unsigned)s.nlines), block(64 * Q);
    if ( io.src_real ) {
        if ( io.src_rt == RealT::f32 )
            launch_kernel<k_fwd_full_wave<P, Q, float, true>>(grid, block, 0, st, static_cast<const float*>(io.src), to_device(s.im),
                                                                       static_cast<float2*>(io.dst), to_device(s.om),
                                                                       s.nout, s.nlines, io.scale, tw);
        else
            launch_kernel<k_fwd_full_wave<P, Q, __half, true>>(grid, block, 0, st, static_cast<const __half*>(io.src), to_device(s.im),
                                                                        static_cast<float2*>(io.dst), to_device(s.om),
                                                                        s.nout, s.nlines, io.scale, tw);
    }
    else {
        if ( io.src_ct == CpxT::c64 )
            launch_kernel<k_fwd_full_wave<P, Q, float2, false>>(grid, block, 0, st, static_cast<const float2*>(io.src), to_device(s.im),
                                                                         static_cast<float2*>(io.dst), to_device(s.om),
                                                                         s.nout, s.nlines, io.scale, tw);
        else
            launch_kernel<k_fwd_full_wave<P, Q, __half2, false>>(grid, block, 0, st, static_cast<const __half2*>(io.src),
                                                                          to_device(s.im), static_cast<float2*>(io.dst),
                                                                          to_device(s.om), s.nout, s.nlines, io.scale, tw);
    }
    return hipGetLastError( );
}

// Full-length real lines into a transposed spectrum (lines adjacent, even element
// step, 16-byte aligned output): two lines per workgroup as one complex line.
// PPW pairs per workgroup (FASTFFT_PAIR_GROUP = 1, 2 or 4 caps it; by default 4 for
// 4096-point rows, 2 for the other lengths where the pair count allows and a workgroup
// of 2 Q waves fits).
template <int P, int Q, int PPW>
static hipError_t go_full_wave_pair_g(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    const int  npairs = s.nlines / 2;
    const dim3 grid((unsigned)(npairs / PPW)), block(64 * Q * PPW);
    if ( io.src_rt == RealT::f32 )
        launch_kernel<k_fwd_full_wave_pair<P, Q, float, PPW>>(grid, block, 0, st, static_cast<const float*>(io.src),
                to_device(s.im), static_cast<float2*>(io.dst), to_device(s.om), s.nout, npairs, io.scale, tw);
    else
        launch_kernel<k_fwd_full_wave_pair<P, Q, __half, PPW>>(grid, block, 0, st, static_cast<const __half*>(io.src),
                to_device(s.im), static_cast<float2*>(io.dst), to_device(s.om), s.nout, npairs, io.scale, tw);
    return hipGetLastError( );
}

template <int P, int Q>
static hipError_t go_full_wave_pair(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    if constexpr ( Q <= 8 ) {
        if ( env_int("FASTFFT_PAIR_GROUP", 4) >= 2 && (s.nlines / 2) % 2 == 0 )
            return go_full_wave_pair_g<P, Q, 2>(s, io, tw, st);
    }
    return go_full_wave_pair_g<P, Q, 1>(s, io, tw, st);
}

static bool full_wave_pair_covers(const Stage& s, const StageIO& io) {
    return io.src_real && s.nlines >= 2 && s.nlines % 2 == 0 && s.om.n2 == 1 && s.om.s1 == 1 && s.om.es % 2 == 0 &&
           s.im.n2 == 1 && reinterpret_cast<std::uintptr_t>(io.dst) % 16 == 0 && s.N >= 1024 &&
           ! env_int("FASTFFT_NO_PAIR_LINES", 0);
}

static hipError_t try_full_wave(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    if ( s.L != s.N || io.dst_real || io.dst_ct != CpxT::c64 || s.nout > s.N )
        return hipErrorNotSupported;
    if ( full_wave_pair_covers(s, io) ) {
        switch ( s.N ) {
            case 1024: return go_full_wave_pair<512, 2>(s, io, tw, st);
            case 2048: return go_full_wave_pair<512, 4>(s, io, tw, st);
            case 4096:
                // four pairs of 4096-point rows per workgroup on 1024-point sub-transforms (Q = 4)
                if ( env_int("FASTFFT_PAIR_GROUP", 4) >= 4 && (s.nlines / 2) % 4 == 0 )
                    return go_full_wave_pair_g<1024, 4, 4>(s, io, tw, st);
                return go_full_wave_pair<512, 8>(s, io, tw, st);
            case 8192: return go_full_wave_pair<512, 16>(s, io, tw, st);
            default: break;
        }
    }
    switch ( s.N ) {
        case 256: return go_full_wave<256, 1>(s, io, tw, st);
        case 512: return go_full_wave<512, 1>(s, io, tw, st);
        case 1024: return go_full_wave<512, 2>(s, io, tw, st);
        case 2048: return go_full_wave<512, 4>(s, io, tw, st);
        case 4096: return go_full_wave<512, 8>(s, io, tw, st);
        case 8192: return go_full_wave<512, 16>(s, io, tw, st);
        default: return hipErrorNotSupported;
    }
}

hipError_t launch_fwd(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    // a unit-stride output row is written with 16-byte stores: even line offsets, aligned base
    const bool rows_ok = s.om.es != 1 || (s.om.s1 % 2 == 0 && s.om.s2 % 2 == 0 &&
                                          reinterpret_cast<std::uintptr_t>(io.dst) % 16 == 0);
    if ( io.src_real && ! io.dst_real && io.dst_ct == CpxT::c64 && ! env_int("FASTFFT_NO_WAVE_ROWS", 0) ) {
        hipError_t e = launch_r2c_packed(s, io.src, io.dst, io.src_rt, s.nlines, io.scale, tw, nullptr, 1, st);
        if ( e != hipErrorNotSupported )
            return e;
    }
    if ( rows_ok && ! env_int("FASTFFT_NO_WAVE_ROWS", 0) && ! env_int("FASTFFT_NO_FULL_LINES", 0) ) {
        hipError_t e = try_full_wave(s, io, tw, st);
        if ( e != hipErrorNotSupported )
            return e;
    }
    // C2C passes, and real rows whose spectra are stored transposed (a 3-D x-pass): 16 lines per
    // workgroup, transposed through LDS into 128-byte segments
    if ( (s.kind == StageKind::fwd_cpx || (s.kind == StageKind::fwd_real && s.om.es != 1)) &&
         ! env_int("FASTFFT_NO_WAVE_ROWS", 0) ) {
        hipError_t e = launch_tile_lines(s, io, tw, st);
        if ( e != hipErrorNotSupported )
            return e;
    }
    if ( io.src_real && ! io.dst_real && io.dst_ct == CpxT::c64 && rows_ok && ! env_int("FASTFFT_NO_WAVE_ROWS", 0) ) {
        hipError_t e = io.src_rt == RealT::f32 ? try_r2c_wave<float>(s, io, tw, st) : try_r2c_wave<__half>(s, io, tw, st);
        if ( e != hipErrorNotSupported )
            return e;
    }
    return dispatch_P(s.P, [&](auto pc) -> hipError_t {
        constexpr int P = decltype(pc)::value;
        if ( io.src_real ) {
            if ( io.dst_real || io.dst_ct != CpxT::c64 )
                return hipErrorInvalidValue;
            if ( io.src_rt == RealT::f32 )
                return go_fwd<P, float, true, float2>(s, io, tw, st);
            return go_fwd<P, __half, true, float2>(s, io, tw, st);
        }
        if ( io.src_ct != CpxT::c64 || io.dst_real )
            return hipErrorInvalidValue;
        if ( io.dst_ct == CpxT::c64 )
            return go_fwd<P, float2, false, float2>(s, io, tw, st);
        return go_fwd<P, float2, false, __half2>(s, io, tw, st);
    });
}

size_t stage_lds_bytes(int P) {
    size_t out = 0;
    (void)dispatch_P(P, [&](auto pc) -> hipError_t {
        out = sizeof(float2) * (size_t)LineShape<decltype(pc)::value>::LDS;
        return hipSuccess;
    });
    return out;
}

} // namespace fastfft_amd
