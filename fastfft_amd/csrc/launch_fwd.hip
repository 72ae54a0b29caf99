// Forward line stages: R2C x-pass (real input) and C2C y/z passes.
#include <cstdint>

#include "launch_common.hpp"
#include "rows_wave.hpp"

namespace fastfft_amd {

template <int P, typename TIn, bool REAL, typename TOut>
static hipError_t go_fwd(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    dim3       grid, block;
    size_t     lds;
    hipError_t e = prepare_launch<P>(k_fwd<P, TIn, REAL, TOut>, s.nlines, grid, block, lds);
    if ( e != hipSuccess )
        return e;
    k_fwd<P, TIn, REAL, TOut><<<grid, block, lds, st>>>(static_cast<const TIn*>(io.src), to_device(s.im), s.L,
                                                        static_cast<TOut*>(io.dst), to_device(s.om), s.N, s.Q, s.nout,
                                                        s.nlines, io.scale, tw, twiddle_stride(s.N));
    return hipGetLastError( );
}

// R2C rows with P in {256, 512, 1024}: one block per row, one wave per sub-transform.
template <int P, int Q, typename TIn>
static hipError_t go_r2c_wave(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    k_r2c_rows_wave<P, Q, TIn><<<dim3((unsigned)s.nlines), dim3(64 * Q), 0, st>>>(
            static_cast<const TIn*>(io.src), to_device(s.im), s.L, static_cast<float2*>(io.dst), to_device(s.om), s.nout,
            s.nlines, io.scale, tw);
    return hipGetLastError( );
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

hipError_t launch_fwd(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    // a unit-stride output row is written with 16-byte stores: even line offsets, aligned base
    const bool rows_ok = s.om.es != 1 || (s.om.s1 % 2 == 0 && s.om.s2 % 2 == 0 &&
                                          reinterpret_cast<std::uintptr_t>(io.dst) % 16 == 0);
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
