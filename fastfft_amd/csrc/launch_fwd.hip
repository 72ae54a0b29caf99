// Forward line stages: R2C x-pass (real input) and C2C y/z passes.
#include "launch_common.hpp"

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

hipError_t launch_fwd(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
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
