// Inverse line stages: C2C y/z passes (output pruned) and the C2R x-pass.
#include "launch_common.hpp"

namespace fastfft_amd {

template <int P, typename TIn, bool HERM, typename TOut, bool REAL_OUT>
static hipError_t go_inv(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    dim3       grid, block;
    size_t     lds;
    hipError_t e = prepare_launch<P>(k_inv<P, TIn, HERM, TOut, REAL_OUT>, s.nlines, grid, block, lds);
    if ( e != hipSuccess )
        return e;
    k_inv<P, TIn, HERM, TOut, REAL_OUT><<<grid, block, lds, st>>>(static_cast<const TIn*>(io.src), to_device(s.im),
                                                                  static_cast<TOut*>(io.dst), to_device(s.om), s.N, s.M,
                                                                  s.o, s.Q, s.nlines, io.scale, tw, twiddle_stride(s.N));
    return hipGetLastError( );
}

hipError_t launch_inv(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    return dispatch_P(s.P, [&](auto pc) -> hipError_t {
        constexpr int P = decltype(pc)::value;
        if ( s.kind == StageKind::inv_c2r ) {
            if ( io.src_real || io.src_ct != CpxT::c64 || ! io.dst_real )
                return hipErrorInvalidValue;
            if ( io.dst_rt == RealT::f32 )
                return go_inv<P, float2, true, float, true>(s, io, tw, st);
            return go_inv<P, float2, true, __half, true>(s, io, tw, st);
        }
        if ( io.src_real || io.dst_real || io.dst_ct != CpxT::c64 )
            return hipErrorInvalidValue;
        if ( io.src_ct == CpxT::c64 )
            return go_inv<P, float2, false, float2, false>(s, io, tw, st);
        return go_inv<P, __half2, false, float2, false>(s, io, tw, st);
    });
}

} // namespace fastfft_amd
