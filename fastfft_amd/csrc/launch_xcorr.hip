// Fused correlation pass of FwdImageInvFFT.
#include "launch_common.hpp"

namespace fastfft_amd {

template <int P, typename TImg>
static hipError_t go_xcorr(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    dim3       grid, block;
    size_t     lds;
    hipError_t e = prepare_launch<P>(k_xcorr<P, TImg>, s.nlines, grid, block, lds);
    if ( e != hipSuccess )
        return e;
    k_xcorr<P, TImg><<<grid, block, lds, st>>>(static_cast<const float2*>(io.src), to_device(s.im), s.L,
                                                static_cast<const TImg*>(io.img), to_device(s.gm),
                                                static_cast<float2*>(io.dst), to_device(s.om), s.N, s.M, s.o, s.Q,
                                                s.nlines, io.scale, tw, twiddle_stride(s.N));
    return hipGetLastError( );
}

hipError_t launch_xcorr(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    return dispatch_P(s.P, [&](auto pc) -> hipError_t {
        constexpr int P = decltype(pc)::value;
        if ( io.img_ct == CpxT::c64 )
            return go_xcorr<P, float2>(s, io, tw, st);
        return go_xcorr<P, __half2>(s, io, tw, st);
    });
}

} // namespace fastfft_amd
