// Tiled z-pass line transforms (tile_lines.hpp) for the strided passes of the
// 3-D transforms.
#include "launch_common.hpp"
#include "tile_lines.hpp"

namespace fastfft_amd {

template <int N, int DIR, typename TIn>
static hipError_t go_tile(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st, int M, int o) {
    k_tile_lines<N, DIR, TIn><<<dim3((unsigned)(s.nlines / kTileLines)), dim3(512), 0, st>>>(
            static_cast<const TIn*>(io.src), to_device(s.im), s.L, static_cast<float2*>(io.dst), to_device(s.om), M, o,
            s.nlines, io.scale, tw);
    return hipGetLastError( );
}

bool tile_lines_cover(const Stage& s) {
    // adjacent lines (unit line stride), far-apart elements, whole tiles, N <= 512 (one wave per transform)
    return (s.N == 256 || s.N == 512) && s.im.n2 == 1 && s.om.n2 == 1 && s.im.s1 == 1 && s.om.s1 == 1 &&
           s.im.es >= kTileLines && s.om.es >= kTileLines && s.nlines % kTileLines == 0 && s.nlines > 0;
}

hipError_t launch_tile_lines(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    if ( env_int("FASTFFT_NO_TILE_LINES", 0) || ! tile_lines_cover(s) )
        return hipErrorNotSupported;
    const bool fwd = s.kind == StageKind::fwd_cpx, inv = s.kind == StageKind::inv_cpx;
    if ( ! (fwd || inv) || io.src_real || io.dst_real || io.dst_ct != CpxT::c64 )
        return hipErrorNotSupported;
    const int M = fwd ? s.nout : s.M, o = fwd ? 0 : s.o;
    if ( M < 1 || M > s.N )
        return hipErrorNotSupported;
    auto go = [&](auto nc) -> hipError_t {
        constexpr int N = decltype(nc)::value;
        if ( fwd )
            return io.src_ct == CpxT::c64 ? go_tile<N, -1, float2>(s, io, tw, st, M, o)
                                          : go_tile<N, -1, __half2>(s, io, tw, st, M, o);
        return io.src_ct == CpxT::c64 ? go_tile<N, +1, float2>(s, io, tw, st, M, o)
                                      : go_tile<N, +1, __half2>(s, io, tw, st, M, o);
    };
    return s.N == 256 ? go(std::integral_constant<int, 256>{ }) : go(std::integral_constant<int, 512>{ });
}

} // namespace fastfft_amd
