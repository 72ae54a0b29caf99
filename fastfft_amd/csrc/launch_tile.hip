// Line transforms of 256 or 512 points, 16 lines per workgroup (tile_lines.hpp),
// for the passes of the 3-D transforms.
#include <cstdint>

#include "launch_common.hpp"
#include "tile_lines.hpp"

namespace fastfft_amd {

// 16-byte accesses to two neighbouring lines of a tile: even line offsets within
// the tile's rows and a 16-byte aligned buffer
static bool pair_map(const HostLineMap& m, const void* p) {
    return m.es % 2 == 0 && (m.n2 == 1 || m.s1 % 2 == 0) && (reinterpret_cast<uintptr_t>(p) & 15) == 0;
}

template <int N, int DIR, typename TIn, bool HERM, typename TOut, bool TI, bool TO>
static hipError_t go_tile(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st, int M, int o) {
    // bit 0: XCD-contiguous tiles, measured faster where a tile writes more rows than it reads (a
    // forward that pads) and slower for reading tiles; bit 1: 16-byte tile accesses where aligned.
    // FASTFFT_TILE_FLAGS overrides both.
    const int tf    = env_int("FASTFFT_TILE_FLAGS", (DIR < 0 && M > s.L) ? 3 : 2);
    int       flags = (TI || TO) ? (tf & 1) : 0;
    if ( (tf & 2) && (! TI || pair_map(s.im, io.src)) && (! TO || pair_map(s.om, io.dst)) )
        flags |= 2;
    launch_kernel<k_tile_lines<N, DIR, TIn, HERM, TOut, TI, TO>>(
            dim3((unsigned)(s.nlines / tile_workgroup_lines<TI, TO>)), dim3(512), 0, st, static_cast<const TIn*>(io.src),
            to_device(s.im), s.L, static_cast<TOut*>(io.dst), to_device(s.om), M, o, s.nlines, io.scale, tw, flags);
    return hipGetLastError( );
}

// lines of a tile next to each other (unit stride between lines, inside one
// group of a two-index map) with far-apart elements
static bool tile_map(const HostLineMap& m) {
    const bool adjacent = (m.n2 == 1 && m.s1 == 1) || (m.n2 > 1 && m.s2 == 1 && m.n2 % kTileLines == 0);
    return adjacent && m.es >= kTileLines;
}

static bool row_map(const HostLineMap& m) { return m.es == 1; }

bool tile_lines_cover(const Stage& s) {
    if ( ! (s.N == 256 || s.N == 512) || s.nlines <= 0 || s.nlines % kTileLines != 0 )
        return false;
    if ( s.kind == StageKind::inv_c2r )
        return tile_map(s.im) && row_map(s.om);
    if ( s.kind == StageKind::fwd_real ) // real rows in, transposed spectrum rows out
        return row_map(s.im) && tile_map(s.om) && s.nout <= s.N;
    if ( s.kind != StageKind::fwd_cpx && s.kind != StageKind::inv_cpx )
        return false;
    return (tile_map(s.im) && tile_map(s.om)) || (row_map(s.im) && row_map(s.om));
}

hipError_t launch_tile_lines(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    if ( env_int("FASTFFT_NO_TILE_LINES", 0) || ! tile_lines_cover(s) )
        return hipErrorNotSupported;
    const bool r2c = s.kind == StageKind::fwd_real;
    const bool fwd = s.kind == StageKind::fwd_cpx || r2c, c2r = s.kind == StageKind::inv_c2r;
    const int  M = fwd ? s.nout : s.M, o = fwd ? 0 : s.o;
    if ( M < 1 || M > s.N || io.src_real != r2c )
        return hipErrorNotSupported;
    auto go = [&](auto nc) -> hipError_t {
        constexpr int N = decltype(nc)::value;
        if ( r2c ) {
            if ( io.dst_real || io.dst_ct != CpxT::c64 )
                return hipErrorNotSupported;
            return io.src_rt == RealT::f32 ? go_tile<N, -1, float, false, float2, false, true>(s, io, tw, st, M, o)
                                           : go_tile<N, -1, __half, false, float2, false, true>(s, io, tw, st, M, o);
        }
        if ( c2r ) {
            if ( ! io.dst_real || io.src_ct != CpxT::c64 )
                return hipErrorNotSupported;
            return io.dst_rt == RealT::f32 ? go_tile<N, +1, float2, true, float, true, false>(s, io, tw, st, M, o)
                                           : go_tile<N, +1, float2, true, __half, true, false>(s, io, tw, st, M, o);
        }
        if ( io.dst_real || io.dst_ct != CpxT::c64 )
            return hipErrorNotSupported;
        const bool tiled = tile_map(s.im) && tile_map(s.om);
        auto       dir   = [&](auto dc) -> hipError_t {
            constexpr int D = decltype(dc)::value;
            if ( tiled )
                return io.src_ct == CpxT::c64 ? go_tile<N, D, float2, false, float2, true, true>(s, io, tw, st, M, o)
                                              : go_tile<N, D, __half2, false, float2, true, true>(s, io, tw, st, M, o);
            return io.src_ct == CpxT::c64 ? go_tile<N, D, float2, false, float2, false, false>(s, io, tw, st, M, o)
                                          : go_tile<N, D, __half2, false, float2, false, false>(s, io, tw, st, M, o);
        };
        return fwd ? dir(std::integral_constant<int, -1>{ }) : dir(std::integral_constant<int, 1>{ });
    };
    return s.N == 256 ? go(std::integral_constant<int, 256>{ }) : go(std::integral_constant<int, 512>{ });
}

} // namespace fastfft_amd
