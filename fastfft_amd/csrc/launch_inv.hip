// Inverse line stages: C2C y/z passes (output pruned) and the C2R x-pass.
#include "full_wave.hpp"
#include "launch_common.hpp"
#include "rows_packed.hpp"
#include "rows_wave.hpp"
#include "window_rows.hpp"

namespace fastfft_amd {

template <int P, typename TIn, bool HERM, typename TOut, bool REAL_OUT>
static hipError_t go_inv(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    dim3       grid, block;
    size_t     lds;
    hipError_t e = prepare_launch<P>(k_inv<P, TIn, HERM, TOut, REAL_OUT>, s.nlines, grid, block, lds);
    if ( e != hipSuccess )
        return e;
    launch_kernel<k_inv<P, TIn, HERM, TOut, REAL_OUT>>(grid, block, lds, st, static_cast<const TIn*>(io.src), to_device(s.im),
                                                                  static_cast<TOut*>(io.dst), to_device(s.om), s.N, s.M,
                                                                  s.o, s.Q, s.nlines, io.scale, tw, twiddle_stride(s.N));
    return hipGetLastError( );
}

// C2R rows of 512 or 1024 real points, any window of M <= N outputs, one wave
// per row (rows_packed.hpp); keys: one argmax key per row, or null.
static bool c2r_packed_covers(const Stage& s) {
    return s.kind == StageKind::inv_c2r && (s.N == 512 || s.N == 1024) && s.M >= 1 && s.M <= s.N &&
           ! env_int("FASTFFT_NO_PACKED_ROWS", 0);
}

template <int P, typename TOut>
static hipError_t go_c2r_packed(const Stage& s, const void* src, void* dst, float post, const float2* tw,
                                unsigned long long* keys, const RowBatchItem* items, int n, hipStream_t st) {
    const dim3 grid((unsigned)((s.nlines + kPackedRows - 1) / kPackedRows), (unsigned)n);
    launch_kernel<k_c2r_rows_packed<P, TOut>>(grid, dim3(64 * kPackedRows), 0, st, static_cast<const float2*>(src), to_device(s.im),
                                                                       static_cast<TOut*>(dst), to_device(s.om), s.M, s.o,
                                                                       s.nlines, post, tw, keys, items);
    return hipGetLastError( );
}

static hipError_t launch_c2r_packed(const Stage& s, const void* src, void* dst, RealT rt, float post, const float2* tw,
                                    unsigned long long* keys, const RowBatchItem* items, int n, hipStream_t st) {
    if ( ! c2r_packed_covers(s) || s.nlines < 1 || n < 1 )
        return hipErrorNotSupported;
    if ( s.N == 512 )
        return rt == RealT::f32 ? go_c2r_packed<256, float>(s, src, dst, post, tw, keys, items, n, st)
                                : go_c2r_packed<256, __half>(s, src, dst, post, tw, keys, items, n, st);
    return rt == RealT::f32 ? go_c2r_packed<512, float>(s, src, dst, post, tw, keys, items, n, st)
                            : go_c2r_packed<512, __half>(s, src, dst, post, tw, keys, items, n, st);
}

// Template-matching statistics of a chunk (k_c2r_stats_packed) and their reset
// (a fill kernel, so that a captured graph holds no memset).
// Column-major batch input: row r at r, element k at k * es (es even, every column
// holding whole groups of kCmRows rows).
static bool c2r_cm_input(const Stage& s) {
    return s.im.n2 == 1 && s.im.s1 == 1 && s.im.es > 1 && s.im.es % 2 == 0 &&
           s.im.es >= (s.nlines + kCmRows - 1) / kCmRows * kCmRows;
}

bool c2r_stats_covers(const Stage& s) {
    return c2r_packed_covers(s) && s.im.n2 == 1 && (s.im.es == 1 || c2r_cm_input(s)) && s.om.n2 == 1;
}

// Partial maps of template group g are merged in group order, after the maps
// of earlier chunks and calls, so the first maximum keeps its template.
__global__ void __launch_bounds__(256) k_stats_merge(const float* __restrict__ part, int G, long npix, float* __restrict__ mip,
                                                     int* __restrict__ best, float* __restrict__ sum, float* __restrict__ sum_sq) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if ( i >= npix )
        return;
    const auto* pg = as_global(part);
    float       m = as_global(mip)[i], a = as_global(sum)[i], q = as_global(sum_sq)[i];
    int         b = as_global(best)[i];
    for ( int g = 0; g < G; g++ ) {
        const long  base = 4L * g * npix; // mip, best (as bits), sum, sum_sq of group g
        const float v    = pg[base + i];
        if ( v > m ) {
            m = v;
            b = __float_as_int(pg[base + npix + i]);
        }
        a += pg[base + 2 * npix + i];
        q += pg[base + 3 * npix + i];
    }
    as_global(mip)[i]    = m;
    as_global(best)[i]   = b;
    as_global(sum)[i]    = a;
    as_global(sum_sq)[i] = q;
}

hipError_t launch_c2r_stats_batch(const Stage& s, const void* items, int n, float post, const float2* tw, int first_id,
                                  const BatchStats& bs, hipStream_t st) {
    if ( ! c2r_stats_covers(s) || n < 1 )
        return hipErrorNotSupported;
    // template groups spread a chunk over the GPU when the scratch is there: up to 4 groups of
    // at least 4 templates, and groups of 8 for larger chunks (64 templates: 8 groups of 8, 1 %
    // above 4 of 16 although twice the partial maps), or groups of FASTFFT_STATS_PER_GROUP
    // templates; at most kStatsGroups
    const int tpg = env_int("FASTFFT_STATS_PER_GROUP", 0);
    int       G   = tpg > 0 ? (n + tpg - 1) / tpg : std::max(std::min(4, (n + 3) / 4), (n + 7) / 8);
    G             = bs.part ? std::max(1, std::min(kStatsGroups, G)) : 1;
    const int per = (n + G - 1) / G;
    G             = (n + per - 1) / per; // no empty group
    const dim3  grid((unsigned)((s.nlines + kPackedRows - 1) / kPackedRows), (unsigned)G), block(64 * kPackedRows);
    const auto* it = static_cast<const RowBatchItem*>(items);
    const bool  cm = c2r_cm_input(s);
    auto        go = [&](auto pc, auto hc) {
        if ( cm )
            launch_kernel<k_c2r_stats_packed<decltype(pc)::value, decltype(hc)::value, true>>(grid, block, 0, st,
                    to_device(s.im), s.M, s.o, s.nlines, post, tw, it, n, first_id, bs.mip, bs.best, bs.sum, bs.sum_sq,
                    bs.part);
        else
            launch_kernel<k_c2r_stats_packed<decltype(pc)::value, decltype(hc)::value>>(grid, block, 0, st,
                    to_device(s.im), s.M, s.o, s.nlines, post, tw, it, n, first_id, bs.mip, bs.best, bs.sum, bs.sum_sq,
                    bs.part);
        return hipGetLastError( );
    };
    const bool half = 2 * s.M <= s.N;
    using P256      = std::integral_constant<int, 256>;
    using P512      = std::integral_constant<int, 512>;
    hipError_t e    = s.N == 512 ? (half ? go(P256{ }, std::true_type{ }) : go(P256{ }, std::false_type{ }))
                                 : (half ? go(P512{ }, std::true_type{ }) : go(P512{ }, std::false_type{ }));
    if ( e != hipSuccess || G == 1 )
        return e;
    const long npix = (long)s.nlines * s.M;
    launch_kernel<k_stats_merge>(dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, st, bs.part, G, npix, bs.mip, bs.best, bs.sum,
                                                                             bs.sum_sq);
    return hipGetLastError( );
}

__global__ void __launch_bounds__(256) k_stats_reset(float* mip, int* best, float* sum, float* sum_sq, long n) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if ( i < n ) {
        as_global(mip)[i]    = -__builtin_huge_valf( );
        as_global(best)[i]   = -1;
        as_global(sum)[i]    = 0.f;
        as_global(sum_sq)[i] = 0.f;
    }
}

hipError_t launch_stats_reset(const BatchStats& bs, long n, hipStream_t st) {
    if ( n < 1 )
        return hipSuccess;
    launch_kernel<k_stats_reset>(dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, bs.mip, bs.best, bs.sum, bs.sum_sq, n);
    return hipGetLastError( );
}

// single C2R pass on the packed kernel, with its argmax keys reduced after it
static hipError_t try_c2r_packed(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    if ( ! c2r_packed_covers(s) || io.src_real || io.src_ct != CpxT::c64 || ! io.dst_real )
        return hipErrorNotSupported;
    // rows whose elements lie far apart (a 3-D inverse's x-pass, a 2-D InvFFT's 512-point rows):
    // the tiled line kernel reads 16 neighbouring rows per 128-byte access, where one wave per
    // row would gather 8 bytes per element
    if ( s.im.es != 1 && ! io.amax && ! env_int("FASTFFT_NO_TILE_LINES", 0) ) {
        const hipError_t e = launch_tile_lines(s, io, tw, st);
        if ( e != hipErrorNotSupported )
            return e;
    }
    const bool keys = io.amax && s.nlines <= kArgmaxPartials;
    hipError_t e    = launch_c2r_packed(s, io.src, io.dst, io.dst_rt, io.scale, tw, keys ? io.amax->keys : nullptr,
                                        nullptr, 1, st);
    if ( e == hipSuccess && keys ) {
        e = launch_argmax_keys(io.amax->keys, s.nlines, s.M, io.amax->value, io.amax->xy, st);
        if ( e == hipSuccess && io.amax_fused )
            *io.amax_fused = true;
    }
    return e;
}

// Trimmed C2R passes of 2048, 4096 or 8192 points (and full-length ones asked
// for their argmax): the packed windowed kernel, Q = N / 1024 waves per row
// (rows_packed.hpp), with its argmax keys reduced after it.  Full-length passes
// without argmax keep k_inv_full_wave.  FASTFFT_NO_PACKED_WINDOW switches this
// kernel off alone, FASTFFT_NO_WINDOW_ROWS together with the windowed row
// kernel (the generic line kernel then takes these shapes).
static hipError_t try_c2r_window_packed(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    if ( s.kind != StageKind::inv_c2r || ! (s.N == 2048 || s.N == 4096 || s.N == 8192) || s.M < 1 || s.M > s.N ||
         (s.M == s.N && s.o % s.N == 0 && ! io.amax) ||
         s.nlines < 1 || io.src_real || io.src_ct != CpxT::c64 || ! io.dst_real ||
         env_int("FASTFFT_NO_PACKED_ROWS", 0) || env_int("FASTFFT_NO_PACKED_WINDOW", 0) ||
         env_int("FASTFFT_NO_WINDOW_ROWS", 0) )
        return hipErrorNotSupported;
    const bool          keys = io.amax && s.nlines <= kArgmaxPartials;
    unsigned long long* rk   = keys ? io.amax->keys : nullptr;
    auto                go   = [&](auto qc) -> hipError_t {
        constexpr int Q = decltype(qc)::value;
        const dim3    grid((unsigned)s.nlines), block(64 * Q);
        if ( io.dst_rt == RealT::f32 )
            launch_kernel<k_c2r_window_packed<512, Q, float>>(grid, block, 0, st, static_cast<const float2*>(io.src), to_device(s.im),
                                                                       static_cast<float*>(io.dst), to_device(s.om), s.M,
                                                                       s.o, s.nlines, io.scale, tw, rk);
        else
            launch_kernel<k_c2r_window_packed<512, Q, __half>>(grid, block, 0, st, static_cast<const float2*>(io.src), to_device(s.im),
                                                                        static_cast<__half*>(io.dst), to_device(s.om), s.M,
                                                                        s.o, s.nlines, io.scale, tw, rk);
        return hipGetLastError( );
    };
    hipError_t e = s.N == 2048   ? go(std::integral_constant<int, 2>{ })
                   : s.N == 4096 ? go(std::integral_constant<int, 4>{ })
                                 : go(std::integral_constant<int, 8>{ });
    if ( e == hipSuccess && keys ) {
        e = launch_argmax_keys(io.amax->keys, s.nlines, s.M, io.amax->value, io.amax->xy, st);
        if ( e == hipSuccess && io.amax_fused )
            *io.amax_fused = true;
    }
    return e;
}

// A C2R pass of 2048, 4096 or 8192 points cut to M <= 1024 outputs whose rows are
// the rows of a 1024-point R2C pass (k_c2r_r2c_rows): the real rows stay in LDS.
static bool same_map(const HostLineMap& a, const HostLineMap& b) {
    return a.s1 == b.s1 && a.s2 == b.s2 && a.es == b.es && a.n2 == b.n2;
}

bool c2r_r2c_rows_covers(const Stage& c, const Stage& f) {
    return c.kind == StageKind::inv_c2r && (c.N == 2048 || c.N == 4096 || c.N == 8192) && c.M >= 1 && c.M <= 1024 &&
           c.nlines >= 1 && f.kind == StageKind::fwd_real && f.N == 1024 && f.L == c.M && f.nlines == c.nlines &&
           f.nout >= 1 && f.nout <= 513 && same_map(c.om, f.im) && ! env_int("FASTFFT_NO_FRAME_FUSION", 0) &&
           ! env_int("FASTFFT_NO_PACKED_ROWS", 0) && ! env_int("FASTFFT_NO_PACKED_WINDOW", 0) &&
           ! env_int("FASTFFT_NO_WAVE_ROWS", 0);
}

hipError_t launch_c2r_r2c_rows(const Stage& c, const Stage& f, const float2* src, float2* dst, float post, float pre,
                               const float2* tw, hipStream_t st) {
    if ( ! c2r_r2c_rows_covers(c, f) )
        return hipErrorNotSupported;
    // rows per workgroup: neighbouring rows (line stride one element, even element
    // strides, 16-byte aligned buffers) are moved together by k_c2r_r2c_rows_tile;
    // FASTFFT_FRAME_ROWS = 1, 2 or 4 sets the count (at most 16 waves per workgroup)
    const auto paired = [](const HostLineMap& m) { return m.s1 == 1 && m.n2 == 1 && m.es % 2 == 0; };
    int        rb     = env_int("FASTFFT_FRAME_ROWS", 4);
    if ( rb != 2 && rb != 4 )
        rb = 1;
    if ( rb * (c.N / 1024) > 16 )
        rb = 2;
    if ( ! paired(c.im) || ! paired(f.om) || c.nlines % rb != 0 || (reinterpret_cast<uintptr_t>(src) & 15) != 0 ||
         (reinterpret_cast<uintptr_t>(dst) & 15) != 0 )
        rb = 1;
    auto go = [&](auto qc) -> hipError_t {
        constexpr int Q = decltype(qc)::value;
        if ( rb == 1 )
            launch_kernel<k_c2r_r2c_rows<Q>>(dim3((unsigned)c.nlines), dim3(64 * Q), 0, st,
                    src, to_device(c.im), c.M, c.o, c.nlines, post, dst, to_device(f.om), f.nout, pre, tw);
        else if ( rb == 2 )
            launch_kernel<k_c2r_r2c_rows_tile<Q, 2>>(dim3((unsigned)(c.nlines / 2)), dim3(64 * Q * 2), 0, st,
                    src, to_device(c.im), c.M, c.o, c.nlines, post, dst, to_device(f.om), f.nout, pre, tw);
        else if constexpr ( Q <= 4 )
            launch_kernel<k_c2r_r2c_rows_tile<Q, 4>>(dim3((unsigned)(c.nlines / 4)), dim3(64 * Q * 4), 0, st,
                    src, to_device(c.im), c.M, c.o, c.nlines, post, dst, to_device(f.om), f.nout, pre, tw);
        return hipGetLastError( );
    };
    return c.N == 2048 ? go(std::integral_constant<int, 2>{ })
           : c.N == 4096 ? go(std::integral_constant<int, 4>{ })
                         : go(std::integral_constant<int, 8>{ });
}

// Trimmed C2R rows with P in {256, 512, 1024} and a window origin at register
// rotation 0 or E/2: one block per row, one wave per sub-transform.
template <int P, int Q, int RS, typename TOut>
static hipError_t go_c2r_wave(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    // the argmax keys need one slot per row
    const bool keys = io.amax && s.nlines <= kArgmaxPartials;
    launch_kernel<k_c2r_rows_wave<P, Q, RS, TOut>>(dim3((unsigned)s.nlines), dim3(64 * Q), 0, st,
            static_cast<const float2*>(io.src), to_device(s.im), static_cast<TOut*>(io.dst), to_device(s.om), s.M, s.o,
            s.nlines, io.scale, tw, keys ? io.amax->keys : nullptr, nullptr);
    hipError_t e = hipGetLastError( );
    if ( e == hipSuccess && keys ) {
        e = launch_argmax_keys(io.amax->keys, s.nlines, s.M, io.amax->value, io.amax->xy, st);
        if ( e == hipSuccess && io.amax_fused )
            *io.amax_fused = true;
    }
    return e;
}

template <typename TOut>
static hipError_t try_c2r_wave(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    if ( s.o % 64 != 0 )
        return hipErrorNotSupported;
    auto by_p = [&](auto pc) -> hipError_t {
        constexpr int P  = decltype(pc)::value;
        constexpr int E  = P / 64;
        const int     rs = (((s.o / 64) % E) + E) % E;
        if ( rs != 0 && rs != E / 2 )
            return hipErrorNotSupported;
        auto by_q = [&](auto qc) -> hipError_t {
            constexpr int Q = decltype(qc)::value;
            if constexpr ( P * Q > kTwiddleN ) {
                return hipErrorNotSupported;
            }
            else {
                return rs == 0 ? go_c2r_wave<P, Q, 0, TOut>(s, io, tw, st) : go_c2r_wave<P, Q, E / 2, TOut>(s, io, tw, st);
            }
        };
        switch ( s.Q ) {
            case 1: return by_q(std::integral_constant<int, 1>{ });
            case 2: return by_q(std::integral_constant<int, 2>{ });
            case 4: return by_q(std::integral_constant<int, 4>{ });
            case 8: return by_q(std::integral_constant<int, 8>{ });
            case 16: return by_q(std::integral_constant<int, 16>{ });
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

template <int P, int Q, int RS, typename TOut>
static hipError_t go_c2r_wave_batch(const Stage& s, const RowBatchItem* items, int n, float post, const float2* tw,
                                    hipStream_t st) {
    launch_kernel<k_c2r_rows_wave<P, Q, RS, TOut>>(dim3((unsigned)s.nlines, (unsigned)n), dim3(64 * Q), 0, st,
            nullptr, to_device(s.im), nullptr, to_device(s.om), s.M, s.o, s.nlines, post, tw, nullptr, items);
    return hipGetLastError( );
}

// window origins and (P, Q) that k_c2r_rows_wave covers
static int c2r_rows_rotation(const Stage& s) {
    if ( s.o % 64 != 0 || ! (s.P == 256 || s.P == 512 || s.P == 1024) || s.P * s.Q > kTwiddleN ||
         ! (s.Q == 1 || s.Q == 2 || s.Q == 4 || s.Q == 8 || s.Q == 16) )
        return -1;
    const int E  = s.P / 64;
    const int rs = (((s.o / 64) % E) + E) % E;
    return rs == 0 || rs == E / 2 ? rs : -1;
}

bool c2r_rows_batch_covers(const Stage& s) {
    if ( c2r_cm_input(s) )
        return c2r_packed_covers(s) && s.nlines <= kArgmaxPartials && s.om.n2 == 1;
    return (c2r_rows_rotation(s) >= 0 || c2r_packed_covers(s)) && s.nlines <= kArgmaxPartials && s.im.n2 == 1 &&
           s.om.n2 == 1;
}

template <int P, typename TOut>
static hipError_t go_c2r_packed_cm(const Stage& s, const RowBatchItem* items, int n, float post, const float2* tw,
                                   hipStream_t st) {
    const dim3 grid((unsigned)((s.nlines + kCmRows - 1) / kCmRows), (unsigned)n);
    // a window at origin 0 of at most 128 KEEP outputs needs the first KEEP registers only
    constexpr int E  = P / 64;
    const bool    o0 = ((s.o % s.N) + s.N) % s.N == 0;
    auto          go = [&](auto kc) {
        launch_kernel<k_c2r_rows_packed_cm<P, TOut, decltype(kc)::value>>(grid, dim3(64 * kCmWaves), 0, st,
                s.im.es, to_device(s.om), s.M, s.o, s.nlines, post, tw, items);
        return hipGetLastError( );
    };
    if ( o0 && s.M <= 256 && E > 2 )
        return go(std::integral_constant<int, (E > 2 ? 2 : E)>{ });
    if ( o0 && s.M <= 512 && E > 4 )
        return go(std::integral_constant<int, (E > 4 ? 4 : E)>{ });
    return go(std::integral_constant<int, E>{ });
}

hipError_t launch_c2r_rows_batch(const Stage& s, const void* items, int n, RealT out_rt, float post, const float2* tw,
                                 hipStream_t st) {
    if ( ! c2r_rows_batch_covers(s) )
        return hipErrorNotSupported;
    const auto* it = static_cast<const RowBatchItem*>(items);
    if ( c2r_cm_input(s) ) {
        if ( s.N == 512 )
            return out_rt == RealT::f32 ? go_c2r_packed_cm<256, float>(s, it, n, post, tw, st)
                                        : go_c2r_packed_cm<256, __half>(s, it, n, post, tw, st);
        return out_rt == RealT::f32 ? go_c2r_packed_cm<512, float>(s, it, n, post, tw, st)
                                    : go_c2r_packed_cm<512, __half>(s, it, n, post, tw, st);
    }
    {
        hipError_t e = launch_c2r_packed(s, nullptr, nullptr, out_rt, post, tw, nullptr, it, n, st);
        if ( e != hipErrorNotSupported )
            return e;
    }
    const int rs = c2r_rows_rotation(s);
    auto      by_p = [&](auto pc) -> hipError_t {
        constexpr int P  = decltype(pc)::value;
        constexpr int E  = P / 64;
        auto          go = [&](auto qc) -> hipError_t {
            constexpr int Q = decltype(qc)::value;
            if constexpr ( P * Q > kTwiddleN ) {
                return hipErrorNotSupported;
            }
            else {
                if ( out_rt == RealT::f32 )
                    return rs == 0 ? go_c2r_wave_batch<P, Q, 0, float>(s, it, n, post, tw, st)
                                   : go_c2r_wave_batch<P, Q, E / 2, float>(s, it, n, post, tw, st);
                return rs == 0 ? go_c2r_wave_batch<P, Q, 0, __half>(s, it, n, post, tw, st)
                               : go_c2r_wave_batch<P, Q, E / 2, __half>(s, it, n, post, tw, st);
            }
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

// Full-length inverse lines (M = N outputs from 0, 256 <= N <= 8192): one
// workgroup per line, P = 512 (256 for N = 256) and Q = N / P.
template <int P, int Q>
static hipError_t go_full_inv(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    const dim3 grid((unsigned)s.nlines), block(64 * Q);
    const auto im = to_device(s.im), om = to_device(s.om);
    if ( s.kind == StageKind::inv_c2r ) {
        const auto* in = static_cast<const float2*>(io.src);
        if ( io.dst_rt == RealT::f32 )
            launch_kernel<k_inv_full_wave<P, Q, float2, true, float, true>>(grid, block, 0, st, in, im, static_cast<float*>(io.dst), om, s.nlines, io.scale, tw);
        else
            launch_kernel<k_inv_full_wave<P, Q, float2, true, __half, true>>(grid, block, 0, st, in, im, static_cast<__half*>(io.dst), om, s.nlines, io.scale, tw);
    }
    else if ( io.src_ct == CpxT::c64 ) {
        launch_kernel<k_inv_full_wave<P, Q, float2, false, float2, false>>(grid, block, 0, st,
                static_cast<const float2*>(io.src), im, static_cast<float2*>(io.dst), om, s.nlines, io.scale, tw);
    }
    else {
        launch_kernel<k_inv_full_wave<P, Q, __half2, false, float2, false>>(grid, block, 0, st,
                static_cast<const __half2*>(io.src), im, static_cast<float2*>(io.dst), om, s.nlines, io.scale, tw);
    }
    return hipGetLastError( );
}

static hipError_t try_full_inv(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    if ( s.M != s.N || s.o % s.N != 0 || io.src_real || (io.src_ct == CpxT::c32h && s.kind == StageKind::inv_c2r) )
        return hipErrorNotSupported;
    if ( s.kind == StageKind::inv_c2r ? ! io.dst_real : (io.dst_real || io.dst_ct != CpxT::c64) )
        return hipErrorNotSupported;
    switch ( s.N ) {
        case 256: return go_full_inv<256, 1>(s, io, tw, st);
        case 512: return go_full_inv<512, 1>(s, io, tw, st);
        case 1024: return go_full_inv<512, 2>(s, io, tw, st);
        case 2048: return go_full_inv<512, 4>(s, io, tw, st);
        case 4096: return go_full_inv<512, 8>(s, io, tw, st);
        case 8192: return go_full_inv<512, 16>(s, io, tw, st);
        default: return hipErrorNotSupported;
    }
}

// Output-pruned inverse lines of N in {256, ..., 8192} with any window (P = 512,
// or 256 for N = 256, whatever the stage's own sub-transform size).
template <int P, int Q, typename TIn, bool HERM, typename TOut, bool REAL_OUT>
static hipError_t go_window(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    launch_kernel<k_inv_window_rows<P, Q, TIn, HERM, TOut, REAL_OUT>>(dim3((unsigned)s.nlines), dim3(64 * Q), 0, st,
            static_cast<const TIn*>(io.src), to_device(s.im), static_cast<TOut*>(io.dst), to_device(s.om), s.M, s.o,
            s.nlines, io.scale, tw);
    return hipGetLastError( );
}

static hipError_t try_window(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    if ( env_int("FASTFFT_NO_WINDOW_ROWS", 0) || s.M < 1 || s.M > s.N || s.nlines < 1 )
        return hipErrorNotSupported;
    auto go = [&](auto pc, auto qc) -> hipError_t {
        constexpr int P = decltype(pc)::value;
        constexpr int Q = decltype(qc)::value;
        if ( s.kind == StageKind::inv_c2r ) {
            if ( io.src_real || io.src_ct != CpxT::c64 || ! io.dst_real )
                return hipErrorInvalidValue;
            return io.dst_rt == RealT::f32 ? go_window<P, Q, float2, true, float, true>(s, io, tw, st)
                                           : go_window<P, Q, float2, true, __half, true>(s, io, tw, st);
        }
        if ( io.src_real || io.dst_real || io.dst_ct != CpxT::c64 )
            return hipErrorInvalidValue;
        return io.src_ct == CpxT::c64 ? go_window<P, Q, float2, false, float2, false>(s, io, tw, st)
                                      : go_window<P, Q, __half2, false, float2, false>(s, io, tw, st);
    };
    using I = std::integral_constant<int, 1>;
    switch ( s.N ) {
        case 256: return go(std::integral_constant<int, 256>{ }, I{ });
        case 512: return go(std::integral_constant<int, 512>{ }, I{ });
        case 1024: return go(std::integral_constant<int, 512>{ }, std::integral_constant<int, 2>{ });
        case 2048: return go(std::integral_constant<int, 512>{ }, std::integral_constant<int, 4>{ });
        case 4096: return go(std::integral_constant<int, 512>{ }, std::integral_constant<int, 8>{ });
        case 8192: return go(std::integral_constant<int, 512>{ }, std::integral_constant<int, 16>{ });
        default: return hipErrorNotSupported;
    }
}

hipError_t launch_inv(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    if ( ! env_int("FASTFFT_NO_WAVE_ROWS", 0) ) {
        hipError_t e = try_c2r_packed(s, io, tw, st);
        if ( e == hipErrorNotSupported )
            e = try_c2r_window_packed(s, io, tw, st);
        if ( e != hipErrorNotSupported )
            return e;
    }
    if ( ! env_int("FASTFFT_NO_WAVE_ROWS", 0) && ! env_int("FASTFFT_NO_FULL_LINES", 0) && ! io.amax ) {
        hipError_t e = try_full_inv(s, io, tw, st);
        if ( e != hipErrorNotSupported )
            return e;
    }
    // 256 / 512-point lines, 16 per workgroup (a C2R pass asked for its argmax keeps the row kernel)
    if ( (s.kind == StageKind::inv_cpx || (s.kind == StageKind::inv_c2r && ! io.amax)) && ! env_int("FASTFFT_NO_WAVE_ROWS", 0) ) {
        hipError_t e = launch_tile_lines(s, io, tw, st);
        if ( e != hipErrorNotSupported )
            return e;
    }
    if ( s.kind == StageKind::inv_c2r && io.dst_real && ! io.src_real && io.src_ct == CpxT::c64 &&
         ! env_int("FASTFFT_NO_WAVE_ROWS", 0) ) {
        hipError_t e = io.dst_rt == RealT::f32 ? try_c2r_wave<float>(s, io, tw, st) : try_c2r_wave<__half>(s, io, tw, st);
        if ( e != hipErrorNotSupported )
            return e;
    }
    // (an argmax asked of this C2R pass stays unfused: the caller runs it separately)
    if ( ! env_int("FASTFFT_NO_WAVE_ROWS", 0) ) {
        hipError_t e = try_window(s, io, tw, st);
        if ( e != hipErrorNotSupported )
            return e;
    }
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
