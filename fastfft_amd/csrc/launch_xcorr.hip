// Fused correlation pass of FwdImageInvFFT.
//
// Lines of a single index (the 2-D second axis) with P in {256, 512, 1024}, a
// window origin that is a multiple of 64 placed at register rotation 0 or E/2,
// N = P*Q <= 8192 and a contiguous, 16-byte aligned image column run the
// workgroup-per-column kernel of xcorr_col.hpp; every other shape runs the
// generic line kernel.
#include <algorithm>
#include <cstdint>

#include "launch_common.hpp"
#include "xcorr_col.hpp"
#include "xcorr_toeplitz.hpp"

namespace fastfft_amd {

template <int P, typename TImg>
static hipError_t go_xcorr(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    dim3       grid, block;
    size_t     lds;
    hipError_t e = prepare_launch<P>(k_xcorr<P, TImg>, s.nlines, grid, block, lds);
    if ( e != hipSuccess )
        return e;
    launch_kernel<k_xcorr<P, TImg>>(grid, block, lds, st, static_cast<const float2*>(io.src), to_device(s.im), s.L,
                                                static_cast<const TImg*>(io.img), to_device(s.gm),
                                                static_cast<float2*>(io.dst), to_device(s.om), s.N, s.M, s.o, s.Q,
                                                s.nlines, io.scale, tw, twiddle_stride(s.N));
    return hipGetLastError( );
}

// One workgroup per column; columns 0 and nlines-1 share workgroup 0 when they
// are the DC and Nyquist columns of a real transform (FASTFFT_XCORR_PACK=0
// gives them a workgroup each, for tests).
// Several templates per launch (io.xitems): the templates are split into groups
// of kXcorrGroupTemplates along gridDim.y (each group stages its image column
// again; FASTFFT_XCORR_GROUPS overrides the count).  Two templates per group
// were fastest at 64 x 384^2 -> 512^2 on the 1024 frame: 133 us per 64-template
// launch against 152 us with 8 per group and 142 us with 1.
constexpr int kXcorrGroupTemplates = 2;

template <int P, int Q, int RS, typename TImg>
static hipError_t go_xcorr_col(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    const int pack   = s.real_end_lines && s.nlines >= 2 && env_int("FASTFFT_XCORR_PACK", 1) ? 1 : 0;
    const int blocks = s.nlines - pack;
    int       groups = 1;
    if ( io.xitems && io.nxitems > 1 && blocks % 8 == 0 ) {
        const int per = io.xgroup > 0 ? io.xgroup : kXcorrGroupTemplates;
        groups        = env_int("FASTFFT_XCORR_GROUPS", (io.nxitems + per - 1) / per);
        groups        = std::max(1, std::min(groups, io.nxitems));
    }
    XcorrItem one{io.src, io.dst, s.L, 0};
    const dim3 grid((unsigned)blocks, (unsigned)groups), block(64 * Q);
    // one template per launch: the single-template form (FASTFFT_XCORR_ONE=0 keeps the loop form)
    if ( ! io.xitems && env_int("FASTFFT_XCORR_ONE", 1) )
        launch_kernel<k_xcorr_col<P, Q, RS, TImg, true>>(grid, block, 0, st, one, nullptr, 1, s.im.s1, s.im.es,
                static_cast<const TImg*>(io.img), s.gm.s1, s.om.s1, s.om.es, s.M, s.o, io.scale, tw, s.nlines - 1, pack);
    else
        launch_kernel<k_xcorr_col<P, Q, RS, TImg>>(grid, block, 0, st, one, static_cast<const XcorrItem*>(io.xitems),
                io.nxitems, s.im.s1, s.im.es, static_cast<const TImg*>(io.img), s.gm.s1, s.om.s1, s.om.es, s.M, s.o, io.scale,
                tw, s.nlines - 1, pack);
    return hipGetLastError( );
}

template <int P, typename TImg>
static hipError_t try_xcorr_col(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    constexpr int E = P / 64;
    if ( s.o % 64 != 0 )
        return hipErrorNotSupported;
    const int rs = (((s.o / 64) % E) + E) % E;
    auto      by_q = [&](auto qc) -> hipError_t {
        constexpr int Q = decltype(qc)::value;
        if constexpr ( P * Q > kTwiddleN || (Q == 1 && P != 1024) ) {
            return hipErrorNotSupported;
        }
        else {
            if ( rs == 0 )
                return go_xcorr_col<P, Q, 0, TImg>(s, io, tw, st);
            if ( rs == E / 2 )
                return go_xcorr_col<P, Q, E / 2, TImg>(s, io, tw, st);
            return hipErrorNotSupported;
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
}

static bool single_index_lines(const HostLineMap& m) { return m.n2 == 1 && m.s2 == 0; }

bool xcorr_col_covers(const Stage& s, CpxT img_ct) {
    const int per = img_ct == CpxT::c32h ? 4 : 2;
    if ( ! (single_index_lines(s.im) && single_index_lines(s.om) && single_index_lines(s.gm) && s.gm.es == 1 &&
            s.gm.s1 % per == 0) )
        return false;
    if ( ! (s.P == 256 || s.P == 512 || s.P == 1024) || s.o % 64 != 0 || s.P * s.Q > kTwiddleN || s.P < s.L ||
         ! ((s.Q == 1 && s.P == 1024) || s.Q == 2 || s.Q == 4 || s.Q == 8 || s.Q == 16) )
        return false;
    const int E  = s.P / 64;
    const int rs = (((s.o / 64) % E) + E) % E;
    return rs == 0 || rs == E / 2;
}

bool xcorr_toeplitz_covers(const Stage& s) {
    // column-major template and output slots (rows contiguous, template columns of a
    // multiple of 16 rows)
    return s.kind == StageKind::xcorr && s.N == kTzFy && s.L >= 1 && s.L <= kTzMaxL && s.M >= 1 && s.M <= kTzMaxM &&
           s.L + s.M - 1 <= s.N && s.o == 0 && s.real_end_lines && s.nlines >= 2 && single_index_lines(s.im) &&
           s.im.es == 1 && s.im.s1 % 16 == 0 && s.im.s1 >= s.L && single_index_lines(s.om) && s.om.es == 1 &&
           s.om.s1 >= s.M && single_index_lines(s.gm) && s.gm.es == 1 && ! env_int("FASTFFT_NO_TOEPLITZ", 0);
}

// One workgroup per column (the DC and Nyquist columns share workgroup 0), each running
// every template group of 32; the frame's N-point y transforms are not taken, so the
// products carry N.
static hipError_t go_xcorr_toeplitz(const Stage& s, const StageIO& io, hipStream_t st) {
    const int n = io.nxitems;
    if ( ! io.xitems || n < 1 || io.xout_step <= 0 )
        return hipErrorNotSupported;
    launch_kernel<k_xcorr_toeplitz>(dim3((unsigned)(s.nlines - 1)), dim3(kTzWaves * 64), 0, st,
            static_cast<const XcorrItem*>(io.xitems), n, s.im.s1, static_cast<const float2*>(io.img), s.gm.s1, s.om.s1,
            io.xout_step, s.M, s.L, io.scale * (float)s.N, s.nlines - 1);
    return hipGetLastError( );
}

bool batch_kernels_cover(const Stage& r2c, const Stage& xcorr, const Stage& c2r, CpxT img_ct) {
    return r2c.kind == StageKind::fwd_real && xcorr.kind == StageKind::xcorr && c2r.kind == StageKind::inv_c2r &&
           r2c_rows_batch_covers(r2c) && xcorr_col_covers(xcorr, img_ct) && c2r_rows_batch_covers(c2r) &&
           ! env_int("FASTFFT_NO_WAVE_XCORR", 0) && ! env_int("FASTFFT_NO_WAVE_ROWS", 0);
}

hipError_t launch_xcorr(const Stage& s, const StageIO& io, const float2* tw, hipStream_t st) {
    if ( io.img_mixed ) {
        // a frame in mixed form is only taken by the matrix-core kernel
        if ( ! (io.split_bf16 && io.img_ct == CpxT::c64 && xcorr_toeplitz_covers(s)) )
            return hipErrorNotSupported;
        return go_xcorr_toeplitz(s, io, st);
    }
    // 16-byte image loads: aligned image base (the column starts are checked with the shape)
    const bool col_ok = xcorr_col_covers(s, io.img_ct) && reinterpret_cast<std::uintptr_t>(io.img) % 16 == 0 &&
                        ! env_int("FASTFFT_NO_WAVE_XCORR", 0);
    if ( col_ok ) {
        hipError_t e    = hipErrorNotSupported;
        auto       pick = [&](auto pc) -> hipError_t {
            constexpr int P = decltype(pc)::value;
            return io.img_ct == CpxT::c64 ? try_xcorr_col<P, float2>(s, io, tw, st) : try_xcorr_col<P, __half2>(s, io, tw, st);
        };
        if ( s.P == 256 )
            e = pick(std::integral_constant<int, 256>{ });
        else if ( s.P == 512 )
            e = pick(std::integral_constant<int, 512>{ });
        else if ( s.P == 1024 )
            e = pick(std::integral_constant<int, 1024>{ });
        if ( e != hipErrorNotSupported )
            return e;
    }
    // the generic kernel correlates one template per launch
    if ( io.xitems )
        return hipErrorNotSupported;
    return dispatch_P(s.P, [&](auto pc) -> hipError_t {
        constexpr int P = decltype(pc)::value;
        if ( io.img_ct == CpxT::c64 )
            return go_xcorr<P, float2>(s, io, tw, st);
        return go_xcorr<P, __half2>(s, io, tw, st);
    });
}

} // namespace fastfft_amd
