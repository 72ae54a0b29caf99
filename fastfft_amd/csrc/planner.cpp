// GPU-free planner: validation, padded sizes and stage lists.
#include "plan.hpp"

#include <algorithm>
#include <cstdlib>
#include <sstream>

#include "../../include/fastfft/fastfft.h"

namespace fastfft_amd {

static bool is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }

// A set, non-zero environment switch (read when a plan's stage lists are built).
static bool env_flag_planner(const char* name) {
    const char* v = std::getenv(name);
    return v && *v && std::atoi(v) != 0;
}

static int next_pow2(int n) {
    int p = 1;
    while ( p < n )
        p <<= 1;
    return p;
}

long padded_memory_size(Dims& d) {
    // Real input: x/2+1 complex per row when x is even, (x-1)/2+1 when odd, two
    // floats each; w is the number of complex values along x after R2C.
    long w = (d.x % 2 == 0) ? d.x / 2 + 1 : (d.x - 1) / 2 + 1;
    d.w    = (int)w;
    return 2 * w * (long)d.y * (long)d.z;
}

static SizeChange classify(const Dims& in, const Dims& out, bool& mixed) {
    mixed             = false;
    const bool any_up = out.x > in.x || out.y > in.y || out.z > in.z;
    const bool any_dn = out.x < in.x || out.y < in.y || out.z < in.z;
    if ( any_up && any_dn )
        mixed = true;
    if ( any_up )
        return SizeChange::increase;
    if ( any_dn )
        return SizeChange::decrease;
    return SizeChange::no_change;
}

PlanError validate(PlanState& p) {
    PlanError e;
    if ( ! p.fwd_set || ! p.inv_set ) {
        e.code = FASTFFT_NOT_PLANNED;
        e.msg  = "both SetForwardFFTPlan and SetInverseFFTPlan must be called before a transform";
        return e;
    }
    if ( ! p.padded ) {
        e.code = FASTFFT_UNSUPPORTED;
        e.msg  = "input arrays that are not FFTW padded are not supported (FastFFT.cu:147)";
        return e;
    }
    if ( p.fwd_out.x != p.inv_in.x || p.fwd_out.y != p.inv_in.y || p.fwd_out.z != p.inv_in.z ) {
        e.code = FASTFFT_INVALID_DIMENSION;
        e.msg  = "all fwd output dimensions must match the inv input dimensions";
        return e;
    }
    bool mixed;
    p.fwd_change = classify(p.fwd_in, p.fwd_out, mixed);
    if ( mixed ) {
        e.code = FASTFFT_INVALID_DIMENSION;
        e.msg  = "fwd plan: all dimensions must either increase, decrease or stay the same";
        return e;
    }
    p.inv_change = classify(p.inv_in, p.inv_out, mixed);
    if ( mixed ) {
        e.code = FASTFFT_INVALID_DIMENSION;
        e.msg  = "inv plan: all dimensions must either increase, decrease or stay the same";
        return e;
    }
    if ( p.fwd_in.z == 1 && p.fwd_out.z == 1 ) {
        if ( p.inv_in.z != 1 || p.inv_out.z != 1 ) {
            e.code = FASTFFT_INVALID_DIMENSION;
            e.msg  = "fwd/inv dimensionality may not change (z dimension)";
            return e;
        }
        if ( p.fwd_in.y == 1 && p.fwd_out.y == 1 ) {
            if ( p.inv_in.y != 1 || p.inv_out.y != 1 ) {
                e.code = FASTFFT_INVALID_DIMENSION;
                e.msg  = "fwd/inv dimensionality may not change (y dimension)";
                return e;
            }
            p.transform_dimension = 1;
        }
        else {
            p.transform_dimension = 2;
        }
    }
    else {
        p.transform_dimension = 3;
        const Dims* all[4]    = {&p.fwd_in, &p.fwd_out, &p.inv_in, &p.inv_out};
        for ( const Dims* d : all )
            if ( d->x > kMaxSize3D || d->y > kMaxSize3D || d->z > kMaxSize3D ) {
                e.code = FASTFFT_INVALID_DIMENSION;
                e.msg  = "all dimensions must be <= 512 for 3d transforms";
                return e;
            }
    }
    if ( p.transform_dimension != p.rank ) {
        e.code = p.transform_dimension == 1 ? FASTFFT_UNSUPPORTED : FASTFFT_INVALID_DIMENSION;
        e.msg  = "plan dimensionality (" + std::to_string(p.transform_dimension) + "d) does not match the transformer Rank " +
                std::to_string(p.rank);
        return e;
    }
    // Non-power-of-two sizes are allowed only on the small side of a size change.
    const int fin[3]  = {p.fwd_in.x, p.fwd_in.y, p.fwd_in.z};
    const int fout[3] = {p.fwd_out.x, p.fwd_out.y, p.fwd_out.z};
    const int iout[3] = {p.inv_out.x, p.inv_out.y, p.inv_out.z};
    const char ax[3]  = {'x', 'y', 'z'};
    for ( int a = 0; a < 3; a++ ) {
        if ( ! is_pow2(fin[a]) && p.fwd_change != SizeChange::increase ) {
            e.code = FASTFFT_INVALID_DIMENSION;
            e.msg  = std::string("fwd input ") + ax[a] + " dimension must be a power of 2 unless the forward transform pads";
            return e;
        }
        if ( ! is_pow2(fout[a]) ) {
            e.code = FASTFFT_INVALID_DIMENSION;
            e.msg  = std::string("fwd output (= inv input) ") + ax[a] + " dimension must be a power of 2";
            return e;
        }
        if ( ! is_pow2(iout[a]) && p.inv_change != SizeChange::decrease ) {
            e.code = FASTFFT_INVALID_DIMENSION;
            e.msg  = std::string("inv output ") + ax[a] + " dimension must be a power of 2 unless the inverse trims";
            return e;
        }
    }
    const int nmax = p.transform_dimension == 3 ? kMaxSize3D : kMaxSize2D;
    if ( p.fwd_out.x > nmax || p.fwd_out.y > nmax || p.fwd_in.x > nmax || p.fwd_in.y > nmax ) {
        e.code = FASTFFT_INVALID_DIMENSION;
        e.msg  = "transform sizes above " + std::to_string(nmax) + " are not supported";
        return e;
    }
    if ( p.fwd_out.x < kMinSize || p.fwd_out.y < kMinSize ) {
        e.code = FASTFFT_INVALID_DIMENSION;
        e.msg  = "transform sizes below 2 are not supported";
        return e;
    }
    p.validated = true;
    return e;
}

int pick_sub_size(int N, int L, int M) {
    int need = std::max(std::max(L, M), 1);
    int P    = std::max(16, next_pow2(need));
    return std::min(P, N);
}

// Leading dimension (complex elements) of a workspace array of lines holding n
// values each.  The x-passes access these arrays across lines (one element per
// line); FASTFFT_WS_PAD adds elements per line to move those accesses off a
// single memory channel.
static long ws_ld(int n) {
    const char* v   = std::getenv("FASTFFT_WS_PAD");
    const long  pad = v ? std::max(0L, std::atol(v)) : 0L;
    return n + pad;
}

// Leading dimension of a row-major workspace array of n complex values per row:
// even, so that every row starts 16-byte aligned for the 16-byte row accesses.
static long ws_ld_rows(int n) {
    const long ld = ws_ld(n);
    return ld + (ld & 1);
}

static HostLineMap lm(long s1, long s2, long es, int n2) {
    HostLineMap m;
    m.s1 = s1;
    m.s2 = s2;
    m.es = es;
    m.n2 = n2;
    return m;
}

static Stage make_stage(StageKind k, const char* ref, int N, int L, int M, int o, int nout, int nlines) {
    Stage s;
    s.kind       = k;
    s.ref_kernel = ref;
    s.N          = N;
    s.L          = L;
    s.M          = M;
    s.o          = o;
    s.nout       = nout;
    s.nlines     = nlines;
    s.P          = pick_sub_size(N, L, M);
    s.Q          = N / s.P;
    return s;
}

// Window origin along an axis: only meaningful for a trimmed inverse.
static int axis_origin(const PlanState& p, int a, int n_in, int n_out) { return n_out < n_in ? p.origin[a] : 0; }

static PlanError fwd_stages_2d(const PlanState& p, std::vector<Stage>& st, Buf final_dst) {
    PlanError  e;
    const Dims &in = p.fwd_in, &out = p.fwd_out;
    const int  pitch = 2 * in.w;
    // x-pass: each of the in.y real rows -> A[kx][y], ldA = in.y
    // fwd decrease keeps the first out.x/2+1 frequencies of the full in.x transform
    const int Nx = std::max(in.x, out.x);
    Stage     s1 = make_stage(StageKind::fwd_real, p.fwd_change == SizeChange::decrease ? "r2c_decrease_XY" :
                                                   p.fwd_change == SizeChange::increase ? "r2c_increase_XY" : "r2c_none_XY",
                              Nx, in.x, 0, 0, out.w, in.y);
    s1.src      = Buf::user_in;
    s1.src_elem = Elem::real_input_type;
    s1.im       = lm(pitch, 0, 1, 1);
    s1.dst      = Buf::ws_a;
    s1.dst_elem = Elem::cpx_f32;
    s1.om       = lm(1, 0, ws_ld(in.y), 1);
    st.push_back(s1);
    // y-pass: each of the out.w columns A[kx][0..in.y) -> C[kx][ky]
    const int Ny = std::max(in.y, out.y);
    Stage     s2 = make_stage(StageKind::fwd_cpx, p.fwd_change == SizeChange::decrease ? "c2c_fwd_decrease" :
                                                  p.fwd_change == SizeChange::increase ? "c2c_fwd_increase" : "c2c_fwd_none",
                              Ny, in.y, 0, 0, out.y, out.w);
    s2.src      = Buf::ws_a;
    s2.src_elem = Elem::cpx_f32;
    s2.im       = lm(ws_ld(in.y), 0, 1, 1);
    s2.dst      = final_dst;
    s2.dst_elem = Elem::cpx_input_type;
    s2.om       = lm(out.y, 0, 1, 1);
    st.push_back(s2);
    return e;
}

static PlanError fwd_stages_3d(const PlanState& p, std::vector<Stage>& st) {
    PlanError   e;
    const Dims &in = p.fwd_in, &out = p.fwd_out;
    if ( p.fwd_change == SizeChange::decrease ) {
        e.code = FASTFFT_UNSUPPORTED;
        e.msg  = "3D FFT fwd decrease not yet supported (FastFFT.cu:449-453)";
        return e;
    }
    const int pitch = 2 * in.w, w = out.w;
    // x-pass on the in.z*in.y rows -> T1[z][kx][y]
    Stage s1    = make_stage(StageKind::fwd_real, p.fwd_change == SizeChange::increase ? "r2c_increase_XZ" : "r2c_none_XZ", out.x,
                             in.x, 0, 0, w, in.z * in.y);
    s1.src      = Buf::user_in;
    s1.src_elem = Elem::real_input_type;
    s1.im       = lm((long)in.y * pitch, pitch, 1, in.y);
    s1.dst      = Buf::ws_a;
    s1.dst_elem = Elem::cpx_f32;
    s1.om       = lm((long)w * in.y, 1, in.y, in.y);
    st.push_back(s1);
    // y-pass on the in.z*w lines -> T2[z][kx][ky]
    Stage s2    = make_stage(StageKind::fwd_cpx, p.fwd_change == SizeChange::increase ? "c2c_fwd_increase_XYZ" : "c2c_fwd_none_XYZ",
                             out.y, in.y, 0, 0, out.y, in.z * w);
    s2.src      = Buf::ws_a;
    s2.src_elem = Elem::cpx_f32;
    s2.im       = lm(in.y, 0, 1, 1);
    s2.dst      = Buf::ws_b;
    s2.dst_elem = Elem::cpx_f32;
    s2.om       = lm(out.y, 0, 1, 1);
    st.push_back(s2);
    // z-pass on the w*out.y lines (stride w*out.y) -> C[kz][kx][ky]
    Stage s3    = make_stage(StageKind::fwd_cpx, p.fwd_change == SizeChange::increase ? "c2c_fwd_increase" : "c2c_fwd_none", out.z,
                             in.z, 0, 0, out.z, w * out.y);
    s3.src      = Buf::ws_b;
    s3.src_elem = Elem::cpx_f32;
    s3.im       = lm(1, 0, (long)w * out.y, 1);
    s3.dst      = Buf::user_out;
    s3.dst_elem = Elem::cpx_input_type;
    s3.om       = lm(1, 0, (long)w * out.y, 1);
    st.push_back(s3);
    return e;
}

// Inverse stages.  src_cpx is the buffer holding C[kx][ky] (2-D) / C[kz][kx][ky] (3-D).
static PlanError inv_stages_2d(const PlanState& p, std::vector<Stage>& st, Buf src, Elem src_elem) {
    PlanError   e;
    const Dims &in = p.inv_in, &out = p.inv_out;
    if ( p.inv_change == SizeChange::increase ) {
        e.code = FASTFFT_UNSUPPORTED;
        e.msg  = "2D inverse size increase is not supported";
        return e;
    }
    const bool dec = p.inv_change == SizeChange::decrease;
    // y-pass on the in.w columns, keeping out.y rows of the window -> B[kx][j], ldB = out.y
    Stage s1    = make_stage(StageKind::inv_cpx, dec ? "c2c_inv_decrease" : "c2c_inv_none", in.y, 0, out.y,
                             axis_origin(p, 1, in.y, out.y), 0, in.w);
    s1.src      = src;
    s1.src_elem = src_elem;
    s1.im       = lm(in.y, 0, 1, 1);
    s1.dst      = Buf::ws_b;
    s1.dst_elem = Elem::cpx_f32;
    s1.om       = lm(ws_ld(out.y), 0, 1, 1);
    st.push_back(s1);
    // x-pass C2R on the out.y rows -> out[j][pitch]
    Stage s2    = make_stage(StageKind::inv_c2r, dec ? "c2r_decrease_XY" : "c2r_none_XY", in.x, 0, out.x,
                             axis_origin(p, 0, in.x, out.x), 0, out.y);
    s2.src      = Buf::ws_b;
    s2.src_elem = Elem::cpx_f32;
    s2.im       = lm(1, 0, ws_ld(out.y), 1);
    s2.dst      = Buf::user_out;
    s2.dst_elem = Elem::real_input_type;
    s2.om       = lm(2 * out.w, 0, 1, 1);
    st.push_back(s2);
    return e;
}

static PlanError inv_stages_3d(const PlanState& p, std::vector<Stage>& st, Buf src, Elem src_elem) {
    PlanError   e;
    const Dims &in = p.inv_in, &out = p.inv_out;
    if ( p.inv_change == SizeChange::increase ) {
        e.code = FASTFFT_UNSUPPORTED;
        e.msg  = "3D FFT inv increase not yet supported (FastFFT.cu:589-594)";
        return e;
    }
    const bool dec = p.inv_change == SizeChange::decrease;
    const int  w   = in.w;
    if ( out.y < in.y && ! env_flag_planner("FASTFFT_INV3D_Z_FIRST") ) {
        // the y window is trimmed: the y-pass first, on the contiguous ky lines of the whole
        // input, so that the strided z-pass reads out.y / in.y of the input's volume (both
        // orders move the same bytes; this one moves the most of them with unit stride)
        // y-pass on the in.z*w lines -> T1[kz][kx][jy]
        Stage s1    = make_stage(StageKind::inv_cpx, "c2c_inv_decrease_XYZ", in.y, 0, out.y, axis_origin(p, 1, in.y, out.y),
                                 0, in.z * w);
        s1.src      = src;
        s1.src_elem = src_elem;
        s1.im       = lm(in.y, 0, 1, 1);
        s1.dst      = Buf::ws_a;
        s1.dst_elem = Elem::cpx_f32;
        s1.om       = lm(out.y, 0, 1, 1);
        st.push_back(s1);
        // z-pass on the w*out.y lines -> T2[jz][kx][jy]
        Stage s2    = make_stage(StageKind::inv_cpx, dec && out.z < in.z ? "c2c_inv_decrease_XZ" : "c2c_inv_none_XZ", in.z, 0,
                                 out.z, axis_origin(p, 2, in.z, out.z), 0, w * out.y);
        s2.src      = Buf::ws_a;
        s2.src_elem = Elem::cpx_f32;
        s2.im       = lm(1, 0, (long)w * out.y, 1);
        s2.dst      = Buf::ws_b;
        s2.dst_elem = Elem::cpx_f32;
        s2.om       = lm(1, 0, (long)w * out.y, 1);
        st.push_back(s2);
    }
    else {
        // z-pass on the w*in.y lines -> T1[jz][kx][ky] (out.z planes)
        Stage s1    = make_stage(StageKind::inv_cpx, dec ? "c2c_inv_decrease_XZ" : "c2c_inv_none_XZ", in.z, 0, out.z,
                                 axis_origin(p, 2, in.z, out.z), 0, w * in.y);
        s1.src      = src;
        s1.src_elem = src_elem;
        s1.im       = lm(1, 0, (long)w * in.y, 1);
        s1.dst      = Buf::ws_a;
        s1.dst_elem = Elem::cpx_f32;
        s1.om       = lm(1, 0, (long)w * in.y, 1);
        st.push_back(s1);
        // y-pass on the out.z*w lines -> T2[jz][kx][jy]
        Stage s2    = make_stage(StageKind::inv_cpx, dec ? "c2c_inv_decrease_XYZ" : "c2c_inv_none_XYZ", in.y, 0, out.y,
                                 axis_origin(p, 1, in.y, out.y), 0, out.z * w);
        s2.src      = Buf::ws_a;
        s2.src_elem = Elem::cpx_f32;
        s2.im       = lm(in.y, 0, 1, 1);
        s2.dst      = Buf::ws_b;
        s2.dst_elem = Elem::cpx_f32;
        s2.om       = lm(out.y, 0, 1, 1);
        st.push_back(s2);
    }
    // x-pass C2R on the out.z*out.y rows -> out[jz][jy][pitch]
    const int pitch = 2 * out.w;
    Stage     s3    = make_stage(StageKind::inv_c2r, dec ? "c2r_decrease" : "c2r_none", in.x, 0, out.x,
                                 axis_origin(p, 0, in.x, out.x), 0, out.z * out.y);
    s3.src          = Buf::ws_b;
    s3.src_elem     = Elem::cpx_f32;
    s3.im           = lm((long)w * out.y, 1, out.y, out.y);
    s3.dst          = Buf::user_out;
    s3.dst_elem     = Elem::real_input_type;
    s3.om           = lm((long)out.y * pitch, pitch, 1, out.y);
    st.push_back(s3);
    return e;
}

static PlanError xcorr_stages_2d(const PlanState& p, std::vector<Stage>& st) {
    PlanError e;
    if ( p.fwd_change == SizeChange::decrease ) {
        e.code = FASTFFT_UNSUPPORTED;
        e.msg  = "2D FFT cross correlation with a fwd size decrease is not supported";
        return e;
    }
    if ( p.inv_change == SizeChange::increase ) {
        e.code = FASTFFT_UNSUPPORTED;
        e.msg  = "2D FFT Cross correlation with an inv size increase is not supported (FastFFT.cu:648-652)";
        return e;
    }
    const Dims &in = p.fwd_in, &mid = p.fwd_out, &out = p.inv_out;
    const int   pitch = 2 * in.w;
    // The two intermediates are row-major, A[y][kx] (ldA) and B[j][kx] (ldB),
    // so that both x-passes read and write whole rows; the fused y-pass
    // gathers its template column from A and scatters its output column to B.
    const long ldA = ws_ld_rows(mid.w), ldB = ws_ld_rows(mid.w);
    Stage      s1  = make_stage(StageKind::fwd_real, p.fwd_change == SizeChange::increase ? "r2c_increase_XY" : "r2c_none_XY",
                                mid.x, in.x, 0, 0, mid.w, in.y);
    s1.src         = Buf::user_in;
    s1.src_elem    = Elem::real_input_type;
    s1.im          = lm(pitch, 0, 1, 1);
    s1.dst         = Buf::ws_a;
    s1.dst_elem    = Elem::cpx_f32;
    s1.om          = lm(ldA, 0, 1, 1);
    st.push_back(s1);
    const bool dec = p.inv_change == SizeChange::decrease;
    Stage      s2  = make_stage(StageKind::xcorr,
                                dec ? (p.fwd_change == SizeChange::increase ? "xcorr_fwd_increase_inv_decrease" : "xcorr_fwd_none_inv_decrease")
                                    : "generic_fwd_increase_op_inv_none",
                                mid.y, in.y, out.y, axis_origin(p, 1, mid.y, out.y), 0, mid.w);
    s2.src            = Buf::ws_a;
    s2.src_elem       = Elem::cpx_f32;
    s2.im             = lm(1, 0, ldA, 1);
    s2.gm             = lm(mid.y, 0, 1, 1);
    s2.dst            = Buf::ws_b;
    s2.dst_elem       = Elem::cpx_f32;
    s2.om             = lm(1, 0, ldB, 1);
    s2.real_end_lines = mid.x % 2 == 0 && mid.w >= 2;
    st.push_back(s2);
    Stage s3    = make_stage(StageKind::inv_c2r, dec ? "c2r_decrease_XY" : "c2r_none_XY", mid.x, 0, out.x,
                             axis_origin(p, 0, mid.x, out.x), 0, out.y);
    s3.src      = Buf::ws_b;
    s3.src_elem = Elem::cpx_f32;
    s3.im       = lm(ldB, 0, 1, 1);
    s3.dst      = Buf::user_out;
    s3.dst_elem = Elem::real_input_type;
    s3.om       = lm(2 * out.w, 0, 1, 1);
    st.push_back(s3);
    return e;
}

static PlanError xcorr_stages_3d(const PlanState& p, std::vector<Stage>& st) {
    PlanError e;
    if ( p.fwd_change == SizeChange::decrease || p.inv_change == SizeChange::increase ) {
        e.code = FASTFFT_UNSUPPORTED;
        e.msg  = "3D FFT cross correlation supports fwd none/increase with inv none/decrease";
        return e;
    }
    const Dims &in = p.fwd_in, &mid = p.fwd_out, &out = p.inv_out;
    const int   pitch = 2 * in.w, w = mid.w;
    Stage s1    = make_stage(StageKind::fwd_real, "r2c_increase_XZ", mid.x, in.x, 0, 0, w, in.z * in.y);
    s1.src      = Buf::user_in;
    s1.src_elem = Elem::real_input_type;
    s1.im       = lm((long)in.y * pitch, pitch, 1, in.y);
    s1.dst      = Buf::ws_a;
    s1.dst_elem = Elem::cpx_f32;
    s1.om       = lm((long)w * in.y, 1, in.y, in.y);
    st.push_back(s1);
    Stage s2    = make_stage(StageKind::fwd_cpx, "c2c_fwd_increase_XYZ", mid.y, in.y, 0, 0, mid.y, in.z * w);
    s2.src      = Buf::ws_a;
    s2.src_elem = Elem::cpx_f32;
    s2.im       = lm(in.y, 0, 1, 1);
    s2.dst      = Buf::ws_b;
    s2.dst_elem = Elem::cpx_f32;
    s2.om       = lm(mid.y, 0, 1, 1);
    st.push_back(s2);
    // fused z-pass: T2[z][kx][ky] (in.z planes) with image C[kz][kx][ky] -> T3[jz][kx][ky]
    Stage s3    = make_stage(StageKind::xcorr, "generic_fwd_increase_op_inv_none", mid.z, in.z, out.z,
                             axis_origin(p, 2, mid.z, out.z), 0, w * mid.y);
    s3.src      = Buf::ws_b;
    s3.src_elem = Elem::cpx_f32;
    s3.im       = lm(1, 0, (long)w * mid.y, 1);
    s3.gm       = lm(1, 0, (long)w * mid.y, 1);
    s3.dst      = Buf::ws_a;
    s3.dst_elem = Elem::cpx_f32;
    s3.om       = lm(1, 0, (long)w * mid.y, 1);
    st.push_back(s3);
    Stage s4    = make_stage(StageKind::inv_cpx, "c2c_inv_none_XYZ", mid.y, 0, out.y, axis_origin(p, 1, mid.y, out.y), 0,
                             out.z * w);
    s4.src      = Buf::ws_a;
    s4.src_elem = Elem::cpx_f32;
    s4.im       = lm(mid.y, 0, 1, 1);
    s4.dst      = Buf::ws_b;
    s4.dst_elem = Elem::cpx_f32;
    s4.om       = lm(out.y, 0, 1, 1);
    st.push_back(s4);
    const int opitch = 2 * out.w;
    Stage     s5     = make_stage(StageKind::inv_c2r, "c2r_none", mid.x, 0, out.x, axis_origin(p, 0, mid.x, out.x), 0,
                                  out.z * out.y);
    s5.src           = Buf::ws_b;
    s5.src_elem      = Elem::cpx_f32;
    s5.im            = lm((long)w * out.y, 1, out.y, out.y);
    s5.dst           = Buf::user_out;
    s5.dst_elem      = Elem::real_input_type;
    s5.om            = lm((long)out.y * opitch, opitch, 1, out.y);
    st.push_back(s5);
    return e;
}

PlanError build_stages(const PlanState& p, const std::string& op, std::vector<Stage>& st) {
    PlanError e;
    st.clear( );
    if ( ! p.validated ) {
        e.code = FASTFFT_NOT_PLANNED;
        e.msg  = "plan not validated";
        return e;
    }
    const bool r3 = p.transform_dimension == 3;
    if ( op == "fwd" )
        return r3 ? fwd_stages_3d(p, st) : fwd_stages_2d(p, st, Buf::user_out);
    if ( op == "inv" )
        return r3 ? inv_stages_3d(p, st, Buf::user_in, Elem::cpx_input_type) : inv_stages_2d(p, st, Buf::user_in, Elem::cpx_input_type);
    if ( op == "xcorr" )
        return r3 ? xcorr_stages_3d(p, st) : xcorr_stages_2d(p, st);
    e.code = FASTFFT_INVALID_ARGUMENT;
    e.msg  = "unknown operation '" + op + "' (expected fwd, inv or xcorr)";
    return e;
}

// Elements touched in a buffer by a stage: max over lines and elements of the address + 1.
static size_t extent(const HostLineMap& m, int nlines, int nelem) {
    if ( nlines <= 0 || nelem <= 0 )
        return 0;
    long last_line = nlines - 1;
    long b         = (last_line / m.n2) * m.s1 + (last_line % m.n2) * m.s2;
    // the largest base is not always the last line when s2 > s1; take the max corner
    long b2 = (m.n2 > 1) ? ((last_line / m.n2) * m.s1 + (long)(m.n2 - 1) * m.s2) : b;
    return (size_t)(std::max(b, b2) + (long)(nelem - 1) * m.es + 1);
}

void stage_list_workspace(const std::vector<Stage>& st, size_t& ws_a, size_t& ws_b) {
    for ( const Stage& s : st ) {
        int    nw = s.kind == StageKind::fwd_real || s.kind == StageKind::fwd_cpx ? s.nout : s.M;
        size_t ex = extent(s.om, s.nlines, nw);
        if ( s.dst == Buf::ws_a )
            ws_a = std::max(ws_a, ex);
        if ( s.dst == Buf::ws_b )
            ws_b = std::max(ws_b, ex);
    }
}

void workspace_elems(const PlanState& p, size_t& ws_a, size_t& ws_b) {
    ws_a = ws_b = 0;
    const char* ops[3] = {"fwd", "inv", "xcorr"};
    for ( const char* op : ops ) {
        std::vector<Stage> st;
        if ( build_stages(p, op, st).code != 0 )
            continue;
        stage_list_workspace(st, ws_a, ws_b);
    }
}

static const char* kind_name(StageKind k) {
    switch ( k ) {
        case StageKind::fwd_real: return "fwd_real";
        case StageKind::fwd_cpx: return "fwd_cpx";
        case StageKind::inv_cpx: return "inv_cpx";
        case StageKind::inv_c2r: return "inv_c2r";
        case StageKind::xcorr: return "xcorr";
    }
    return "?";
}

static const char* buf_name(Buf b) {
    switch ( b ) {
        case Buf::user_in: return "in";
        case Buf::user_out: return "out";
        case Buf::image: return "image";
        case Buf::ws_a: return "ws_a";
        case Buf::ws_b: return "ws_b";
    }
    return "?";
}

std::string describe(const PlanState& p, const std::vector<Stage>& stages) {
    std::ostringstream os;
    os << "rank=" << p.transform_dimension << " fwd=" << (int)p.fwd_change << " inv=" << (int)p.inv_change << "\n";
    for ( const Stage& s : stages ) {
        os << kind_name(s.kind) << " ref=" << s.ref_kernel << " N=" << s.N << " P=" << s.P << " Q=" << s.Q << " L=" << s.L
           << " M=" << s.M << " o=" << s.o << " nout=" << s.nout << " lines=" << s.nlines << " " << buf_name(s.src) << "->"
           << buf_name(s.dst) << " im=(" << s.im.s1 << "," << s.im.s2 << "," << s.im.es << "," << s.im.n2 << ")"
           << " om=(" << s.om.s1 << "," << s.om.s2 << "," << s.om.es << "," << s.om.n2 << ")"
           << " gm=(" << s.gm.s1 << "," << s.gm.s2 << "," << s.gm.es << "," << s.gm.n2 << ")\n";
    }
    return os.str( );
}

} // namespace fastfft_amd
