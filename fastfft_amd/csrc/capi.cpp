// C ABI of the FastFFT path: plan object, execution of planned stages,
// host memory helpers.  Host code only; kernels live in launch_*.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fastfft/fastfft.h"
#include "launch.hpp"
#include "plan.hpp"

using namespace fastfft_amd;

namespace {

thread_local std::string g_last_error;

fastfft_status fail(int code, const std::string& msg) {
    g_last_error = msg;
    return static_cast<fastfft_status>(code);
}

fastfft_status ok( ) { return FASTFFT_SUCCESS; }

fastfft_status hip_fail(hipError_t e, const char* what) {
    return fail(FASTFFT_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

constexpr size_t kMaxLdsBytes = 160 * 1024; // gfx950 LDS per CU
constexpr int    kMaxDevices  = 64;

// Per-device FP32 twiddle table, generated once in FP64 and kept until exit.
std::mutex g_tw_mutex;
float2*    g_tw[kMaxDevices] = {nullptr};

hipError_t twiddles_for_current_device(float2** out, hipStream_t stream) {
    int        dev = 0;
    hipError_t e   = hipGetDevice(&dev);
    if ( e != hipSuccess )
        return e;
    if ( dev < 0 || dev >= kMaxDevices )
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    if ( ! g_tw[dev] ) {
        float2* tw = nullptr;
        e          = hipMalloc(&tw, sizeof(float2) * kTwiddleTableN);
        if ( e != hipSuccess )
            return e;
        e = launch_twiddle_init(tw, stream);
        if ( e == hipSuccess )
            e = hipStreamSynchronize(stream);
        if ( e != hipSuccess ) {
            (void)hipFree(tw);
            return e;
        }
        g_tw[dev] = tw;
    }
    *out = g_tw[dev];
    return hipSuccess;
}

} // namespace

struct fastfft_plan_s {
    PlanState   st;
    hipStream_t stream = hipStreamPerThread;
    void*       ws     = nullptr; // workspace: ws_a then ws_b (float2)
    size_t      ws_a = 0, ws_b = 0;
    int         ws_device = -1;
    void*       argmax_scratch = nullptr; // argmax row keys, partials and result slots
    // CopyDeviceToHostAndSynchronize bookkeeping (FastFFT.cu:947-1009)
    void* last_buffer     = nullptr;
    int   stage_completed = 0; // 0: none, 3: forward, 7: inverse / correlation
    // stage lists per operation (fwd, inv, xcorr), built once per validated plan
    std::vector<Stage> stages[3];
    bool               stages_built[3] = {false, false, false};
    float2*            tw              = nullptr; // twiddle table of ws_device
};

static int op_index(const char* op) {
    if ( ! op )
        return -1;
    if ( std::strcmp(op, "fwd") == 0 )
        return 0;
    if ( std::strcmp(op, "inv") == 0 )
        return 1;
    if ( std::strcmp(op, "xcorr") == 0 )
        return 2;
    return -1;
}

static const char* kOpNames[3] = {"fwd", "inv", "xcorr"};

static void invalidate(fastfft_plan p) {
    p->st.validated = false;
    for ( int i = 0; i < 3; i++ ) {
        p->stages_built[i] = false;
        p->stages[i].clear( );
    }
}

static size_t elem_bytes(const PlanState& s) { return s.input_type == FASTFFT_INPUT_F16 ? 2 : 4; }

static fastfft_status check_plan(fastfft_plan p) {
    if ( ! p )
        return fail(FASTFFT_INVALID_ARGUMENT, "null plan handle");
    return ok( );
}

static fastfft_status ensure_validated(fastfft_plan p) {
    if ( p->st.validated )
        return ok( );
    PlanError e = validate(p->st);
    if ( e.code )
        return fail(e.code, e.msg);
    return ok( );
}

static fastfft_status ensure_device_state(fastfft_plan p, float2** tw) {
    int        dev = 0;
    hipError_t e   = hipGetDevice(&dev);
    if ( e != hipSuccess )
        return hip_fail(e, "hipGetDevice");
    if ( p->ws && p->ws_device == dev && p->tw ) {
        *tw = p->tw;
        return ok( );
    }
    e = twiddles_for_current_device(tw, p->stream);
    if ( e != hipSuccess )
        return hip_fail(e, "twiddle table");
    p->tw = *tw;
    if ( ! p->ws || p->ws_device != dev ) {
        if ( p->ws ) {
            (void)hipFree(p->ws);
            p->ws = nullptr;
        }
        workspace_elems(p->st, p->ws_a, p->ws_b);
        size_t bytes = sizeof(float2) * (p->ws_a + p->ws_b);
        if ( bytes < 256 )
            bytes = 256;
        e = hipMalloc(&p->ws, bytes);
        if ( e != hipSuccess ) {
            p->ws = nullptr;
            return e == hipErrorOutOfMemory ? fail(FASTFFT_OUT_OF_MEMORY, "workspace allocation failed")
                                            : hip_fail(e, "hipMalloc(workspace)");
        }
        p->ws_device = dev;
    }
    return ok( );
}

static fastfft_status op_kind_check(const fastfft_op& op, bool allow_scale, bool allow_conj, const char* which) {
    if ( op.kind == FASTFFT_OP_NOOP )
        return ok( );
    if ( op.kind == FASTFFT_OP_SCALE && allow_scale )
        return ok( );
    if ( (op.kind == FASTFFT_OP_CONJ_MUL || op.kind == FASTFFT_OP_CONJ_MUL_THEN_SCALE) && allow_conj )
        return ok( );
    return fail(FASTFFT_INVALID_ARGUMENT, std::string("operator kind ") + std::to_string(op.kind) + " is not valid as " + which);
}

static float op_scale(const fastfft_op& op) {
    return (op.kind == FASTFFT_OP_SCALE || op.kind == FASTFFT_OP_CONJ_MUL_THEN_SCALE) ? op.scale : 1.f;
}

static fastfft_status get_stages(fastfft_plan p, int opi, const std::vector<Stage>** out) {
    if ( ! p->stages_built[opi] ) {
        std::vector<Stage> st;
        PlanError          pe = build_stages(p->st, kOpNames[opi], st);
        if ( pe.code )
            return fail(pe.code, pe.msg);
        for ( const Stage& s : st )
            if ( stage_lds_bytes(s.P) > kMaxLdsBytes )
                return fail(FASTFFT_LDS_EXCEEDED, "stage " + s.ref_kernel + " needs more LDS than a gfx950 CU has");
        p->stages[opi]       = std::move(st);
        p->stages_built[opi] = true;
    }
    *out = &p->stages[opi];
    return ok( );
}

static fastfft_status run_stages(fastfft_plan p, int opi, void* in, void* out, const void* image, float pre, float intra,
                                 float post, int first = 0, int last = -1, const ArgmaxSlots* amax = nullptr,
                                 bool* amax_fused = nullptr) {
    const std::vector<Stage>* sp = nullptr;
    if ( auto rc = get_stages(p, opi, &sp) )
        return rc;
    const std::vector<Stage>& stages = *sp;
    if ( last < 0 )
        last = (int)stages.size( ) - 1;
    if ( first < 0 || last >= (int)stages.size( ) || first > last )
        return fail(FASTFFT_INVALID_ARGUMENT,
                    "stage range outside the " + std::to_string(stages.size( )) + " stages of " + kOpNames[opi]);
    float2*        tw = nullptr;
    fastfft_status rc = ensure_device_state(p, &tw);
    if ( rc )
        return rc;
    float2* wa = static_cast<float2*>(p->ws);
    float2* wb = wa + p->ws_a;
    auto    ptr = [&](Buf b) -> void* {
        switch ( b ) {
            case Buf::user_in: return in;
            case Buf::user_out: return out;
            case Buf::image: return const_cast<void*>(image);
            case Buf::ws_a: return wa;
            case Buf::ws_b: return wb;
        }
        return nullptr;
    };
    const bool half_in  = p->st.input_type == FASTFFT_INPUT_F16;
    const bool half_img = p->st.image_type == FASTFFT_IMAGE_C32H;
    auto       set_elem = [&](Elem el, bool& is_real, RealT& rt, CpxT& ct) {
        is_real = el == Elem::real_input_type;
        rt      = half_in ? RealT::f16 : RealT::f32;
        ct      = el == Elem::cpx_f32 ? CpxT::c64 : el == Elem::cpx_image_type ? (half_img ? CpxT::c32h : CpxT::c64)
                                                                            : (half_in ? CpxT::c32h : CpxT::c64);
    };
    for ( size_t i = (size_t)first; i <= (size_t)last; i++ ) {
        const Stage& s = stages[i];
        StageIO      io;
        io.src = ptr(s.src);
        io.dst = ptr(s.dst);
        io.img = image;
        set_elem(s.src_elem, io.src_real, io.src_rt, io.src_ct);
        set_elem(s.dst_elem, io.dst_real, io.dst_rt, io.dst_ct);
        io.img_ct = half_img ? CpxT::c32h : CpxT::c64;
        hipError_t e;
        switch ( s.kind ) {
            case StageKind::fwd_real:
            case StageKind::fwd_cpx:
                io.scale = (i == 0) ? pre : 1.f;
                e        = launch_fwd(s, io, tw, p->stream);
                break;
            case StageKind::inv_cpx:
            case StageKind::inv_c2r:
                io.scale = (i + 1 == stages.size( )) ? post : 1.f;
                if ( s.kind == StageKind::inv_c2r && i + 1 == stages.size( ) ) {
                    io.amax       = amax;
                    io.amax_fused = amax_fused;
                }
                e = launch_inv(s, io, tw, p->stream);
                break;
            case StageKind::xcorr:
                io.scale = intra;
                e        = launch_xcorr(s, io, tw, p->stream);
                break;
            default: e = hipErrorInvalidValue;
        }
        if ( e != hipSuccess )
            return hip_fail(e, ("launch of stage " + s.ref_kernel).c_str( ));
    }
    return ok( );
}

// Device scratch of the argmax epilogues:
//   kArgmaxPartials 64-bit row keys of the argmax computed by the C2R pass,
//   then the partial values and indices of the two-kernel epilogue and its result slots.
struct ArgmaxScratch {
    ArgmaxSlots fused;
    float*      pv  = nullptr;
    int*        pi  = nullptr;
    float*      rv  = nullptr;
    int*        rxy = nullptr;
};

static fastfft_status argmax_scratch(fastfft_plan plan, ArgmaxScratch& a) {
    const size_t keys  = sizeof(unsigned long long) * kArgmaxPartials;
    const size_t bytes = keys + sizeof(float) * (kArgmaxPartials + 1) + sizeof(int) * (kArgmaxPartials + 2);
    if ( ! plan->argmax_scratch ) {
        hipError_t e = hipMalloc(&plan->argmax_scratch, bytes);
        if ( e != hipSuccess ) {
            plan->argmax_scratch = nullptr;
            return hip_fail(e, "hipMalloc(argmax scratch)");
        }
    }
    char* base   = static_cast<char*>(plan->argmax_scratch);
    a.fused.keys = reinterpret_cast<unsigned long long*>(base);
    a.pv         = reinterpret_cast<float*>(base + keys);
    a.rv         = a.pv + kArgmaxPartials;
    a.pi         = reinterpret_cast<int*>(a.rv + 1);
    a.rxy        = a.pi + kArgmaxPartials;
    return ok( );
}

extern "C" {

fastfft_status fastfft_create(fastfft_plan* plan, int input_type, int image_type, int rank) {
    if ( ! plan )
        return fail(FASTFFT_INVALID_ARGUMENT, "null output handle");
    *plan = nullptr;
    if ( input_type != FASTFFT_INPUT_F32 && input_type != FASTFFT_INPUT_F16 )
        return fail(FASTFFT_INVALID_ARGUMENT, "input type must be F32 or F16");
    if ( image_type != FASTFFT_IMAGE_C64 && image_type != FASTFFT_IMAGE_C32H )
        return fail(FASTFFT_INVALID_ARGUMENT, "image type must be C64 (float2) or C32H (__half2)");
    if ( rank != 2 && rank != 3 )
        return fail(FASTFFT_INVALID_ARGUMENT, "Only 2D and 3D FFTs are supported");
    fastfft_plan p     = new (std::nothrow) fastfft_plan_s;
    if ( ! p )
        return fail(FASTFFT_OUT_OF_MEMORY, "plan allocation failed");
    p->st.input_type = input_type;
    p->st.image_type = image_type;
    p->st.rank       = rank;
    *plan            = p;
    return ok( );
}

fastfft_status fastfft_destroy(fastfft_plan plan) {
    if ( ! plan )
        return ok( );
    if ( plan->ws )
        (void)hipFree(plan->ws);
    if ( plan->argmax_scratch )
        (void)hipFree(plan->argmax_scratch);
    delete plan;
    return ok( );
}

static fastfft_status set_dims(fastfft_plan plan, Dims& din, Dims& dout, size_t ix, size_t iy, size_t iz, size_t ox, size_t oy,
                               size_t oz) {
    const size_t lim = 1u << 15; // dims are short4 in the reference
    if ( ix == 0 || iy == 0 || iz == 0 || ox == 0 || oy == 0 || oz == 0 )
        return fail(FASTFFT_INVALID_DIMENSION, "logical dimensions must be > 0");
    if ( ix >= lim || iy >= lim || iz >= lim || ox >= lim || oy >= lim || oz >= lim )
        return fail(FASTFFT_INVALID_DIMENSION, "logical dimensions must fit in a short");
    din.x  = (int)ix;
    din.y  = (int)iy;
    din.z  = (int)iz;
    dout.x = (int)ox;
    dout.y = (int)oy;
    dout.z = (int)oz;
    invalidate(plan);
    return ok( );
}

fastfft_status fastfft_set_fwd_plan(fastfft_plan plan, size_t in_x, size_t in_y, size_t in_z, size_t out_x, size_t out_y,
                                    size_t out_z, int is_padded_input) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( auto rc = set_dims(plan, plan->st.fwd_in, plan->st.fwd_out, in_x, in_y, in_z, out_x, out_y, out_z) )
        return rc;
    plan->st.padded            = is_padded_input != 0;
    plan->st.input_memory      = padded_memory_size(plan->st.fwd_in);
    plan->st.fwd_output_memory = padded_memory_size(plan->st.fwd_out);
    plan->st.compute_memory    = std::max(plan->st.compute_memory, std::max(plan->st.input_memory, plan->st.fwd_output_memory));
    plan->st.fwd_set           = true;
    if ( ! plan->st.padded )
        return fail(FASTFFT_UNSUPPORTED, "Support for input arrays that are not FFTW padded needs to be implemented.");
    return ok( );
}

fastfft_status fastfft_set_inv_plan(fastfft_plan plan, size_t in_x, size_t in_y, size_t in_z, size_t out_x, size_t out_y,
                                    size_t out_z, int is_padded_output) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( plan->st.fwd_set && (is_padded_output != 0) != plan->st.padded )
        return fail(FASTFFT_INVALID_ARGUMENT, "If the input data are FFTW padded, so must the output.");
    if ( auto rc = set_dims(plan, plan->st.inv_in, plan->st.inv_out, in_x, in_y, in_z, out_x, out_y, out_z) )
        return rc;
    long in_mem                = padded_memory_size(plan->st.inv_in);
    plan->st.inv_output_memory = padded_memory_size(plan->st.inv_out);
    plan->st.compute_memory    = std::max(plan->st.compute_memory, std::max(in_mem, plan->st.inv_output_memory));
    plan->st.inv_set           = true;
    return ok( );
}

fastfft_status fastfft_set_output_window(fastfft_plan plan, int origin_x, int origin_y, int origin_z) {
    if ( auto rc = check_plan(plan) )
        return rc;
    plan->st.origin[0] = origin_x;
    plan->st.origin[1] = origin_y;
    plan->st.origin[2] = origin_z;
    invalidate(plan);
    return ok( );
}

fastfft_status fastfft_set_stream(fastfft_plan plan, void* stream) {
    if ( auto rc = check_plan(plan) )
        return rc;
    plan->stream = stream ? static_cast<hipStream_t>(stream) : hipStreamPerThread;
    return ok( );
}

fastfft_status fastfft_fwd(fastfft_plan plan, void* in, void* out, fastfft_op pre_op, fastfft_op intra_op) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( ! in )
        return fail(FASTFFT_INVALID_ARGUMENT, "null input pointer");
    if ( auto rc = op_kind_check(pre_op, true, false, "the FwdFFT pre op") )
        return rc;
    if ( auto rc = op_kind_check(intra_op, false, false, "the FwdFFT intra op") )
        return rc;
    if ( auto rc = ensure_validated(plan) )
        return rc;
    void* dst = out ? out : in;
    if ( auto rc = run_stages(plan, 0, in, dst, nullptr, op_scale(pre_op), 1.f, 1.f) )
        return rc;
    plan->last_buffer     = dst;
    plan->stage_completed = 3;
    return ok( );
}

fastfft_status fastfft_inv(fastfft_plan plan, void* in, void* out, fastfft_op intra_op, fastfft_op post_op) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( ! in )
        return fail(FASTFFT_INVALID_ARGUMENT, "null input pointer");
    if ( auto rc = op_kind_check(intra_op, false, false, "the InvFFT intra op") )
        return rc;
    if ( auto rc = op_kind_check(post_op, true, false, "the InvFFT post op") )
        return rc;
    if ( auto rc = ensure_validated(plan) )
        return rc;
    void* dst = out ? out : in;
    if ( auto rc = run_stages(plan, 1, in, dst, nullptr, 1.f, 1.f, op_scale(post_op)) )
        return rc;
    plan->last_buffer     = dst;
    plan->stage_completed = 7;
    return ok( );
}

fastfft_status fastfft_fwd_image_inv(fastfft_plan plan, void* in, const void* image_to_search, void* out, fastfft_op pre_op,
                                     fastfft_op intra_op, fastfft_op post_op) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( ! in || ! image_to_search )
        return fail(FASTFFT_INVALID_ARGUMENT, "null input or image_to_search pointer");
    if ( auto rc = op_kind_check(pre_op, true, false, "the FwdImageInvFFT pre op") )
        return rc;
    if ( intra_op.kind != FASTFFT_OP_CONJ_MUL && intra_op.kind != FASTFFT_OP_CONJ_MUL_THEN_SCALE )
        return fail(FASTFFT_INVALID_ARGUMENT, "FwdImageInvFFT needs an intra op (CONJ_MUL or CONJ_MUL_THEN_SCALE)");
    if ( auto rc = op_kind_check(post_op, true, false, "the FwdImageInvFFT post op") )
        return rc;
    if ( auto rc = ensure_validated(plan) )
        return rc;
    void* dst = out ? out : in;
    if ( auto rc = run_stages(plan, 2, in, dst, image_to_search, op_scale(pre_op), op_scale(intra_op), op_scale(post_op)) )
        return rc;
    plan->last_buffer     = dst;
    plan->stage_completed = 7;
    return ok( );
}

fastfft_status fastfft_fwd_image_inv_argmax(fastfft_plan plan, void* in, const void* image_to_search, void* out,
                                            fastfft_op pre_op, fastfft_op intra_op, fastfft_op post_op, int* peak_xy,
                                            float* value) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( ! in || ! image_to_search || ! peak_xy || ! value )
        return fail(FASTFFT_INVALID_ARGUMENT, "null input, image_to_search, peak_xy or value pointer");
    if ( auto rc = op_kind_check(pre_op, true, false, "the FwdImageInvFFT pre op") )
        return rc;
    if ( intra_op.kind != FASTFFT_OP_CONJ_MUL && intra_op.kind != FASTFFT_OP_CONJ_MUL_THEN_SCALE )
        return fail(FASTFFT_INVALID_ARGUMENT, "FwdImageInvFFT needs an intra op (CONJ_MUL or CONJ_MUL_THEN_SCALE)");
    if ( auto rc = op_kind_check(post_op, true, false, "the FwdImageInvFFT post op") )
        return rc;
    if ( auto rc = ensure_validated(plan) )
        return rc;
    if ( plan->st.transform_dimension != 2 )
        return fail(FASTFFT_UNSUPPORTED, "FwdImageInvFFT with argmax is implemented for 2D plans");
    ArgmaxScratch a;
    if ( auto rc = argmax_scratch(plan, a) )
        return rc;
    a.fused.value = value;
    a.fused.xy    = peak_xy;
    void* dst     = out ? out : in;
    bool  fused   = false;
    if ( auto rc = run_stages(plan, 2, in, dst, image_to_search, op_scale(pre_op), op_scale(intra_op), op_scale(post_op), 0, -1,
                              &a.fused, &fused) )
        return rc;
    plan->last_buffer     = dst;
    plan->stage_completed = 7;
    if ( ! fused ) {
        const Dims& o = plan->st.inv_out;
        hipError_t  e = launch_argmax(dst, plan->st.input_type == FASTFFT_INPUT_F16 ? RealT::f16 : RealT::f32, o.x, o.y,
                                      2 * o.w, a.pv, a.pi, value, peak_xy, plan->stream);
        if ( e != hipSuccess )
            return hip_fail(e, "argmax");
    }
    return ok( );
}

fastfft_status fastfft_wait(fastfft_plan plan) {
    if ( auto rc = check_plan(plan) )
        return rc;
    hipError_t e = hipStreamSynchronize(plan->stream);
    return e == hipSuccess ? ok( ) : hip_fail(e, "hipStreamSynchronize");
}

fastfft_status fastfft_copy_device_to_host_and_synchronize(fastfft_plan plan, void* host, int n_elements) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( ! host )
        return fail(FASTFFT_INVALID_ARGUMENT, "null host pointer");
    if ( ! plan->last_buffer )
        return fail(FASTFFT_INVALID_ARGUMENT, "no transform has run, so no device buffer is current");
    long n = n_elements;
    if ( n <= 0 )
        n = plan->stage_completed < 5 ? plan->st.fwd_output_memory : plan->st.inv_output_memory;
    hipError_t e = hipMemcpyAsync(host, plan->last_buffer, (size_t)n * elem_bytes(plan->st), hipMemcpyDeviceToHost, plan->stream);
    if ( e == hipSuccess )
        e = hipStreamSynchronize(plan->stream);
    return e == hipSuccess ? ok( ) : hip_fail(e, "CopyDeviceToHostAndSynchronize");
}

fastfft_status fastfft_get_memory_sizes(fastfft_plan plan, int* input, int* fwd_output, int* inv_output) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( input )
        *input = (int)plan->st.input_memory;
    if ( fwd_output )
        *fwd_output = (int)plan->st.fwd_output_memory;
    if ( inv_output )
        *inv_output = (int)plan->st.inv_output_memory;
    return ok( );
}

fastfft_status fastfft_get_dims(fastfft_plan plan, int which, int dims[4]) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( ! dims )
        return fail(FASTFFT_INVALID_ARGUMENT, "null dims");
    const Dims* d = nullptr;
    switch ( which ) {
        case FASTFFT_FWD_INPUT: d = &plan->st.fwd_in; break;
        case FASTFFT_FWD_OUTPUT: d = &plan->st.fwd_out; break;
        case FASTFFT_INV_INPUT: d = &plan->st.inv_in; break;
        case FASTFFT_INV_OUTPUT: d = &plan->st.inv_out; break;
        default: return fail(FASTFFT_INVALID_ARGUMENT, "unknown dims selector");
    }
    dims[0] = d->x;
    dims[1] = d->y;
    dims[2] = d->z;
    dims[3] = d->w;
    return ok( );
}

fastfft_status fastfft_get_buffer_pointer(fastfft_plan plan, void** ptr) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( ! ptr )
        return fail(FASTFFT_INVALID_ARGUMENT, "null output pointer");
    *ptr = plan->ws;
    return ok( );
}

fastfft_status fastfft_workspace_bytes(fastfft_plan plan, size_t* bytes) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( ! bytes )
        return fail(FASTFFT_INVALID_ARGUMENT, "null output pointer");
    if ( auto rc = ensure_validated(plan) )
        return rc;
    size_t a = 0, b = 0;
    workspace_elems(plan->st, a, b);
    *bytes = sizeof(float2) * (a + b);
    return ok( );
}

fastfft_status fastfft_argmax(fastfft_plan plan, const void* img, int nx, int ny, int pitch, int peak[2], float* value) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( ! img || ! peak || nx <= 0 || ny <= 0 || pitch < nx )
        return fail(FASTFFT_INVALID_ARGUMENT, "argmax: null pointer or bad extent");
    ArgmaxScratch a;
    if ( auto rc = argmax_scratch(plan, a) )
        return rc;
    float*     rv  = a.rv;
    int*       rxy = a.rxy;
    hipError_t e   = launch_argmax(img, plan->st.input_type == FASTFFT_INPUT_F16 ? RealT::f16 : RealT::f32, nx, ny, pitch,
                                   a.pv, a.pi, rv, rxy, plan->stream);
    if ( e != hipSuccess )
        return hip_fail(e, "argmax launch");
    float hv = 0.f;
    int   hxy[2];
    e = hipMemcpyAsync(&hv, rv, sizeof(float), hipMemcpyDeviceToHost, plan->stream);
    if ( e == hipSuccess )
        e = hipMemcpyAsync(hxy, rxy, 2 * sizeof(int), hipMemcpyDeviceToHost, plan->stream);
    if ( e == hipSuccess )
        e = hipStreamSynchronize(plan->stream);
    if ( e != hipSuccess )
        return hip_fail(e, "argmax copy");
    peak[0] = hxy[0];
    peak[1] = hxy[1];
    if ( value )
        *value = hv;
    return ok( );
}

fastfft_status fastfft_argmax_device(fastfft_plan plan, const void* img, int nx, int ny, int pitch, int* peak_xy, float* value) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( ! img || ! peak_xy || ! value || nx <= 0 || ny <= 0 || pitch < nx )
        return fail(FASTFFT_INVALID_ARGUMENT, "argmax: null pointer or bad extent");
    ArgmaxScratch a;
    if ( auto rc = argmax_scratch(plan, a) )
        return rc;
    hipError_t e = launch_argmax(img, plan->st.input_type == FASTFFT_INPUT_F16 ? RealT::f16 : RealT::f32, nx, ny, pitch, a.pv,
                                 a.pi, value, peak_xy, plan->stream);
    return e == hipSuccess ? ok( ) : hip_fail(e, "argmax");
}

fastfft_status fastfft_run_stages(fastfft_plan plan, const char* op, int first, int last, void* in, const void* image_to_search,
                                  void* out, fastfft_op pre_op, fastfft_op intra_op, fastfft_op post_op) {
    if ( auto rc = check_plan(plan) )
        return rc;
    const int opi = op_index(op);
    if ( opi < 0 || ! in )
        return fail(FASTFFT_INVALID_ARGUMENT, "unknown op (expected fwd, inv or xcorr) or null input pointer");
    if ( opi == 2 && ! image_to_search )
        return fail(FASTFFT_INVALID_ARGUMENT, "xcorr stages need image_to_search");
    if ( opi == 2 && intra_op.kind != FASTFFT_OP_CONJ_MUL && intra_op.kind != FASTFFT_OP_CONJ_MUL_THEN_SCALE )
        return fail(FASTFFT_INVALID_ARGUMENT, "xcorr stages need an intra op (CONJ_MUL or CONJ_MUL_THEN_SCALE)");
    if ( auto rc = op_kind_check(pre_op, true, false, "a pre op") )
        return rc;
    if ( auto rc = op_kind_check(post_op, true, false, "a post op") )
        return rc;
    if ( auto rc = ensure_validated(plan) )
        return rc;
    void* dst = out ? out : in;
    return run_stages(plan, opi, in, dst, image_to_search, op_scale(pre_op), op_scale(intra_op), op_scale(post_op), first, last);
}

fastfft_status fastfft_describe(fastfft_plan plan, const char* op, char* buf, size_t buf_len) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( ! op || ! buf || buf_len == 0 )
        return fail(FASTFFT_INVALID_ARGUMENT, "describe: null argument");
    if ( auto rc = ensure_validated(plan) )
        return rc;
    const int opi = op_index(op);
    if ( opi < 0 )
        return fail(FASTFFT_INVALID_ARGUMENT, std::string("unknown operation '") + op + "' (expected fwd, inv or xcorr)");
    const std::vector<Stage>* sp = nullptr;
    if ( auto rc = get_stages(plan, opi, &sp) )
        return rc;
    std::string s = describe(plan->st, *sp);
    std::snprintf(buf, buf_len, "%s", s.c_str( ));
    if ( s.size( ) + 1 > buf_len )
        return fail(FASTFFT_INVALID_ARGUMENT, "describe: buffer too small");
    return ok( );
}

const char* fastfft_last_error_string(void) { return g_last_error.c_str( ); }

const char* fastfft_version(void) { return "fastfft_amd 0.1.0 (gfx950)"; }

} // extern "C"
