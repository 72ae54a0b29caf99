// C ABI of the FastFFT path: plan object, execution of planned stages,
// host memory helpers.  Host code only; kernels live in launch_*.hip.
#include <hip/hip_runtime.h>

#include <cxxabi.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fastfft/fastfft.h"
#include "launch.hpp"
#include "plan.hpp"
#include "tuning.hpp"

using namespace fastfft_amd;

namespace fastfft_amd {

bool log_enabled( ) {
    static const bool on = [] {
        const char* v = std::getenv("FASTFFT_LOG");
        return v && v[0] && std::atoi(v) != 0;
    }( );
    return on;
}

void log_text(const std::string& text) {
    static std::mutex m;
    std::lock_guard<std::mutex> lock(m);
    std::fprintf(stderr, "fastfft: %s%s", text.c_str( ), (! text.empty( ) && text.back( ) == '\n') ? "" : "\n");
    std::fflush(stderr);
}

// The demangled tag reads "fastfft_amd::KernelTag<&(void fastfft_amd::k<args>(params))>"
// for a kernel template and "fastfft_amd::KernelTag<&fastfft_amd::k>" otherwise; the
// kernel's name with its template arguments is what follows the '&', without the
// return type and the parameter list.
void log_launch(const char* tag, dim3 grid, dim3 block, size_t lds) {
    int         status = 0;
    char*       d      = abi::__cxa_demangle(tag, nullptr, nullptr, &status);
    std::string name   = d ? d : tag;
    std::free(d);
    const size_t amp = name.find('&');
    if ( amp != std::string::npos && amp + 1 < name.size( ) && name.back( ) == '>' ) {
        if ( name[amp + 1] == '(' ) {
            size_t b = amp + 2;
            if ( name.compare(b, 5, "void ") == 0 )
                b += 5;
            int    depth = 0;
            size_t e     = b;
            while ( e < name.size( ) && ! (name[e] == '(' && depth == 0) ) {
                depth += name[e] == '<' ? 1 : name[e] == '>' ? -1 : 0;
                e++;
            }
            name = name.substr(b, e - b);
        }
        else {
            name = name.substr(amp + 1, name.size( ) - amp - 2);
        }
    }
    for ( size_t at; (at = name.find("fastfft_amd::")) != std::string::npos; )
        name.erase(at, 13);
    static const char* const kVectorNames[][2] = {{"HIP_vector_type<float, 2u>", "float2"},
                                                  {"HIP_vector_type<__half, 2u>", "__half2"}};
    for ( const auto& r : kVectorNames )
        for ( size_t at; (at = name.find(r[0])) != std::string::npos; )
            name.replace(at, std::strlen(r[0]), r[1]);
    for ( size_t at; (at = name.find(" >")) != std::string::npos; ) // "float2 >" of a nested "<...> >"
        name.erase(at, 1);
    char cfg[128];
    std::snprintf(cfg, sizeof(cfg), " grid=(%u,%u,%u) block=%u lds=%zu", grid.x, grid.y, grid.z, block.x * block.y * block.z,
                  lds);
    log_text("launch " + name + cfg);
}

} // namespace fastfft_amd

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

// Device allocation of a plan's internal buffers (workspace, argmax scratch,
// batch slots and frame).  With FASTFFT_POISON=1 (a debugging switch the GPU
// tests turn on) every new buffer is filled with 0xff bytes, a NaN in every
// float and float2 slot and -1 in every int, so that a kernel that reads
// workspace nobody wrote puts NaN into its outputs.  Never called while a
// stream is capturing (allocation is not allowed there either).
hipError_t dev_alloc(void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if ( e != hipSuccess ) {
        *p = nullptr;
        return e;
    }
    const char* v = std::getenv("FASTFFT_POISON");
    if ( v && std::atoi(v) != 0 ) {
        e = hipMemset(*p, 0xff, bytes);
        if ( e == hipSuccess )
            e = hipDeviceSynchronize( );
        if ( e != hipSuccess ) {
            (void)hipFree(*p);
            *p = nullptr;
        }
    }
    return e;
}

constexpr size_t kMaxLdsBytes = 160 * 1024; // gfx950 LDS per CU
constexpr int    kMaxDevices  = 64;

// Per-device FP32 twiddle table (a static device array), generated once in FP64 at the
// first transform on the device and kept until exit.
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
        e          = twiddle_table_address(&tw);
        if ( e != hipSuccess )
            return e;
        e = launch_twiddle_init(tw, stream);
        if ( e == hipSuccess )
            e = hipStreamSynchronize(stream);
        if ( e != hipSuccess )
            return e;
        g_tw[dev] = tw;
    }
    *out = g_tw[dev];
    return hipSuccess;
}

} // namespace

struct fastfft_plan_s {
    PlanState   st;
    hipStream_t stream    = hipStreamPerThread;
    int         precision = FASTFFT_PRECISION_FP32;
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
    // batch of correlations set by fastfft_set_batch: one entry per template, and
    // the xcorr stage list of every distinct template size
    struct BatchItem {
        const void* in;
        void*       out;
        int         x, y;
    };
    std::vector<BatchItem>                       batch;
    int*                                         batch_peaks  = nullptr; // device int[2n] or null
    float*                                       batch_values = nullptr; // device float[n] or null
    // template-matching statistics (fastfft_set_batch_stats): when on, the C2R part of
    // a batched call updates these maps in place of the outputs and peaks
    BatchStats batch_stats;
    int        stats_first_id = 0;
    bool       stats_on       = false;
    std::map<std::pair<int, int>, std::vector<Stage>> batch_stages;
    // batched kernels (when they cover the plan's stages): chunks of batch_chunk
    // templates share one launch per stage; batch_dev holds the A and B slots of
    // a chunk, the argmax keys and the per-template launch descriptors
    int           batch_chunk = 0; // 0: per-template launches
    int           batch_group = 0; // templates per template group of the correlation pass (0: the launcher's default)
    int           batch_p     = 0; // sub-transform length of the correlation pass (0: the planner's)
    std::string   batch_tuned;     // where the launch shape came from: env, cache or timed
    void*         batch_dev   = nullptr;
    RowBatchItem* batch_r2c   = nullptr;
    XcorrItem*    batch_xc    = nullptr;
    RowBatchItem* batch_c2r   = nullptr;
    // Sub-frame form of the batch (set_batch decides).  With templates of at most
    // Lx x Ly, a window of Mx x My lags at origin (ox, oy) and Fx >= Lx + Mx - 1,
    // Fy >= Ly + My - 1, every kept lag of the N-frame correlation,
    //   out[x][y] = sum_{u<Lx, v<Ly} t[u][v] g[(ox + x + u) mod N][(oy + y + v) mod N]
    // (g = unnormalised 2-D inverse of the image spectrum), reads g only inside
    // the Fx x Fy frame starting at (ox, oy).  Per call the image is reduced once
    // to that frame (sub_inv: trimmed inverse, scaled by 1/(Fx Fy)) and its
    // spectrum (sub_fwd); the templates are then correlated on the small frame
    // (sub_xc, window origin 0) with identical results up to rounding.
    bool               sub_active = false;
    int                sub_fx = 0, sub_fy = 0;
    std::vector<Stage> sub_inv, sub_fwd, sub_xc;
    void*              sub_dev = nullptr; // frame (real, FFTW padded), its spectrum, workspace a, b
    float*             sub_s   = nullptr;
    float2*            sub_h = nullptr, *sub_wa = nullptr, *sub_wb = nullptr;
    // The library's own HIP graph of one batched call: captured on graph_stream at the
    // first call with a given image and scales, then launched as one graph per call on
    // the plan's stream (when that stream is not itself being captured).
    hipStream_t    graph_stream = nullptr;
    hipGraph_t     graph        = nullptr;
    hipGraphExec_t graph_exec   = nullptr;
    std::string    graph_key;
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

// the template-group scratch of the batch statistics, and the statistics setting
static void clear_stats(fastfft_plan p) {
    if ( p->batch_stats.part )
        (void)hipFree(p->batch_stats.part);
    p->batch_stats = BatchStats{ };
    p->stats_on    = false;
}

static void free_batch_graph(fastfft_plan p) {
    if ( p->graph_exec )
        (void)hipGraphExecDestroy(p->graph_exec);
    if ( p->graph )
        (void)hipGraphDestroy(p->graph);
    p->graph_exec = nullptr;
    p->graph      = nullptr;
    p->graph_key.clear( );
}

static void free_batch_device(fastfft_plan p) {
    free_batch_graph(p);
    clear_stats(p); // the statistics belong to the batch's kernels
    if ( p->batch_dev )
        (void)hipFree(p->batch_dev);
    p->batch_dev   = nullptr;
    p->batch_r2c   = nullptr;
    p->batch_xc    = nullptr;
    p->batch_c2r   = nullptr;
    p->batch_chunk = 0;
    p->batch_group = 0;
    p->batch_p     = 0;
    p->batch_tuned.clear( );
}

static void free_subframe(fastfft_plan p) {
    if ( p->sub_dev )
        (void)hipFree(p->sub_dev);
    p->sub_dev    = nullptr;
    p->sub_s      = nullptr;
    p->sub_h      = nullptr;
    p->sub_wa     = nullptr;
    p->sub_wb     = nullptr;
    p->sub_active = false;
    p->sub_inv.clear( );
    p->sub_fwd.clear( );
    p->sub_xc.clear( );
}

static void invalidate(fastfft_plan p) {
    p->st.validated = false;
    p->batch.clear( );
    p->batch_stages.clear( );
    free_batch_device(p);
    free_subframe(p);
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
        e = dev_alloc(&p->ws, bytes);
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

static bool env_flag(const char* name) {
    const char* v = std::getenv(name);
    return v && v[0] && std::atoi(v) != 0;
}

static int env_chunk( ) {
    const char* v = std::getenv("FASTFFT_BATCH_CHUNK");
    return v ? std::atoi(v) : 0;
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
        if ( log_enabled( ) )
            log_text(std::string("plan ") + kOpNames[opi] + " stages:\n" + describe(p->st, st));
        p->stages[opi]       = std::move(st);
        p->stages_built[opi] = true;
    }
    *out = &p->stages[opi];
    return ok( );
}

static fastfft_status run_stage_list(fastfft_plan p, const std::vector<Stage>& stages, const char* opname, void* in,
                                     void* out, const void* image, float pre, float intra, float post, int first = 0,
                                     int last = -1, const ArgmaxSlots* amax = nullptr, bool* amax_fused = nullptr,
                                     float2* ws_a = nullptr, float2* ws_b = nullptr, const float2* tw_in = nullptr);

static fastfft_status run_stages(fastfft_plan p, int opi, void* in, void* out, const void* image, float pre, float intra,
                                 float post, int first = 0, int last = -1, const ArgmaxSlots* amax = nullptr,
                                 bool* amax_fused = nullptr) {
    const std::vector<Stage>* sp = nullptr;
    if ( auto rc = get_stages(p, opi, &sp) )
        return rc;
    return run_stage_list(p, *sp, kOpNames[opi], in, out, image, pre, intra, post, first, last, amax, amax_fused);
}

// ws_a / ws_b (both or neither): workspace of an internal stage list instead of the plan's.
// tw_in with ws_a: the twiddle table too, so that nothing of the plan's device state is
// allocated or used (the allocation-free entry point).
static fastfft_status run_stage_list(fastfft_plan p, const std::vector<Stage>& stages, const char* opname, void* in,
                                     void* out, const void* image, float pre, float intra, float post, int first, int last,
                                     const ArgmaxSlots* amax, bool* amax_fused, float2* ws_a, float2* ws_b,
                                     const float2* tw_in) {
    if ( last < 0 )
        last = (int)stages.size( ) - 1;
    if ( first < 0 || last >= (int)stages.size( ) || first > last )
        return fail(FASTFFT_INVALID_ARGUMENT,
                    "stage range outside the " + std::to_string(stages.size( )) + " stages of " + opname);
    float2* tw = const_cast<float2*>(tw_in);
    if ( ! (tw_in && ws_a) ) {
        fastfft_status rc = ensure_device_state(p, &tw);
        if ( rc )
            return rc;
    }
    float2* wa = ws_a ? ws_a : static_cast<float2*>(p->ws);
    float2* wb = ws_a ? ws_b : wa + p->ws_a;
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
        is_real = el == Elem::real_input_type || el == Elem::real_f32;
        rt      = (half_in && el != Elem::real_f32) ? RealT::f16 : RealT::f32;
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
                io.scale      = intra;
                io.split_bf16 = p->precision == FASTFFT_PRECISION_SPLIT_BF16;
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
        hipError_t e = dev_alloc(&plan->argmax_scratch, bytes);
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

// ---------------------------------------------------------------- batched correlation

// Elements a line map addresses for nlines lines of nelem elements (single-index lines).
static size_t map_extent(const HostLineMap& m, int nlines, int nelem) {
    return (size_t)((long)(nlines - 1) * m.s1 + (long)(nelem - 1) * m.es + 1);
}
// A slot: the R2C rows' output (nlines template rows of nout values); B slot: the
// correlation pass's output as the C2R pass reads it (nlines rows of the spectrum's
// w values), rounded to even counts so that every slot starts 16-byte aligned.
// Column-major slots (s1 = 1, es = the column length) hold whole columns: the row kernels
// move groups of rows past the last row, and the split-BF16 pass reads 16-row groups.
static size_t slot_extent(const HostLineMap& m, int nlines, int nelem) {
    if ( m.s1 == 1 && m.es > 1 )
        return (size_t)nelem * (size_t)m.es;
    return map_extent(m, nlines, nelem);
}
static size_t slot_elems_a(const Stage& r2c) { return (slot_extent(r2c.om, r2c.nlines, r2c.nout) + 1) / 2 * 2; }
static size_t slot_elems_b(const Stage& c2r) { return (slot_extent(c2r.im, c2r.nlines, c2r.N / 2 + 1) + 1) / 2 * 2; }

// Device block for chunks of C templates: A slots, B slots, argmax keys, then the
// descriptor arrays of all n templates (filled by fill_batch_items).
static size_t batch_dev_bytes(const std::vector<Stage>& st, int C, int n) {
    return sizeof(float2) * (size_t)C * (slot_elems_a(st[0]) + slot_elems_b(st[2])) +
           sizeof(unsigned long long) * (size_t)C * kArgmaxPartials + (sizeof(RowBatchItem) * 2 + sizeof(XcorrItem)) * (size_t)n;
}

static fastfft_status fill_batch_items(fastfft_plan p, const std::vector<Stage>& st, int C) {
    const int n     = (int)p->batch.size( );
    char*     base  = static_cast<char*>(p->batch_dev);
    float2*   slotA = reinterpret_cast<float2*>(base);
    float2*   slotB = slotA + (size_t)C * slot_elems_a(st[0]);
    auto*     keys  = reinterpret_cast<unsigned long long*>(slotB + (size_t)C * slot_elems_b(st[2]));
    p->batch_r2c    = reinterpret_cast<RowBatchItem*>(keys + (size_t)C * kArgmaxPartials);
    p->batch_c2r    = p->batch_r2c + n;
    p->batch_xc     = reinterpret_cast<XcorrItem*>(p->batch_c2r + n);
    std::vector<RowBatchItem> r2c(n), c2r(n);
    std::vector<XcorrItem>    xc(n);
    for ( int i = 0; i < n; i++ ) {
        const auto& it = p->batch[i];
        const int   sl = i % C;
        float2*     a  = slotA + (size_t)sl * slot_elems_a(st[0]);
        float2*     b  = slotB + (size_t)sl * slot_elems_b(st[2]);
        r2c[i]         = RowBatchItem{it.in, a, nullptr, 2L * (it.x / 2 + 1), it.x, it.y};
        xc[i]          = XcorrItem{a, b, it.y, 0};
        c2r[i]         = RowBatchItem{b, it.out, p->batch_peaks ? keys + (size_t)sl * kArgmaxPartials : nullptr, 0, 0, 0};
    }
    hipError_t e = hipMemcpy(p->batch_r2c, r2c.data( ), sizeof(RowBatchItem) * n, hipMemcpyHostToDevice);
    if ( e == hipSuccess )
        e = hipMemcpy(p->batch_c2r, c2r.data( ), sizeof(RowBatchItem) * n, hipMemcpyHostToDevice);
    if ( e == hipSuccess )
        e = hipMemcpy(p->batch_xc, xc.data( ), sizeof(XcorrItem) * n, hipMemcpyHostToDevice);
    return e == hipSuccess ? ok( ) : hip_fail(e, "batch descriptors");
}

// Parts of a batched call, for fastfft_run_batch_stage (timing of one part).
enum : unsigned {
    kBatchFrame   = 1u,  // reduction of the image to the frame (frame form only)
    kBatchR2C     = 2u,  // R2C rows of every template
    kBatchXcorr   = 4u,  // correlation pass
    kBatchC2R     = 8u,  // trimmed C2R rows (+ argmax row keys)
    kBatchArgmax  = 16u, // argmax reduction of the row keys
    kBatchAll     = 31u
};

// The correlation stage with sub-transform length P (0 or P = s.P: s itself).
static Stage xcorr_stage_with_p(const Stage& s, int P) {
    Stage x = s;
    if ( P > 0 && P != s.P ) {
        x.P = P;
        x.Q = s.N / P;
    }
    return x;
}

// Launches every chunk of the batch: R2C rows, fused columns, C2R rows (+ argmax keys);
// parts outside `parts` are skipped.
static fastfft_status run_batch_chunks(fastfft_plan p, const std::vector<Stage>& st, const void* image, CpxT img_ct,
                                       float pre, float intra, float post, const float2* tw, unsigned parts = kBatchAll,
                                       bool mixed = false) {
    const int   n    = (int)p->batch.size( );
    const int   C    = p->batch_chunk;
    const bool  half = p->st.input_type == FASTFFT_INPUT_F16;
    const RealT rt   = half ? RealT::f16 : RealT::f32;
    auto*       keys = reinterpret_cast<unsigned long long*>(p->batch_r2c) - (size_t)C * kArgmaxPartials;
    for ( int c0 = 0; c0 < n; c0 += C ) {
        const int  cnt = std::min(C, n - c0);
        hipError_t e   = hipSuccess;
        if ( parts & kBatchR2C )
            e = launch_r2c_rows_batch(st[0], p->batch_r2c + c0, cnt, st[0].nlines, rt, pre, tw, p->stream);
        if ( e == hipSuccess && (parts & kBatchXcorr) ) {
            StageIO io;
            io.img     = image;
            io.img_ct  = img_ct;
            io.scale   = intra;
            io.xitems     = p->batch_xc + c0;
            io.nxitems    = cnt;
            io.xout_step  = (long)slot_elems_b(st[2]);
            io.xgroup     = p->batch_group;
            io.split_bf16 = p->precision == FASTFFT_PRECISION_SPLIT_BF16;
            io.img_mixed  = mixed;
            e             = launch_xcorr(xcorr_stage_with_p(st[1], p->batch_p), io, tw, p->stream);
        }
        if ( e == hipSuccess && (parts & kBatchC2R) )
            e = p->stats_on ? launch_c2r_stats_batch(st[2], p->batch_c2r + c0, cnt, post, tw, p->stats_first_id + c0,
                                                     p->batch_stats, p->stream)
                            : launch_c2r_rows_batch(st[2], p->batch_c2r + c0, cnt, rt, post, tw, p->stream);
        if ( e == hipSuccess && p->batch_peaks && ! p->stats_on && (parts & kBatchArgmax) )
            e = launch_argmax_keys(keys, st[2].nlines, st[2].M, p->batch_values + c0, p->batch_peaks + 2 * c0, p->stream, cnt,
                                   kArgmaxPartials);
        if ( e != hipSuccess )
            return hip_fail(e, "batched correlation launch");
    }
    return ok( );
}

static int next_pow2(int v);

// The split-BF16 form of a batch on a reduced frame: the frame's forward list stops after
// its R2C x-pass, whose output (the frame in mixed form, with the frame spectrum's column
// layout) the matrix-core correlation pass reads in place of the spectrum.
static bool batch_mixed(fastfft_plan p) {
    if ( ! p->sub_active || p->precision != FASTFFT_PRECISION_SPLIT_BF16 || p->sub_fwd.size( ) != 2 ||
         p->sub_xc.size( ) != 3 )
        return false;
    const Stage &f = p->sub_fwd[0], &x = p->sub_xc[1];
    return f.kind == StageKind::fwd_real && (f.dst == Buf::ws_a || f.dst == Buf::ws_b) && f.om.s1 == 1 &&
           f.om.n2 == 1 && f.om.es == x.gm.s1 && f.N == x.N && f.nlines == x.N && f.nout == x.nlines &&
           xcorr_toeplitz_covers(x);
}

// Tuning-cache key of a batch: device architecture, precision, element types, the plan's
// transform size, the frame the templates are correlated on, the largest template, the
// window, the batch size (rounded up to a power of two, at most 128) and whether peaks
// are computed.
static std::string batch_tuning_key(fastfft_plan p, const std::vector<Stage>& st, CpxT img_ct) {
    int lx = 0, ly = 0;
    for ( const auto& it : p->batch ) {
        lx = std::max(lx, it.x);
        ly = std::max(ly, it.y);
    }
    const int   n  = std::min(next_pow2((int)p->batch.size( )), 128);
    const auto& nn = p->st.inv_in;
    const auto& mo = p->st.inv_out;
    char        buf[320];
    std::snprintf(buf, sizeof(buf), "%s %s %s/%s plan=%dx%d frame=%s L=%dx%d M=%dx%d o=%d,%d n=%d peaks=%d",
                  device_arch( ).c_str( ), p->precision == FASTFFT_PRECISION_SPLIT_BF16 ? "split_bf16" : "fp32",
                  p->st.input_type == FASTFFT_INPUT_F16 ? "f16" : "f32", img_ct == CpxT::c32h ? "c32h" : "c64", nn.x, nn.y,
                  p->sub_active ? (std::to_string(p->sub_fx) + "x" + std::to_string(p->sub_fy)).c_str( ) : "full", lx, ly,
                  mo.x, mo.y, p->sub_active ? 0 : p->st.origin[0], p->sub_active ? 0 : p->st.origin[1], n,
                  p->batch_peaks != nullptr ? 1 : 0);
    (void)st;
    return buf;
}

// Chooses the batch's launch shape -- templates per launch (chunk), templates per group
// of the correlation pass, and the pass's sub-transform length -- then allocates
// batch_dev for it and fills the descriptors.  In order: FASTFFT_BATCH_CHUNK (with the
// default group and length), the tuning cache (unless FASTFFT_TUNE=1), else timing every
// candidate on this batch: the whole batched call, one warm-up and the best of three
// runs, against a constant image, with the outputs and peaks redirected to scratch of
// the plan's (the caller's buffers are not written); the fastest shape wins, a larger
// chunk within 3 % of it preferred (fewer, larger launches replay better from a graph).
// The timed choice is recorded in the tuning cache.
static fastfft_status setup_batch_kernels(fastfft_plan p, const std::vector<Stage>& st, CpxT img_ct) {
    const int n  = (int)p->batch.size( );
    float2*   tw = nullptr;
    if ( auto rc = ensure_device_state(p, &tw) )
        return rc;
    free_batch_device(p);
    const std::string key = batch_tuning_key(p, st, img_ct);
    TuneEntry         best;
    std::string       how;
    if ( env_chunk( ) > 0 ) {
        best.chunk = env_chunk( );
        how        = "env";
    }
    else if ( ! env_flag("FASTFFT_TUNE") && tuning_lookup(key, best) && best.chunk > 0 ) {
        how = "cache";
        // an entry from another build may name a length this batch's kernels do not take
        if ( best.p > 0 && ! xcorr_col_covers(xcorr_stage_with_p(st[1], best.p), img_ct) )
            best.p = 0;
    }
    else {
        // candidate shapes (the matrix-core pass has one template group per 32 templates)
        const bool       mixed = batch_mixed(p);
        std::vector<int> chunks, groups = {1, 2, 4, 8, 16}, ps = {0};
        if ( mixed )
            groups = {0};
        const size_t     per = sizeof(float2) * (slot_elems_a(st[0]) + slot_elems_b(st[2]));
        for ( int C = 1; C <= std::min(n, 64); C *= 2 )
            if ( (C >= 8 || C == std::min(n, 64) || n < 8) && (C == 1 || per * (size_t)C <= (size_t(1) << 30)) )
                chunks.push_back(C);
        if ( chunks.empty( ) )
            chunks.push_back(1);
        // the 1024-point correlation pass as one wave per column (P = N) besides the planner's split
        if ( p->precision == FASTFFT_PRECISION_FP32 && st[1].N == 1024 && st[1].P != 1024 &&
             xcorr_col_covers(xcorr_stage_with_p(st[1], 1024), img_ct) )
            ps.push_back(1024);
        const int cmax = chunks.back( );
        // scratch: the slots, descriptors and keys for the largest chunk, then one output
        // per slot and the peaks and values of every template
        const size_t out_bytes = elem_bytes(p->st) * (size_t)p->st.inv_output_memory;
        const size_t bd        = (batch_dev_bytes(st, cmax, n) + 255) / 256 * 256;
        const size_t img_bytes = (img_ct == CpxT::c32h ? 4 : 8) * (size_t)st[1].gm.s1 * st[1].nlines;
        const size_t outs_off  = bd;
        const size_t pk_off    = outs_off + (out_bytes * cmax + 255) / 256 * 256;
        const size_t img_off   = pk_off + ((sizeof(int) * 2 + sizeof(float)) * (size_t)n + 255) / 256 * 256;
        void*        scratch   = nullptr;
        hipError_t   e         = dev_alloc(&scratch, img_off + img_bytes);
        if ( e != hipSuccess )
            return hip_fail(e, "batch tuning scratch");
        char* base = static_cast<char*>(scratch);
        // a constant image (about 0.0115 in every float: no zeros, no denormals)
        e = hipMemset(base + img_off, 0x3c, img_bytes);
        if ( e == hipSuccess )
            e = hipDeviceSynchronize( );
        // redirect the outputs and peaks for the timing runs
        std::vector<void*> user_out(n);
        int*               user_pk  = p->batch_peaks;
        float*             user_val = p->batch_values;
        for ( int i = 0; i < n; i++ ) {
            user_out[i]      = p->batch[i].out;
            p->batch[i].out = base + outs_off + out_bytes * (size_t)(i % cmax);
        }
        if ( user_pk ) {
            p->batch_peaks  = reinterpret_cast<int*>(base + pk_off);
            p->batch_values = reinterpret_cast<float*>(base + pk_off + sizeof(int) * 2 * (size_t)n);
        }
        p->batch_dev      = scratch;
        hipEvent_t     e0 = nullptr, e1 = nullptr;
        fastfft_status run = FASTFFT_SUCCESS; // a failed launch keeps its own message
        if ( e == hipSuccess )
            e = hipEventCreate(&e0);
        if ( e == hipSuccess )
            e = hipEventCreate(&e1);
        best = TuneEntry{ };
        std::vector<TuneEntry> timed;
        for ( int pc : ps )
            for ( int g : groups )
                for ( int C : chunks ) {
                    if ( e != hipSuccess || run || g > C )
                        continue; // (g = 0: the kernel's own grouping)
                    p->batch_chunk = C;
                    p->batch_group = g;
                    p->batch_p     = pc;
                    if ( (run = fill_batch_items(p, st, C)) != FASTFFT_SUCCESS )
                        continue;
                    float ms = 0.f;
                    for ( int rep = 0; rep < 4 && e == hipSuccess && ! run; rep++ ) {
                        e = hipEventRecord(e0, p->stream);
                        if ( e == hipSuccess )
                            run = run_batch_chunks(p, st, base + img_off, img_ct, 1.f, 1.f, 1.f, tw, kBatchAll, mixed);
                        if ( e == hipSuccess && ! run )
                            e = hipEventRecord(e1, p->stream);
                        if ( e == hipSuccess && ! run )
                            e = hipEventSynchronize(e1);
                        float t = 0.f;
                        if ( e == hipSuccess && ! run )
                            e = hipEventElapsedTime(&t, e0, e1);
                        if ( rep == 1 || (rep > 1 && t < ms) )
                            ms = t;
                    }
                    if ( e != hipSuccess || run )
                        continue;
                    timed.push_back(TuneEntry{C, g, pc, (double)ms});
                }
        // the fastest, or a larger chunk within 3 % of it (the fastest of those)
        double tmin = 0.0;
        for ( const TuneEntry& c : timed )
            if ( tmin == 0.0 || c.ms < tmin )
                tmin = c.ms;
        for ( const TuneEntry& c : timed )
            if ( c.ms <= 1.03 * tmin &&
                 (best.chunk == 0 || c.chunk > best.chunk || (c.chunk == best.chunk && c.ms < best.ms)) )
                best = c;
        if ( e0 )
            (void)hipEventDestroy(e0);
        if ( e1 )
            (void)hipEventDestroy(e1);
        for ( int i = 0; i < n; i++ )
            p->batch[i].out = user_out[i];
        p->batch_peaks  = user_pk;
        p->batch_values = user_val;
        free_batch_device(p); // frees the scratch
        if ( run )
            return run;
        if ( e != hipSuccess )
            return hip_fail(e, "batch tuning");
        if ( best.chunk <= 0 )
            return fail(FASTFFT_HIP_ERROR, "batch tuning: no candidate shape ran");
        tuning_store(key, best);
        how = "timed";
    }
    hipError_t e = dev_alloc(&p->batch_dev, batch_dev_bytes(st, best.chunk, n));
    if ( e != hipSuccess ) {
        p->batch_dev = nullptr;
        return hip_fail(e, "hipMalloc(batch)");
    }
    p->batch_chunk = best.chunk;
    p->batch_group = best.group;
    p->batch_p     = best.p;
    p->batch_tuned = how;
    return fill_batch_items(p, st, best.chunk);
}

static int next_pow2(int v) {
    int f = 1;
    while ( f < v )
        f <<= 1;
    return f;
}

// Internal stage lists keep every operand in FP32: the planner's input-type
// operands become FP32 ones.
static void stages_to_f32(std::vector<Stage>& st) {
    for ( Stage& s : st )
        for ( Elem* e : {&s.src_elem, &s.dst_elem} ) {
            if ( *e == Elem::real_input_type )
                *e = Elem::real_f32;
            else if ( *e == Elem::cpx_input_type )
                *e = Elem::cpx_f32;
        }
}

// Fills the derived fields of an internal plan's sizes (w = x/2+1 and the element
// counts) as the plan calls do for a caller's plan.
static void fill_sizes(PlanState& ps) {
    ps.input_memory      = padded_memory_size(ps.fwd_in);
    ps.fwd_output_memory = padded_memory_size(ps.fwd_out);
    (void)padded_memory_size(ps.inv_in);
    ps.inv_output_memory = padded_memory_size(ps.inv_out);
    ps.compute_memory    = std::max(std::max(ps.input_memory, ps.fwd_output_memory),
                                    std::max((long)padded_memory_size(ps.inv_in), ps.inv_output_memory));
}

static PlanState frame_plan(int input_type, const Dims& n, const Dims& out, const int origin[3]) {
    PlanState ps;
    ps.input_type = input_type;
    ps.image_type = FASTFFT_IMAGE_C64;
    ps.rank       = 2;
    ps.fwd_set = ps.inv_set = true;
    ps.fwd_in = ps.fwd_out = ps.inv_in = Dims{n.x, n.y, 1, 0};
    ps.inv_out                          = Dims{out.x, out.y, 1, 0};
    for ( int a = 0; a < 3; a++ )
        ps.origin[a] = origin[a];
    fill_sizes(ps);
    return ps;
}

// Column-major slots of a frame batch's stage list (R2C rows, correlation, C2R rows):
// template spectra A[kx][v] (columns of la >= template rows, a multiple of 16) and
// correlation outputs B[kx][j] (columns of mb >= window rows, a multiple of 8), so
// that the correlation passes read and write whole columns; the row kernels move
// their rows through LDS tiles (rows_packed.hpp).
static void column_major_slots(std::vector<Stage>& xc) {
    Stage &    r = xc[0], &x = xc[1], &c = xc[2];
    const long la = (r.nlines + 15) / 16 * 16;
    const long mb = (c.nlines + 7) / 8 * 8;
    r.om          = HostLineMap{1, 0, la, 1};
    x.im          = HostLineMap{la, 0, 1, 1};
    x.om          = HostLineMap{mb, 0, 1, 1};
    c.im          = HostLineMap{1, 0, mb, 1};
}

// Activates the sub-frame form of the batch (see fastfft_plan_s) when the frame
// is at most a quarter of the plan's, FASTFFT_NO_SUBFRAME is unset and the
// batched kernels cover the frame's correlation stages; otherwise leaves it off.
static fastfft_status setup_subframe(fastfft_plan p) {
    free_subframe(p);
    const char* off = std::getenv("FASTFFT_NO_SUBFRAME");
    if ( off && off[0] && std::atoi(off) != 0 )
        return ok( );
    const PlanState& st = p->st;
    const Dims &     in = st.fwd_in, &nn = st.inv_in, &win = st.inv_out;
    if ( st.transform_dimension != 2 )
        return ok( );
    const int fx = next_pow2(in.x + win.x - 1), fy = next_pow2(in.y + win.y - 1);
    if ( fx >= nn.x || fy >= nn.y || 4L * fx * fy > (long)nn.x * nn.y )
        return ok( );
    const int zero[3] = {0, 0, 0};
    // the templates' correlation on the frame: window origin 0 of the frame is lag origin of the plan
    PlanState ps = st;
    ps.fwd_out = ps.inv_in = Dims{fx, fy, 1, 0};
    ps.image_type          = FASTFFT_IMAGE_C64;
    for ( int a = 0; a < 3; a++ )
        ps.origin[a] = 0;
    fill_sizes(ps);
    ps.validated = false;
    std::vector<Stage> xc, inv, fwd;
    if ( validate(ps).code || build_stages(ps, "xcorr", xc).code || xc.size( ) != 3 )
        return ok( );
    // column-major template and output slots, unless FASTFFT_BATCH_ROW_MAJOR=1 (or the
    // column-major row kernels do not cover the frame)
    if ( ! env_flag("FASTFFT_BATCH_ROW_MAJOR") ) {
        std::vector<Stage> cm = xc;
        column_major_slots(cm);
        if ( batch_kernels_cover(cm[0], cm[1], cm[2], CpxT::c64) )
            xc = std::move(cm);
    }
    if ( ! batch_kernels_cover(xc[0], xc[1], xc[2], CpxT::c64) )
        return ok( );
    // the image reduced to the frame: trimmed inverse of the N-frame spectrum (read in the plan's image type) ...
    PlanState pi = frame_plan(FASTFFT_INPUT_F32, nn, Dims{fx, fy, 1, 0}, st.origin);
    // ... and the frame's spectrum
    PlanState pf = frame_plan(FASTFFT_INPUT_F32, Dims{fx, fy, 1, 0}, Dims{fx, fy, 1, 0}, zero);
    if ( validate(pi).code || build_stages(pi, "inv", inv).code || validate(pf).code || build_stages(pf, "fwd", fwd).code )
        return ok( );
    stages_to_f32(inv);
    stages_to_f32(fwd);
    inv.front( ).src_elem = Elem::cpx_image_type;
    for ( const auto* list : {&inv, &fwd, &xc} )
        for ( const Stage& s : *list )
            if ( stage_lds_bytes(s.P) > kMaxLdsBytes )
                return ok( );
    // device block: frame, frame spectrum, and the larger workspace of the two internal lists
    Dims   fd{fx, fy, 1, 0};
    size_t s_floats = (size_t)padded_memory_size(fd);
    size_t h_elems  = (size_t)(fx / 2 + 1) * (size_t)fy;
    size_t ia, ib, fa, fb;
    workspace_elems(pi, ia, ib);
    workspace_elems(pf, fa, fb);
    const size_t wa = std::max(ia, fa), wb = std::max(ib, fb);
    s_floats        = (s_floats + 63) / 64 * 64;
    const size_t bytes = sizeof(float) * s_floats + sizeof(float2) * (h_elems + wa + wb) + 256;
    hipError_t   e     = dev_alloc(&p->sub_dev, bytes);
    if ( e != hipSuccess ) {
        p->sub_dev = nullptr;
        return hip_fail(e, "hipMalloc(batch frame)");
    }
    p->sub_s      = static_cast<float*>(p->sub_dev);
    p->sub_h      = reinterpret_cast<float2*>(p->sub_s + s_floats);
    p->sub_wa     = p->sub_h + (h_elems + 1) / 2 * 2;
    p->sub_wb     = p->sub_wa + (wa + 1) / 2 * 2;
    p->sub_fx     = fx;
    p->sub_fy     = fy;
    p->sub_inv    = std::move(inv);
    p->sub_fwd    = std::move(fwd);
    p->sub_xc     = std::move(xc);
    p->sub_active = true;
    return ok( );
}

// The batch on the sub-frame: image -> frame -> frame spectrum, then every chunk.
static fastfft_status run_subframe_batch(fastfft_plan p, const void* image, float pre, float intra, float post,
                                         const float2* tw, unsigned parts = kBatchAll) {
    const float frame_scale = 1.f / ((float)p->sub_fx * (float)p->sub_fy);
    const bool  mixed       = batch_mixed(p);
    if ( parts & kBatchFrame ) {
        const auto &inv = p->sub_inv, &fwd = p->sub_fwd;
        // the inverse's C2R rows and the forward's R2C rows are the frame's rows: one kernel
        // keeps them in LDS (the real frame is not stored)
        const bool fuse = inv.size( ) >= 2 && fwd.size( ) >= 2 && inv.back( ).src_elem == Elem::cpx_f32 &&
                          inv.back( ).dst == Buf::user_out && fwd.front( ).src == Buf::user_in &&
                          fwd.front( ).dst_elem == Elem::cpx_f32 &&
                          (inv.back( ).src == Buf::ws_a || inv.back( ).src == Buf::ws_b) &&
                          (fwd.front( ).dst == Buf::ws_a || fwd.front( ).dst == Buf::ws_b) &&
                          inv.back( ).src != fwd.front( ).dst && c2r_r2c_rows_covers(inv.back( ), fwd.front( ));
        const int ilast = (int)inv.size( ) - 1 - (fuse ? 1 : 0);
        if ( auto rc = run_stage_list(p, inv, "batch frame", const_cast<void*>(image), p->sub_s, nullptr, 1.f, 1.f,
                                      frame_scale, 0, ilast, nullptr, nullptr, p->sub_wa, p->sub_wb) )
            return rc;
        if ( fuse ) {
            // the C2R keeps the list's post scale (its last stage), the R2C the forward's pre scale 1
            float2*          tw  = nullptr;
            if ( auto rc = ensure_device_state(p, &tw) )
                return rc;
            const auto*      src = inv.back( ).src == Buf::ws_a ? p->sub_wa : p->sub_wb;
            float2*          dst = fwd.front( ).dst == Buf::ws_a ? p->sub_wa : p->sub_wb;
            const hipError_t e   = launch_c2r_r2c_rows(inv.back( ), fwd.front( ), src, dst, frame_scale, 1.f, tw, p->stream);
            if ( e != hipSuccess )
                return hip_fail(e, "batch frame rows");
        }
        // the frame's spectrum; in the split-BF16 form only its R2C x-pass (already done by
        // the fused kernel)
        if ( ! (mixed && fuse) )
            if ( auto rc = run_stage_list(p, fwd, "batch frame spectrum", p->sub_s, p->sub_h, nullptr, 1.f, 1.f, 1.f,
                                          fuse ? 1 : 0, mixed ? 0 : -1, nullptr, nullptr, p->sub_wa, p->sub_wb) )
                return rc;
    }
    const void* img = mixed ? (p->sub_fwd[0].dst == Buf::ws_a ? p->sub_wa : p->sub_wb) : static_cast<const void*>(p->sub_h);
    return run_batch_chunks(p, p->sub_xc, img, CpxT::c64, pre, intra, post, tw, parts, mixed);
}

// Runs one batched call through the plan's own HIP graph: captured at the first call
// with this image and these scales (on a stream of the plan's, in relaxed capture mode:
// the call issues kernel launches only), then one graph launch on the plan's stream per
// call.  Calls on a stream that is itself being captured, calls in statistics form (whose
// template ids change from call to call) and FASTFFT_NO_BATCH_GRAPH=1 launch the kernels
// directly.
template <class Run>
static fastfft_status run_batch_graph(fastfft_plan p, const void* image, float pre, float intra, float post, Run&& run) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if ( p->stats_on || env_flag("FASTFFT_NO_BATCH_GRAPH") || hipStreamIsCapturing(p->stream, &cs) != hipSuccess ||
         cs != hipStreamCaptureStatusNone )
        return run( );
    char key[160];
    std::snprintf(key, sizeof(key), "%p/%a/%a/%a", image, (double)pre, (double)intra, (double)post);
    if ( ! p->graph_exec || p->graph_key != key ) {
        free_batch_graph(p);
        hipError_t e = hipSuccess;
        if ( ! p->graph_stream )
            e = hipStreamCreateWithFlags(&p->graph_stream, hipStreamNonBlocking);
        if ( e != hipSuccess ) {
            p->graph_stream = nullptr;
            return hip_fail(e, "batch graph stream");
        }
        const hipStream_t user = p->stream;
        p->stream              = p->graph_stream;
        e                      = hipStreamBeginCapture(p->graph_stream, hipStreamCaptureModeRelaxed);
        fastfft_status rc      = e == hipSuccess ? run( ) : FASTFFT_SUCCESS;
        hipGraph_t     g       = nullptr;
        if ( e == hipSuccess ) {
            const hipError_t ee = hipStreamEndCapture(p->graph_stream, &g);
            if ( ! rc && ee != hipSuccess )
                e = ee;
        }
        p->stream = user;
        if ( e == hipSuccess && ! rc )
            e = hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0);
        if ( rc || e != hipSuccess ) {
            if ( g )
                (void)hipGraphDestroy(g);
            p->graph_exec = nullptr;
            return rc ? rc : hip_fail(e, "batch graph capture");
        }
        p->graph     = g;
        p->graph_key = key;
    }
    const hipError_t e = hipGraphLaunch(p->graph_exec, p->stream);
    return e == hipSuccess ? ok( ) : hip_fail(e, "batch graph launch");
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
    free_batch_device(plan);
    free_subframe(plan);
    if ( plan->graph_stream )
        (void)hipStreamDestroy(plan->graph_stream);
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

fastfft_status fastfft_set_precision(fastfft_plan plan, int precision) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( precision != FASTFFT_PRECISION_FP32 && precision != FASTFFT_PRECISION_SPLIT_BF16 )
        return fail(FASTFFT_INVALID_ARGUMENT, "precision must be FASTFFT_PRECISION_FP32 or FASTFFT_PRECISION_SPLIT_BF16");
    // the batch's captured graph holds the other precision's kernels (and, on a frame, its
    // frame passes): it is captured again at the next call
    if ( plan->precision != precision )
        free_batch_graph(plan);
    plan->precision = precision;
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

fastfft_status fastfft_fwd_image_inv_argmax_noalloc(fastfft_plan plan, void* in, const void* image_to_search, void* out,
                                                    fastfft_op pre_op, fastfft_op intra_op, fastfft_op post_op,
                                                    int* peak_xy, float* value) {
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
    const std::vector<Stage>* sp = nullptr;
    if ( auto rc = get_stages(plan, 2, &sp) )
        return rc;
    // scratch layout: the stage list's two intermediates, then the argmax row keys, the
    // partial values and indices of the two-kernel epilogue (each part 256-byte aligned)
    size_t a = 0, b = 0;
    stage_list_workspace(*sp, a, b);
    auto         up    = [](size_t v) { return (v + 255) / 256 * 256; };
    const size_t off_b = up(sizeof(float2) * a);
    const size_t off_k = off_b + up(sizeof(float2) * b);
    const size_t off_v = off_k + up(sizeof(unsigned long long) * kArgmaxPartials);
    const size_t off_i = off_v + up(sizeof(float) * kArgmaxPartials);
    const size_t need  = off_i + up(sizeof(int) * kArgmaxPartials);
    if ( need > kStaticScratchBytes )
        return fail(FASTFFT_UNSUPPORTED, "fwd_image_inv_argmax_noalloc: this plan needs " + std::to_string(need) +
                                             " bytes of scratch, more than the " + std::to_string(kStaticScratchBytes) +
                                             " bytes of the library's static scratch (use fastfft_fwd_image_inv_argmax)");
    // the twiddle table is a static device array as well: the first call on a device fills it
    float2*    tw = nullptr;
    hipError_t e  = twiddles_for_current_device(&tw, plan->stream);
    if ( e != hipSuccess )
        return hip_fail(e, "twiddle table");
    void* base = nullptr;
    e          = static_scratch_for_current_device(&base);
    if ( e != hipSuccess )
        return hip_fail(e, "hipGetSymbolAddress(scratch)");
    char*       sc = static_cast<char*>(base);
    ArgmaxSlots slots;
    slots.keys  = reinterpret_cast<unsigned long long*>(sc + off_k);
    slots.value = value;
    slots.xy    = peak_xy;
    void* dst   = out ? out : in;
    bool  fused = false;
    if ( auto rc = run_stage_list(plan, *sp, "xcorr", in, dst, image_to_search, op_scale(pre_op), op_scale(intra_op),
                                  op_scale(post_op), 0, -1, &slots, &fused, reinterpret_cast<float2*>(sc),
                                  reinterpret_cast<float2*>(sc + off_b), tw) )
        return rc;
    plan->last_buffer     = dst;
    plan->stage_completed = 7;
    if ( ! fused ) {
        const Dims& o = plan->st.inv_out;
        e             = launch_argmax(dst, plan->st.input_type == FASTFFT_INPUT_F16 ? RealT::f16 : RealT::f32, o.x, o.y,
                                      2 * o.w, reinterpret_cast<float*>(sc + off_v), reinterpret_cast<int*>(sc + off_i),
                                      value, peak_xy, plan->stream);
        if ( e != hipSuccess )
            return hip_fail(e, "argmax");
    }
    return ok( );
}

static std::string batch_description(fastfft_plan plan);

fastfft_status fastfft_set_batch(fastfft_plan plan, int n, const void* const* templates, const int* sizes_xy,
                                 void* const* outputs, int* peaks_xy, float* values) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( n < 0 || (n > 0 && (! templates || ! sizes_xy || ! outputs)) || ((peaks_xy == nullptr) != (values == nullptr)) )
        return fail(FASTFFT_INVALID_ARGUMENT, "set_batch: bad count, null array, or only one of peaks_xy / values");
    if ( auto rc = ensure_validated(plan) )
        return rc;
    if ( plan->st.transform_dimension != 2 )
        return fail(FASTFFT_UNSUPPORTED, "batched FwdImageInvFFT is implemented for 2D plans");
    const Dims& in = plan->st.fwd_in;
    std::vector<fastfft_plan_s::BatchItem>            items;
    std::map<std::pair<int, int>, std::vector<Stage>> lists;
    for ( int i = 0; i < n; i++ ) {
        const int x = sizes_xy[2 * i], y = sizes_xy[2 * i + 1];
        if ( ! templates[i] || ! outputs[i] )
            return fail(FASTFFT_INVALID_ARGUMENT, "set_batch: null template or output pointer at " + std::to_string(i));
        if ( x < 1 || y < 1 || x > in.x || y > in.y )
            return fail(FASTFFT_INVALID_ARGUMENT, "set_batch: template " + std::to_string(i) + " is " + std::to_string(x) + "x" +
                                                          std::to_string(y) + ", outside 1.." + std::to_string(in.x) + "x" +
                                                          std::to_string(in.y) + " (the plan's forward input size)");
        items.push_back({templates[i], outputs[i], x, y});
        const auto key = std::make_pair(x, y);
        if ( lists.count(key) )
            continue;
        // the plan with this template size as its forward input: same padded size,
        // window and types, so the workspace of the plan covers it
        PlanState ps       = plan->st;
        ps.fwd_in          = Dims{x, y, 1, 0};
        ps.input_memory    = padded_memory_size(ps.fwd_in); // also sets fwd_in.w (the row pitch)
        PlanError pe       = validate(ps);
        std::vector<Stage> st;
        if ( ! pe.code )
            pe = build_stages(ps, "xcorr", st);
        if ( pe.code )
            return fail(pe.code, "set_batch: template " + std::to_string(i) + ": " + pe.msg);
        lists[key] = std::move(st);
    }
    free_batch_device(plan);
    free_subframe(plan);
    plan->batch        = std::move(items);
    plan->batch_stages = std::move(lists);
    plan->batch_peaks  = peaks_xy;
    plan->batch_values = values;
    clear_stats(plan);
    // one launch per stage for chunks of templates when the fast kernels cover the plan's stages
    const std::vector<Stage>* sp = nullptr;
    if ( auto rc = get_stages(plan, 2, &sp) )
        return rc;
    const CpxT img_ct = plan->st.image_type == FASTFFT_IMAGE_C32H ? CpxT::c32h : CpxT::c64;
    fastfft_status rc = ok( );
    if ( n > 0 )
        rc = setup_subframe(plan);
    if ( ! rc && plan->sub_active )
        rc = setup_batch_kernels(plan, plan->sub_xc, CpxT::c64);
    else if ( ! rc && n > 0 && sp->size( ) == 3 && batch_kernels_cover((*sp)[0], (*sp)[1], (*sp)[2], img_ct) )
        rc = setup_batch_kernels(plan, *sp, img_ct);
    if ( rc ) {
        plan->batch.clear( );
        plan->batch_stages.clear( );
        free_subframe(plan);
    }
    else if ( log_enabled( ) ) {
        log_text("plan " + batch_description(plan));
    }
    return rc;
}

fastfft_status fastfft_fwd_image_inv_batch(fastfft_plan plan, const void* image_to_search, fastfft_op pre_op,
                                           fastfft_op intra_op, fastfft_op post_op) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( ! image_to_search )
        return fail(FASTFFT_INVALID_ARGUMENT, "null image_to_search pointer");
    if ( auto rc = op_kind_check(pre_op, true, false, "the FwdImageInvFFT pre op") )
        return rc;
    if ( intra_op.kind != FASTFFT_OP_CONJ_MUL && intra_op.kind != FASTFFT_OP_CONJ_MUL_THEN_SCALE )
        return fail(FASTFFT_INVALID_ARGUMENT, "FwdImageInvFFT needs an intra op (CONJ_MUL or CONJ_MUL_THEN_SCALE)");
    if ( auto rc = op_kind_check(post_op, true, false, "the FwdImageInvFFT post op") )
        return rc;
    if ( auto rc = ensure_validated(plan) )
        return rc;
    if ( plan->stats_on && (plan->batch_chunk <= 0 || reinterpret_cast<std::uintptr_t>(image_to_search) % 16 != 0) )
        return fail(FASTFFT_UNSUPPORTED, "batch statistics need the batched kernels and a 16-byte aligned image");
    if ( plan->batch_chunk > 0 && reinterpret_cast<std::uintptr_t>(image_to_search) % 16 == 0 ) {
        const std::vector<Stage>* sp = nullptr;
        if ( auto rc = get_stages(plan, 2, &sp) )
            return rc;
        float2* tw = nullptr;
        if ( auto rc = ensure_device_state(plan, &tw) )
            return rc;
        const CpxT  img_ct = plan->st.image_type == FASTFFT_IMAGE_C32H ? CpxT::c32h : CpxT::c64;
        const float pre = op_scale(pre_op), intra = op_scale(intra_op), post = op_scale(post_op);
        auto        run = [&]( ) {
            return plan->sub_active ? run_subframe_batch(plan, image_to_search, pre, intra, post, tw)
                                           : run_batch_chunks(plan, *sp, image_to_search, img_ct, pre, intra, post, tw);
        };
        if ( auto rc = run_batch_graph(plan, image_to_search, pre, intra, post, run) )
            return rc;
        if ( ! plan->batch.empty( ) ) {
            plan->last_buffer     = plan->batch.back( ).out;
            plan->stage_completed = 7;
        }
        return ok( );
    }
    ArgmaxScratch a;
    if ( plan->batch_peaks )
        if ( auto rc = argmax_scratch(plan, a) )
            return rc;
    const Dims& o = plan->st.inv_out;
    for ( size_t i = 0; i < plan->batch.size( ); i++ ) {
        const auto& it = plan->batch[i];
        const auto& st = plan->batch_stages.at(std::make_pair(it.x, it.y));
        ArgmaxSlots slots = a.fused;
        bool        fused = false;
        if ( plan->batch_peaks ) {
            slots.value = plan->batch_values + i;
            slots.xy    = plan->batch_peaks + 2 * i;
        }
        if ( auto rc = run_stage_list(plan, st, "xcorr", const_cast<void*>(it.in), it.out, image_to_search,
                                      op_scale(pre_op), op_scale(intra_op), op_scale(post_op), 0, -1,
                                      plan->batch_peaks ? &slots : nullptr, &fused) )
            return rc;
        if ( plan->batch_peaks && ! fused ) {
            hipError_t e = launch_argmax(it.out, plan->st.input_type == FASTFFT_INPUT_F16 ? RealT::f16 : RealT::f32, o.x,
                                         o.y, 2 * o.w, a.pv, a.pi, slots.value, slots.xy, plan->stream);
            if ( e != hipSuccess )
                return hip_fail(e, "argmax");
        }
        plan->last_buffer = it.out;
    }
    if ( ! plan->batch.empty( ) )
        plan->stage_completed = 7;
    return ok( );
}

fastfft_status fastfft_run_batch_stage(fastfft_plan plan, int stage, const void* image_to_search, fastfft_op pre_op,
                                       fastfft_op intra_op, fastfft_op post_op) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( stage < 0 || stage > 4 || ! image_to_search )
        return fail(FASTFFT_INVALID_ARGUMENT, "run_batch_stage: stage must be 0..4 and the image not null");
    if ( plan->batch_chunk <= 0 || reinterpret_cast<std::uintptr_t>(image_to_search) % 16 != 0 )
        return fail(FASTFFT_UNSUPPORTED, "run_batch_stage: the batch runs per template (no batched kernels)");
    const std::vector<Stage>* sp = nullptr;
    if ( auto rc = get_stages(plan, 2, &sp) )
        return rc;
    float2* tw = nullptr;
    if ( auto rc = ensure_device_state(plan, &tw) )
        return rc;
    const unsigned part = 1u << stage;
    const CpxT     ct   = plan->st.image_type == FASTFFT_IMAGE_C32H ? CpxT::c32h : CpxT::c64;
    if ( plan->sub_active )
        return run_subframe_batch(plan, image_to_search, op_scale(pre_op), op_scale(intra_op), op_scale(post_op), tw, part);
    if ( part == kBatchFrame )
        return ok( );
    return run_batch_chunks(plan, *sp, image_to_search, ct, op_scale(pre_op), op_scale(intra_op), op_scale(post_op), tw,
                            part);
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

// the C2R stage of the batch's active form (frame or full)
static fastfft_status batch_c2r_stage(fastfft_plan p, const Stage** out) {
    if ( p->batch.empty( ) || p->batch_chunk <= 0 )
        return fail(FASTFFT_UNSUPPORTED, "batch statistics need a batch set on the batched kernels (fastfft_set_batch)");
    if ( p->sub_active ) {
        *out = &p->sub_xc[2];
        return ok( );
    }
    const std::vector<Stage>* sp = nullptr;
    if ( auto rc = get_stages(p, 2, &sp) )
        return rc;
    if ( sp->size( ) != 3 )
        return fail(FASTFFT_UNSUPPORTED, "batch statistics: unexpected stage list");
    *out = &(*sp)[2];
    return ok( );
}

fastfft_status fastfft_set_batch_stats(fastfft_plan plan, float* mip, int* best, float* sum, float* sum_sq, int first_id) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( ! mip && ! best && ! sum && ! sum_sq ) {
        clear_stats(plan);
        free_batch_graph(plan);
        return ok( );
    }
    if ( ! mip || ! best || ! sum || ! sum_sq )
        return fail(FASTFFT_INVALID_ARGUMENT, "set_batch_stats: mip, best, sum and sum_sq are all set or all null");
    const Stage* c2r = nullptr;
    if ( auto rc = batch_c2r_stage(plan, &c2r) )
        return rc;
    if ( ! c2r_stats_covers(*c2r) )
        return fail(FASTFFT_UNSUPPORTED, "set_batch_stats: the statistics form covers 512- and 1024-point frames along x "
                                         "(this batch runs " + std::to_string(c2r->N) + "-point rows)");
    // partial maps of the template groups, allocated once per batch (calls allocate nothing)
    float* part = plan->batch_stats.part;
    if ( ! part ) {
        const size_t bytes = sizeof(float) * 4 * (size_t)kStatsGroups * (size_t)c2r->M * (size_t)c2r->nlines;
        if ( dev_alloc(reinterpret_cast<void**>(&part), bytes) != hipSuccess ) {
            (void)hipGetLastError( );
            part = nullptr; // one group per launch
        }
    }
    plan->batch_stats    = BatchStats{mip, best, sum, sum_sq, part};
    plan->stats_first_id = first_id;
    plan->stats_on       = true;
    free_batch_graph(plan);
    return ok( );
}

fastfft_status fastfft_reset_batch_stats(fastfft_plan plan) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( ! plan->stats_on )
        return fail(FASTFFT_INVALID_ARGUMENT, "reset_batch_stats: no statistics set (fastfft_set_batch_stats)");
    const Stage* c2r = nullptr;
    if ( auto rc = batch_c2r_stage(plan, &c2r) )
        return rc;
    const hipError_t e = launch_stats_reset(plan->batch_stats, (long)c2r->M * c2r->nlines, plan->stream);
    return e == hipSuccess ? ok( ) : hip_fail(e, "reset_batch_stats");
}

// The batch set by fastfft_set_batch: its form, chunk size and internal stage lists.
static std::string batch_description(fastfft_plan plan) {
    std::string s;
    s = "batch templates=" + std::to_string(plan->batch.size( )) + " chunk=" + std::to_string(plan->batch_chunk) +
        " group=" + std::to_string(plan->batch_group) + " p=" + std::to_string(plan->batch_p) +
        " tuned=" + (plan->batch_tuned.empty( ) ? std::string("none") : plan->batch_tuned) +
        " precision=" + (batch_mixed(plan) ? std::string("split_bf16")
                         : plan->precision == FASTFFT_PRECISION_SPLIT_BF16 ? std::string("fp32 (split_bf16 does not apply)")
                                                                           : std::string("fp32"));
    if ( plan->sub_active ) {
        const bool cm = plan->sub_xc.size( ) == 3 && plan->sub_xc[1].im.es == 1 && plan->sub_xc[1].im.s1 > 1;
        s += " frame=" + std::to_string(plan->sub_fx) + "x" + std::to_string(plan->sub_fy) +
             (cm ? " slots=column-major" : " slots=row-major") + "\n";
        s += "frame:\n" + describe(plan->st, plan->sub_inv) + "frame spectrum:\n" + describe(plan->st, plan->sub_fwd) +
             "templates:\n" + describe(plan->st, plan->sub_xc);
    }
    else {
        s += " frame=full\n";
    }
    if ( plan->stats_on )
        s.insert(s.find('\n'), " stats=on first_id=" + std::to_string(plan->stats_first_id));
    return s;
}

fastfft_status fastfft_describe(fastfft_plan plan, const char* op, char* buf, size_t buf_len) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( ! op || ! buf || buf_len == 0 )
        return fail(FASTFFT_INVALID_ARGUMENT, "describe: null argument");
    if ( auto rc = ensure_validated(plan) )
        return rc;
    std::string s;
    if ( std::strcmp(op, "batch") == 0 ) {
        s = batch_description(plan);
    }
    else {
        const int opi = op_index(op);
        if ( opi < 0 )
            return fail(FASTFFT_INVALID_ARGUMENT,
                        std::string("unknown operation '") + op + "' (expected fwd, inv, xcorr or batch)");
        const std::vector<Stage>* sp = nullptr;
        if ( auto rc = get_stages(plan, opi, &sp) )
            return rc;
        s = describe(plan->st, *sp);
    }
    std::snprintf(buf, buf_len, "%s", s.c_str( ));
    if ( s.size( ) + 1 > buf_len )
        return fail(FASTFFT_INVALID_ARGUMENT, "describe: buffer too small");
    return ok( );
}

const char* fastfft_last_error_string(void) { return g_last_error.c_str( ); }

const char* fastfft_version(void) { return "fastfft_amd 0.1.0 (gfx950)"; }

} // extern "C"
