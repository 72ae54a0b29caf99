// C ABI of the FastFFT path: plan object, execution of planned stages,
// host memory helpers.  Host code only; kernels live in launch_*.hip.
#include <hip/hip_runtime.h>

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

static void free_batch_device(fastfft_plan p) {
    clear_stats(p); // the statistics belong to the batch's kernels
    if ( p->batch_dev )
        (void)hipFree(p->batch_dev);
    p->batch_dev   = nullptr;
    p->batch_r2c   = nullptr;
    p->batch_xc    = nullptr;
    p->batch_c2r   = nullptr;
    p->batch_chunk = 0;
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
        p->stages[opi]       = std::move(st);
        p->stages_built[opi] = true;
    }
    *out = &p->stages[opi];
    return ok( );
}

static fastfft_status run_stage_list(fastfft_plan p, const std::vector<Stage>& stages, const char* opname, void* in,
                                     void* out, const void* image, float pre, float intra, float post, int first = 0,
                                     int last = -1, const ArgmaxSlots* amax = nullptr, bool* amax_fused = nullptr,
                                     float2* ws_a = nullptr, float2* ws_b = nullptr);

static fastfft_status run_stages(fastfft_plan p, int opi, void* in, void* out, const void* image, float pre, float intra,
                                 float post, int first = 0, int last = -1, const ArgmaxSlots* amax = nullptr,
                                 bool* amax_fused = nullptr) {
    const std::vector<Stage>* sp = nullptr;
    if ( auto rc = get_stages(p, opi, &sp) )
        return rc;
    return run_stage_list(p, *sp, kOpNames[opi], in, out, image, pre, intra, post, first, last, amax, amax_fused);
}

// ws_a / ws_b (both or neither): workspace of an internal stage list instead of the plan's.
static fastfft_status run_stage_list(fastfft_plan p, const std::vector<Stage>& stages, const char* opname, void* in,
                                     void* out, const void* image, float pre, float intra, float post, int first, int last,
                                     const ArgmaxSlots* amax, bool* amax_fused, float2* ws_a, float2* ws_b) {
    if ( last < 0 )
        last = (int)stages.size( ) - 1;
    if ( first < 0 || last >= (int)stages.size( ) || first > last )
        return fail(FASTFFT_INVALID_ARGUMENT,
                    "stage range outside the " + std::to_string(stages.size( )) + " stages of " + opname);
    float2*        tw = nullptr;
    fastfft_status rc = ensure_device_state(p, &tw);
    if ( rc )
        return rc;
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

// ---------------------------------------------------------------- batched correlation
// Chunk sizes chosen by timing, per batch shape, for the life of the process.
static std::mutex                 g_chunk_mutex;
static std::map<std::string, int> g_chunk_cache;

// Elements a line map addresses for nlines lines of nelem elements (single-index lines).
static size_t map_extent(const HostLineMap& m, int nlines, int nelem) {
    return (size_t)((long)(nlines - 1) * m.s1 + (long)(nelem - 1) * m.es + 1);
}
// A slot: the R2C rows' output (nlines template rows of nout values); B slot: the
// correlation pass's output as the C2R pass reads it (nlines rows of the spectrum's
// w values), rounded to even counts so that every slot starts 16-byte aligned.
static size_t slot_elems_a(const Stage& r2c) { return (map_extent(r2c.om, r2c.nlines, r2c.nout) + 1) / 2 * 2; }
static size_t slot_elems_b(const Stage& c2r) { return (map_extent(c2r.im, c2r.nlines, c2r.N / 2 + 1) + 1) / 2 * 2; }

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

// Launches every chunk of the batch: R2C rows, fused columns, C2R rows (+ argmax keys);
// parts outside `parts` are skipped.
static fastfft_status run_batch_chunks(fastfft_plan p, const std::vector<Stage>& st, const void* image, CpxT img_ct,
                                       float pre, float intra, float post, const float2* tw, unsigned parts = kBatchAll) {
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
            io.split_bf16 = p->precision == FASTFFT_PRECISION_SPLIT_BF16;
            e          = launch_xcorr(st[1], io, tw, p->stream);
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

// Picks the chunk size: FASTFFT_BATCH_CHUNK if set, else the largest of 1, 2, 4,
// ... templates per launch whose time on this batch (a zero image) is within 5 %
// of the fastest, cached per batch shape.  Allocates batch_dev for the chosen size and fills the descriptors.
static fastfft_status setup_batch_kernels(fastfft_plan p, const std::vector<Stage>& st, CpxT img_ct) {
    const int n = (int)p->batch.size( );
    float2*   tw = nullptr;
    if ( auto rc = ensure_device_state(p, &tw) )
        return rc;
    std::vector<std::pair<int, int>> sizes;
    for ( const auto& it : p->batch )
        sizes.push_back({it.x, it.y});
    std::sort(sizes.begin( ), sizes.end( ));
    std::string key = std::to_string(st[1].N) + "/" + std::to_string(st[1].P) + "/" + std::to_string(st[0].N) + "/" +
                      std::to_string(st[2].M) + "x" + std::to_string(st[2].nlines) + "/" + std::to_string(st[1].o) + "/" +
                      std::to_string(p->st.input_type) + std::to_string((int)img_ct) + "/" + std::to_string(p->batch_peaks != nullptr) +
                      "/" + std::to_string(p->precision);
    for ( const auto& s : sizes )
        key += ":" + std::to_string(s.first) + "x" + std::to_string(s.second);
    int chosen = env_chunk( );
    if ( chosen <= 0 ) {
        std::lock_guard<std::mutex> lock(g_chunk_mutex);
        auto                        f = g_chunk_cache.find(key);
        if ( f != g_chunk_cache.end( ) )
            chosen = f->second;
    }
    if ( chosen <= 0 ) {
        // candidates up to the batch size and about 1 GiB of slots
        const size_t per = sizeof(float2) * (slot_elems_a(st[0]) + slot_elems_b(st[2]));
        std::vector<int> cand;
        for ( int C = 1; C <= std::min(n, 64); C *= 2 )
            if ( C == 1 || per * (size_t)C <= (size_t(1) << 30) )
                cand.push_back(C);
        const int cmax = cand.back( );
        free_batch_device(p);
        hipError_t e = hipMalloc(&p->batch_dev, batch_dev_bytes(st, cmax, n));
        void*      img = nullptr;
        const size_t img_bytes = (img_ct == CpxT::c32h ? 4 : 8) * (size_t)st[1].gm.s1 * st[1].nlines;
        if ( e == hipSuccess )
            e = hipMalloc(&img, img_bytes);
        if ( e == hipSuccess )
            e = hipMemset(img, 0, img_bytes);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if ( e == hipSuccess )
            e = hipEventCreate(&e0);
        if ( e == hipSuccess )
            e = hipEventCreate(&e1);
        float              best = 0.f;
        std::vector<float> times;
        fastfft_status     run = FASTFFT_SUCCESS; // a failed launch keeps its own message
        chosen                 = 1;
        for ( size_t ci = 0; e == hipSuccess && ci < cand.size( ); ci++ ) {
            p->batch_chunk = cand[ci];
            if ( (run = fill_batch_items(p, st, cand[ci])) != FASTFFT_SUCCESS )
                break;
            // one warm-up run, then the fastest of kTuneReps timed runs (robust to a
            // neighbour's interference on a shared device)
            constexpr int kTuneReps = 5;
            float         ms        = 0.f;
            for ( int rep = 0; rep <= kTuneReps && e == hipSuccess && ! run; rep++ ) {
                e = hipEventRecord(e0, p->stream);
                if ( e == hipSuccess )
                    run = run_batch_chunks(p, st, img, img_ct, 1.f, 1.f, 1.f, tw);
                if ( run )
                    break;
                if ( e == hipSuccess )
                    e = hipEventRecord(e1, p->stream);
                if ( e == hipSuccess )
                    e = hipEventSynchronize(e1);
                float t = 0.f;
                if ( e == hipSuccess )
                    e = hipEventElapsedTime(&t, e0, e1);
                if ( rep == 1 || (rep > 1 && t < ms) )
                    ms = t;
            }
            if ( run )
                break;
            if ( e == hipSuccess )
                times.push_back(ms);
        }
        // the largest chunk within 5 % of the fastest: fewer, larger launches replay
        // better from a graph than eager timing shows (at 64 x 384^2 -> 512^2, chunk 64
        // replays in 286 us against 292-304 us for chunk 32, which times faster eagerly)
        if ( ! run && e == hipSuccess && times.size( ) == cand.size( ) ) {
            best = *std::min_element(times.begin( ), times.end( ));
            for ( size_t ci = 0; ci < cand.size( ); ci++ )
                if ( times[ci] <= 1.05f * best )
                    chosen = cand[ci];
        }
        if ( e0 )
            (void)hipEventDestroy(e0);
        if ( e1 )
            (void)hipEventDestroy(e1);
        if ( img )
            (void)hipFree(img);
        free_batch_device(p);
        if ( run )
            return run;
        if ( e != hipSuccess )
            return hip_fail(e, "batch chunk tuning");
        std::lock_guard<std::mutex> lock(g_chunk_mutex);
        g_chunk_cache[key] = chosen;
    }
    free_batch_device(p);
    hipError_t e = hipMalloc(&p->batch_dev, batch_dev_bytes(st, chosen, n));
    if ( e != hipSuccess ) {
        p->batch_dev = nullptr;
        return hip_fail(e, "hipMalloc(batch)");
    }
    p->batch_chunk = chosen;
    return fill_batch_items(p, st, chosen);
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
    hipError_t   e     = hipMalloc(&p->sub_dev, bytes);
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
        if ( auto rc = run_stage_list(p, fwd, "batch frame spectrum", p->sub_s, p->sub_h, nullptr, 1.f, 1.f, 1.f,
                                      fuse ? 1 : 0, -1, nullptr, nullptr, p->sub_wa, p->sub_wb) )
            return rc;
    }
    return run_batch_chunks(p, p->sub_xc, p->sub_h, CpxT::c64, pre, intra, post, tw, parts);
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
        const CpxT img_ct = plan->st.image_type == FASTFFT_IMAGE_C32H ? CpxT::c32h : CpxT::c64;
        if ( auto rc = plan->sub_active ? run_subframe_batch(plan, image_to_search, op_scale(pre_op), op_scale(intra_op),
                                                             op_scale(post_op), tw)
                                        : run_batch_chunks(plan, *sp, image_to_search, img_ct, op_scale(pre_op),
                                                           op_scale(intra_op), op_scale(post_op), tw) )
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
        if ( hipMalloc(&part, bytes) != hipSuccess ) {
            (void)hipGetLastError( );
            part = nullptr; // one group per launch
        }
    }
    plan->batch_stats    = BatchStats{mip, best, sum, sum_sq, part};
    plan->stats_first_id = first_id;
    plan->stats_on       = true;
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

fastfft_status fastfft_describe(fastfft_plan plan, const char* op, char* buf, size_t buf_len) {
    if ( auto rc = check_plan(plan) )
        return rc;
    if ( ! op || ! buf || buf_len == 0 )
        return fail(FASTFFT_INVALID_ARGUMENT, "describe: null argument");
    if ( auto rc = ensure_validated(plan) )
        return rc;
    std::string s;
    if ( std::strcmp(op, "batch") == 0 ) {
        // the batch set by fastfft_set_batch: its form, chunk size and internal stage lists
        s = "batch templates=" + std::to_string(plan->batch.size( )) + " chunk=" + std::to_string(plan->batch_chunk) +
            " precision=" + (plan->precision == FASTFFT_PRECISION_SPLIT_BF16 ? "split_bf16" : "fp32");
        if ( plan->sub_active ) {
            s += " frame=" + std::to_string(plan->sub_fx) + "x" + std::to_string(plan->sub_fy) + "\n";
            s += "frame:\n" + describe(plan->st, plan->sub_inv) + "frame spectrum:\n" + describe(plan->st, plan->sub_fwd) +
                 "templates:\n" + describe(plan->st, plan->sub_xc);
        }
        else {
            s += " frame=full\n";
        }
        if ( plan->stats_on )
            s.insert(s.find('\n'), " stats=on first_id=" + std::to_string(plan->stats_first_id));
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
