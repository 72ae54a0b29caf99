// Host-side launch interface of the line kernels (implemented in launch_*.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "items.hpp"
#include "plan.hpp"

namespace fastfft_amd {

// Element types of the operands of one stage launch.
enum class RealT { f32 = 0, f16 = 1 };
enum class CpxT { c64 = 0, c32h = 1 };

// Device slots of the argmax that a trimmed C2R x-pass can compute while it
// stores its rows: one key per row (kArgmaxPartials at least), reduced after
// the pass into value and xy.
struct ArgmaxSlots {
    unsigned long long* keys  = nullptr;
    float*              value = nullptr;
    int*                xy    = nullptr;
};

struct StageIO {
    const void* src = nullptr;
    void*       dst = nullptr;
    const void* img = nullptr; // xcorr only
    bool        src_real = false, dst_real = false;
    RealT       src_rt = RealT::f32, dst_rt = RealT::f32;
    CpxT        src_ct = CpxT::c64, dst_ct = CpxT::c64, img_ct = CpxT::c64;
    float       scale = 1.f; // pre (fwd), post (inv), intra (xcorr)
    // inv_c2r only: compute the argmax of the output rows into *amax when the
    // kernel supports it, and set *amax_fused (left unchanged otherwise)
    const ArgmaxSlots* amax       = nullptr;
    bool*              amax_fused = nullptr;
    // xcorr only: correlate with the templates xitems[0 .. nxitems) (device
    // array of XcorrItem, xcorr_col.hpp) instead of src / dst
    const void* xitems  = nullptr;
    int         nxitems = 0;
    // xcorr with xitems: templates per template group along gridDim.y (0: the default)
    int xgroup = 0;
    // xcorr with xitems: elements between the outputs of consecutive items (the batch's B
    // slots are equally spaced), or 0 when they are not
    long xout_step = 0;
    // xcorr only: run the split-BF16 matrix-core kernel where it covers the stage
    // (xcorr_toeplitz_covers, with img_mixed), the FP32 kernels elsewhere
    bool split_bf16 = false;
    // xcorr only: img is the frame in mixed form, the frame's R2C x-pass output R2C_x(s)
    // (column kx of N positions at img + kx * s.gm.s1) instead of its spectrum; only the
    // split-BF16 Toeplitz kernel takes it (the correlation then carries a factor N)
    bool img_mixed = false;
};

// Batched correlation over n templates with the stage parameters of `s` (the
// stages of the largest template); items are device arrays (RowBatchItem,
// rows_wave.hpp; XcorrItem, xcorr_col.hpp).  Each returns hipErrorNotSupported,
// before launching anything, when its fast kernel does not cover the stage.
bool       r2c_rows_batch_covers(const Stage& s);
bool       c2r_rows_batch_covers(const Stage& s);
bool       xcorr_col_covers(const Stage& s, CpxT img_ct);
// complex forward / inverse line stages over adjacent lines with far-apart
// elements (3-D z-passes), N = 256 or 512: tiled kernel (tile_lines.hpp);
// hipErrorNotSupported, before launching anything, when it does not cover s
bool       tile_lines_cover(const Stage& s);
hipError_t launch_tile_lines(const Stage& s, const StageIO& io, const float2* tw, hipStream_t stream);
// the split-BF16 matrix-core correlation pass (xcorr_toeplitz.hpp) on a frame in mixed
// form: N = 1024, L <= 512, M <= 512, L + M - 1 <= 1024, window origin 0, the DC and
// Nyquist columns of a real transform at both ends, row-major template and output
// columns (unit line stride), unit-stride image columns
bool       xcorr_toeplitz_covers(const Stage& s);
bool       batch_kernels_cover(const Stage& r2c, const Stage& xcorr, const Stage& c2r, CpxT img_ct);
hipError_t launch_r2c_rows_batch(const Stage& s, const void* items, int n, int max_rows, RealT in_rt, float pre,
                                 const float2* tw, hipStream_t stream);
hipError_t launch_c2r_rows_batch(const Stage& s, const void* items, int n, RealT out_rt, float post, const float2* tw,
                                 hipStream_t stream);

// Per-pixel template-matching statistics of the batch's trimmed outputs (dense
// M x nlines maps of the C2R stage, index row * M + j): running maximum, the
// template id of its first occurrence, sum and sum of squares.
struct BatchStats {
    float* mip    = nullptr;
    int*   best   = nullptr;
    float* sum    = nullptr;
    float* sum_sq = nullptr;
    // kStatsGroups x 4 partial maps for template groups (null: one group per launch)
    float* part = nullptr;
};
constexpr int kStatsGroups = 8;
// the packed C2R rows (N = 512 or 1024) run the statistics form
bool       c2r_stats_covers(const Stage& s);
hipError_t launch_c2r_stats_batch(const Stage& s, const void* items, int n, float post, const float2* tw, int first_id,
                                  const BatchStats& bs, hipStream_t stream);
// mip = -inf, best = -1, sum = sum_sq = 0 over n positions
hipError_t launch_stats_reset(const BatchStats& bs, long n, hipStream_t stream);

hipError_t launch_fwd(const Stage& s, const StageIO& io, const float2* tw, hipStream_t stream);
hipError_t launch_inv(const Stage& s, const StageIO& io, const float2* tw, hipStream_t stream);

// C2R stage c (2048-8192 points, M <= 1024 outputs) and the 1024-point R2C stage f
// of the same rows in one kernel (the image's reduction to a correlation frame and
// the frame's forward rows): src is c's input, dst f's output, the real rows are
// not stored.  hipErrorNotSupported when the pair does not fit.
bool       c2r_r2c_rows_covers(const Stage& c, const Stage& f);
hipError_t launch_c2r_r2c_rows(const Stage& c, const Stage& f, const float2* src, float2* dst, float post, float pre,
                               const float2* tw, hipStream_t stream);
hipError_t launch_xcorr(const Stage& s, const StageIO& io, const float2* tw, hipStream_t stream);

// Entries of the twiddle table: W_8192^k for k < 8192, then the
// multi-resolution part (W_n^m at 8192 + n + m for every power of two n <= 8192).
constexpr int kTwiddleTableN = 3 * 8192;

// Fills the twiddle table (FP64 on the GPU, rounded to FP32).
hipError_t launch_twiddle_init(float2* tw, hipStream_t stream);
// Device address of the current device's twiddle table (a static device array of
// kTwiddleTableN entries; launch_twiddle_init fills it).
hipError_t twiddle_table_address(float2** out);

// Device scratch of the allocation-free correlation entry point: a static device array
// of kStaticScratchBytes per GPU (scratch.hip).  *out receives the current device's
// copy; the address is looked up once per device and kept.
constexpr size_t kStaticScratchBytes = size_t(32) << 20;
hipError_t       static_scratch_for_current_device(void** out);

// Argmax of a real [ny][pitch] image (first maximum in row-major order).  Two
// kernels; partial_v / partial_i hold kArgmaxPartials entries of device scratch;
// the value goes to result_value[0] and (x, y) to result_xy[0..1].
constexpr int kArgmaxPartials = 512;
hipError_t    launch_argmax(const void* img, RealT rt, int nx, int ny, int pitch, float* partial_v, int* partial_i,
                            float* result_value, int* result_xy, hipStream_t stream);

// Reduces n argmax keys (fft_core.hpp) to the value and (x, y) = (index mod nx, index / nx);
// with batch > 1, set b reduces keys[b * key_stride ..] into result_value[b], result_xy[2b..].
hipError_t launch_argmax_keys(const unsigned long long* keys, int n, int nx, float* result_value, int* result_xy,
                              hipStream_t stream, int batch = 1, long key_stride = 0);

// LDS bytes per workgroup of a line stage with sub-transform size P.
size_t stage_lds_bytes(int P);

// FASTFFT_LOG=1 (read once per process): stderr lines naming each plan's stage
// lists when they are built, the batch form when it is set, and every kernel
// launch from the host with its template arguments, grid, block and LDS bytes.
bool log_enabled( );
void log_text(const std::string& text);
// tag: typeid(KernelTag<kernel>).name() (launch_common.hpp)
void log_launch(const char* tag, dim3 grid, dim3 block, size_t lds);


} // namespace fastfft_amd
