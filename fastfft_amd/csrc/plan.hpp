// Plan state and the GPU-free planner.
//
// The planner turns the two plan calls of a FourierTransformer into validated
// sizes (the reference's ValidateDimensions / ReturnPaddedMemorySize /
// GetTransformSize, src/fastfft/FastFFT.cu:803-930, :2951-3035,
// include/FastFFT.h:363-404, :699-724) and into an explicit list of line
// stages per operation.  It allocates nothing and makes no HIP call, so it is
// exercised by the CPU test-suite through fastfft_describe().
#pragma once

#include <cstddef>
#include <string>
#include <vector>

namespace fastfft_amd {

// Host mirror of LineMap (line_kernels.hpp): element e of line l lives at
// (l / n2) * s1 + (l % n2) * s2 + e * es.
struct HostLineMap {
    long s1 = 0, s2 = 0, es = 1;
    int  n2 = 1;
};

enum class SizeChange { increase = 0, decrease = 1, no_change = 2 }; // FastFFT.h:326 names

enum class StageKind {
    fwd_real, // input-pruned forward of real lines (R2C x-pass)
    fwd_cpx,  // input-pruned forward of complex lines (C2C y/z pass)
    inv_cpx,  // output-pruned inverse of complex lines (C2C y/z pass)
    inv_c2r,  // output-pruned inverse of Hermitian half lines to real (C2R x-pass)
    xcorr     // fused forward * conj-mul * inverse (second/third axis of FwdImageInvFFT)
};

// Buffers a stage reads or writes.
enum class Buf { user_in, user_out, image, ws_a, ws_b };

// Element storage of a stage operand.  The planner emits the first four;
// real_f32 marks an FP32 real operand whatever the plan's input type (internal
// stage lists whose intermediates stay FP32).
enum class Elem { real_input_type, cpx_input_type, cpx_f32, cpx_image_type, real_f32 };

struct Stage {
    StageKind   kind;
    std::string ref_kernel; // reference kernel this stage replaces (for describe())
    int         N = 0, P = 0, Q = 0; // full size, sub-transform size, N/P
    int         L = 0;               // nonzero inputs (forward / xcorr)
    int         M = 0;               // outputs kept (inverse / xcorr)
    int         o = 0;               // window origin (inverse / xcorr)
    int         nout   = 0;          // outputs stored (forward)
    int         nlines = 0;
    Buf         src = Buf::user_in, dst = Buf::user_out;
    Elem        src_elem = Elem::cpx_f32, dst_elem = Elem::cpx_f32;
    HostLineMap im, om, gm; // input, output and (xcorr) image maps
    // xcorr over the kx columns of a real 2-D transform with an even x size:
    // lines 0 and nlines-1 are the DC and Nyquist columns, whose template
    // values are real and of whose outputs the C2R x-pass uses only the real part
    bool real_end_lines = false;
};

struct Dims {
    int x = 0, y = 0, z = 0, w = 0; // w = x/2+1 complex along x (FastFFT.h:379)
};

struct PlanState {
    int  input_type = 0; // fastfft_input_type
    int  image_type = 0; // fastfft_image_type
    int  rank       = 2; // template Rank
    bool fwd_set = false, inv_set = false;
    bool padded   = true;
    Dims fwd_in, fwd_out, inv_in, inv_out;
    int  origin[3] = {0, 0, 0};

    // derived by validate()
    bool       validated = false;
    int        transform_dimension = 0;
    SizeChange fwd_change = SizeChange::no_change, inv_change = SizeChange::no_change;
    long       input_memory = 0, fwd_output_memory = 0, inv_output_memory = 0, compute_memory = 0;
};

// Status codes mirror fastfft_status.
struct PlanError {
    int         code = 0;
    std::string msg;
};

// FFTW padded element count (ReturnPaddedMemorySize, FastFFT.h:363-404); sets d.w.
long padded_memory_size(Dims& d);

// Validation of both plans (ValidateDimensions, FastFFT.cu:803-930).
PlanError validate(PlanState& p);

// Stage lists.  op: "fwd", "inv" or "xcorr".
PlanError build_stages(const PlanState& p, const std::string& op, std::vector<Stage>& stages);

// Workspace (ws_a, ws_b) complex-float counts needed by all operations of p.
void workspace_elems(const PlanState& p, size_t& ws_a, size_t& ws_b);
// Raises ws_a / ws_b to the complex-float counts one stage list writes.
void stage_list_workspace(const std::vector<Stage>& st, size_t& ws_a, size_t& ws_b);

// Sub-transform size for a line of N with L nonzero inputs and M kept outputs.
int pick_sub_size(int N, int L, int M);

std::string describe(const PlanState& p, const std::vector<Stage>& stages);

constexpr int kMaxSize2D = 8192; // twiddle table / largest line transform
constexpr int kMaxSize3D = 512;  // FastFFT.cu:875-888
constexpr int kMinSize   = 2;

} // namespace fastfft_amd
