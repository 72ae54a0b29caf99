/*
 * fastfft.h -- C ABI of the MI355X-native FastFFT size-changing transform path.
 *
 * Plain C: opaque plan handle, sizes and device pointers passed as void*,
 * no HIP or torch types in any signature (streams travel as void* holding a
 * hipStream_t).  Every entry point returns a fastfft_status; the message of
 * the last failure on the calling thread is in fastfft_last_error_string().
 *
 * Each entry point replaces one member of the reference's
 * FastFFT::FourierTransformer<ComputeBaseType, InputType, OtherImageType, Rank>
 * (include/FastFFT.h:84-732, src/fastfft/FastFFT.cu); the citation is given
 * with each declaration.  The C++ drop-in header include/fastfft/FastFFT.h and
 * the Python package fastfft_amd (ctypes) are both thin layers over this ABI.
 *
 * Data layout contract (byte-identical to the reference):
 *   real buffers   FFTW padded: rows of 2*(x/2+1) elements (InputType)
 *   2-D spectrum   C[kx][ky]: x/2+1 rows of y complex values, ky contiguous
 *   3-D spectrum   C[kz][kx][ky]
 *   transforms     unnormalised (Inv(Fwd(x)) = Nx*Ny*Nz * x)
 *   CONJ_MUL       image_to_search * conj(FFT(input))
 */
#ifndef FASTFFT_AMD_FASTFFT_H
#define FASTFFT_AMD_FASTFFT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum fastfft_status {
    FASTFFT_SUCCESS           = 0,
    FASTFFT_INVALID_ARGUMENT  = 1, /* null handle/pointer, bad enum */
    FASTFFT_INVALID_DIMENSION = 2, /* the reference's runtime asserts in ValidateDimensions */
    FASTFFT_UNSUPPORTED       = 3, /* valid request with no kernel path (e.g. 3-D fwd decrease) */
    FASTFFT_NOT_PLANNED       = 4, /* a transform called before both plans were set */
    FASTFFT_HIP_ERROR         = 5, /* a HIP runtime call failed */
    FASTFFT_OUT_OF_MEMORY     = 6,
    FASTFFT_LDS_EXCEEDED      = 7  /* a stage would need more than the 160 KiB of LDS per CU */
} fastfft_status;

/* InputType of the reference template (real position-space data). */
typedef enum fastfft_input_type { FASTFFT_INPUT_F32 = 0, FASTFFT_INPUT_F16 = 1 } fastfft_input_type;

/* OtherImageType: storage of image_to_search in FwdImageInvFFT. */
typedef enum fastfft_image_type { FASTFFT_IMAGE_C64 = 0 /* float2 */, FASTFFT_IMAGE_C32H = 1 /* __half2 */ } fastfft_image_type;

/* Element-wise operators fused into the transforms; the reference's functors
 * KernelFunction::my_functor<T, N, IKF_t> (include/detail/functors.h:15-80). */
/* Arithmetic of the correlation pass (fastfft_set_precision). */
typedef enum fastfft_precision {
    FASTFFT_PRECISION_FP32       = 0, /* FP32 radix passes (default, the drop-in path) */
    FASTFFT_PRECISION_SPLIT_BF16 = 1  /* DFT-32 stages as split (hi + lo) BF16 matrix products, FP32 accumulation */
} fastfft_precision;

typedef enum fastfft_op_kind {
    FASTFFT_OP_NOOP                = 0,
    FASTFFT_OP_SCALE               = 1, /* pre/post op: x * scale */
    FASTFFT_OP_CONJ_MUL            = 2, /* intra op: image * conj(transform) */
    FASTFFT_OP_CONJ_MUL_THEN_SCALE = 3  /* intra op: image * conj(transform) * scale */
} fastfft_op_kind;

typedef struct fastfft_op {
    int   kind;  /* fastfft_op_kind */
    float scale; /* used by SCALE and CONJ_MUL_THEN_SCALE */
} fastfft_op;

/* Which dimensions fastfft_get_dims reports. */
typedef enum fastfft_dims_which {
    FASTFFT_FWD_INPUT  = 0,
    FASTFFT_FWD_OUTPUT = 1,
    FASTFFT_INV_INPUT  = 2,
    FASTFFT_INV_OUTPUT = 3
} fastfft_dims_which;

typedef struct fastfft_plan_s* fastfft_plan;

/* FourierTransformer() constructor (FastFFT.cu:58-75).  compute type is always
 * FP32 (the reference's static_assert at :63).  No GPU work happens here. */
fastfft_status fastfft_create(fastfft_plan* plan, int input_type, int image_type, int rank);

/* ~FourierTransformer() / Deallocate (FastFFT.cu:77-113): frees the workspace. */
fastfft_status fastfft_destroy(fastfft_plan plan);

/* SetForwardFFTPlan (FastFFT.cu:132-161, FastFFT.h:102-104). */
fastfft_status fastfft_set_fwd_plan(fastfft_plan plan, size_t in_x, size_t in_y, size_t in_z, size_t out_x, size_t out_y,
                                    size_t out_z, int is_padded_input);

/* SetInverseFFTPlan (FastFFT.cu:180-200, FastFFT.h:106-108). */
fastfft_status fastfft_set_inv_plan(fastfft_plan plan, size_t in_x, size_t in_y, size_t in_z, size_t out_x, size_t out_y,
                                    size_t out_z, int is_padded_output);

/* Extension (no reference counterpart): origin of the window kept by a trimmed
 * inverse, as wrapped lags.  Default (0,0,0) is the reference's top-left trim
 * (FastFFT.cu:1342, :1675); (-out_x/2, -out_y/2, -out_z/2) centres it on lag 0. */
fastfft_status fastfft_set_output_window(fastfft_plan plan, int origin_x, int origin_y, int origin_z);

/* Stream for all work of this plan (the reference always uses
 * cudaStreamPerThread, docs/_docs/references/usage.md:13).  stream is a
 * hipStream_t; NULL selects the per-thread default stream. */
fastfft_status fastfft_set_stream(fastfft_plan plan, void* stream);

/* Extension, no counterpart in the reference: arithmetic of the fused
 * correlation pass.  FASTFFT_PRECISION_SPLIT_BF16 runs the pass on the matrix
 * cores where it applies (1024-point second axis, at most 512 template rows,
 * window origin 0: the frames of a batched correlation whose window and
 * templates fit 1024 x 1024) with about 1e-5 relative error; every other pass
 * stays FP32.  Takes effect for the next transform or fastfft_set_batch. */
fastfft_status fastfft_set_precision(fastfft_plan plan, int precision);

/* FwdFFT(input, output = nullptr, pre_op, intra_op) (FastFFT.cu:263-277, :315-459).
 * out == NULL transforms in place. */
fastfft_status fastfft_fwd(fastfft_plan plan, void* in, void* out, fastfft_op pre_op, fastfft_op intra_op);

/* InvFFT(input, output = nullptr, intra_op, post_op) (FastFFT.cu:279-293, :461-607). */
fastfft_status fastfft_inv(fastfft_plan plan, void* in, void* out, fastfft_op intra_op, fastfft_op post_op);

/* FwdImageInvFFT(input, image_to_search, output = nullptr, pre, intra, post)
 * (FastFFT.cu:295-313, :609-798).  intra must be CONJ_MUL or CONJ_MUL_THEN_SCALE
 * (the reference's EnableIf<HasIntraOpFunctor<IntraOpType>>, FastFFT.h:668). */
fastfft_status fastfft_fwd_image_inv(fastfft_plan plan, void* in, const void* image_to_search, void* out, fastfft_op pre_op,
                                     fastfft_op intra_op, fastfft_op post_op);

/* Extension: FwdImageInvFFT followed by the argmax of the (trimmed) output:
 * the C2R x-pass reduces each row to a key as it stores it and one small
 * kernel reduces the row keys (a separate two-kernel epilogue when that C2R
 * kernel does not apply).  peak_xy (device int[2]) receives (x, y)
 * of the first maximum in row-major order of the out_x by out_y window, value
 * (device float) its value.  Asynchronous on the plan's stream and free of
 * allocations and memsets after the first call on a plan, so it can be
 * captured in a HIP graph.  2-D plans only. */
fastfft_status fastfft_fwd_image_inv_argmax(fastfft_plan plan, void* in, const void* image_to_search, void* out,
                                            fastfft_op pre_op, fastfft_op intra_op, fastfft_op post_op, int* peak_xy,
                                            float* value);

/* Extension with no counterpart in the reference, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): fastfft_fwd_image_inv_argmax with the same
 * arguments and results, whose intermediates and argmax slots live in the
 * library's static device scratch (32 MiB per GPU, held until the process
 * ends) instead of the plan's workspace.  It calls no hipMalloc, hipFree or
 * memset; the caller passes the inputs and outputs.  Plans whose correlation
 * needs more scratch than that (4096^2 with a 512^2 window and a template up
 * to 512^2 needs about 17 MB) get FASTFFT_UNSUPPORTED.  Every call on one GPU
 * shares that GPU's scratch: calls that overlap in time on one GPU (on
 * different streams or host threads) race.  The first call on a device fills
 * the device's twiddle table (one kernel and a stream synchronisation), so it
 * is not made under graph capture; later calls can be captured.  2-D plans only. */
fastfft_status fastfft_fwd_image_inv_argmax_noalloc(fastfft_plan plan, void* in, const void* image_to_search, void* out,
                                                    fastfft_op pre_op, fastfft_op intra_op, fastfft_op post_op,
                                                    int* peak_xy, float* value);

/* Extension (batched correlation): registers n templates for
 * fastfft_fwd_image_inv_batch.  Template i is a real sizes_xy[2i] x sizes_xy[2i+1]
 * device array, FFTW padded to its own pitch 2*(x/2+1), no larger than the
 * plan's forward input size; it is correlated against the image with the
 * plan's padded size, output window and element types, and its window goes to
 * outputs[i] (the plan's inverse output layout).  peaks_xy (device int[2n]) and
 * values (device float[n]) receive each output's argmax, or are both NULL.
 * Synchronous set-up; the arrays are copied, the device buffers must outlive
 * the batch.  Replaces any previous batch; a new plan or window clears it. */
fastfft_status fastfft_set_batch(fastfft_plan plan, int n, const void* const* templates, const int* sizes_xy,
                                 void* const* outputs, int* peaks_xy, float* values);

/* Extension (batched correlation): runs every correlation of the batch set by
 * fastfft_set_batch against image_to_search, asynchronously on the plan's
 * stream; free of allocations after the first call, so it can be captured in a
 * HIP graph.  The ops apply to every correlation as in FwdImageInvFFT. */
fastfft_status fastfft_fwd_image_inv_batch(fastfft_plan plan, const void* image_to_search, fastfft_op pre_op,
                                           fastfft_op intra_op, fastfft_op post_op);

/* Extension (template matching, no counterpart in the reference; for searches
 * that keep per-pixel statistics over all templates instead of every
 * correlation map): with statistics set, each fastfft_fwd_image_inv_batch call
 * updates, for every position p = y * out_x + x of the out_x by out_y window
 * (dense row-major device maps of out_x * out_y elements),
 *   mip[p]    = max(mip[p], out_t[p]) over the batch's templates t in order,
 *   best[p]   = first_id + t of the first template that raises mip[p] (ties keep
 *               the earlier template, and the earlier call),
 *   sum[p]   += out_t[p],  sum_sq[p] += out_t[p]^2,
 * in place of writing the outputs and peaks of fastfft_set_batch (which stay
 * required there but are not written).  Needs a batch on the batched kernels
 * with 512- or 1024-point rows along x after any frame reduction (the
 * describe("batch") line names the form); first_id may be changed between
 * calls by calling this again.  All four NULL switches statistics off;
 * fastfft_set_batch also clears them. */
fastfft_status fastfft_set_batch_stats(fastfft_plan plan, float* mip, int* best, float* sum, float* sum_sq, int first_id);

/* Extension: mip = -infinity, best = -1, sum = sum_sq = 0 over the maps set by
 * fastfft_set_batch_stats, asynchronously on the plan's stream (a fill kernel,
 * no memset, so it can be captured in a HIP graph). */
fastfft_status fastfft_reset_batch_stats(fastfft_plan plan);

/* Wait() (FastFFT.h:124-126): synchronise the plan's stream. */
fastfft_status fastfft_wait(fastfft_plan plan);

/* CopyDeviceToHostAndSynchronize(host, n = 0) (FastFFT.cu:969-1009): copies n
 * InputType elements (0 = the size of the last completed stage) from the buffer
 * holding the last result. */
fastfft_status fastfft_copy_device_to_host_and_synchronize(fastfft_plan plan, void* host, int n_elements);

/* ReturnInputMemorySize / ReturnFwdOutputMemorySize / ReturnInvOutputMemorySize
 * (FastFFT.h:194-204): element counts (InputType elements) of FFTW padded buffers. */
fastfft_status fastfft_get_memory_sizes(fastfft_plan plan, int* input, int* fwd_output, int* inv_output);

/* ReturnFwdInputDimensions ... ReturnInvOutputDimensions (FastFFT.h:206-220):
 * dims[0..3] = x, y, z, w (w = complex elements along x after R2C). */
fastfft_status fastfft_get_dims(fastfft_plan plan, int which, int dims[4]);

/* GetDeviceBufferPointer (FastFFT.h:128-130): the plan's internal workspace
 * (NULL until the first transform allocates it). */
fastfft_status fastfft_get_buffer_pointer(fastfft_plan plan, void** ptr);

/* Bytes of device workspace the plan allocates at its first transform. */
fastfft_status fastfft_workspace_bytes(fastfft_plan plan, size_t* bytes);

/* Extension: the argmax epilogue of the correlation path.  img is a real
 * [ny][pitch] device array of InputType (as written by FwdImageInvFFT);
 * peak[0..1] receive (x, y) of the first maximum in row-major order, value
 * its value (float).  peak and value are host pointers; the call synchronises. */
fastfft_status fastfft_argmax(fastfft_plan plan, const void* img, int nx, int ny, int pitch, int peak[2], float* value);

/* Extension: the argmax epilogue without a host round trip.  Writes (x, y) to
 * device int[2] peak_xy and the value to device float value, asynchronously on
 * the plan's stream. */
fastfft_status fastfft_argmax_device(fastfft_plan plan, const void* img, int nx, int ny, int pitch, int* peak_xy, float* value);

/* Runs only stages [first, last] of an operation ("fwd", "inv" or "xcorr"),
 * the run-time counterpart of the reference's compile-time FFT_DEBUG_STAGE
 * partial transforms (src/tests/helper_functions.cuh:125-198).  Intermediate
 * stages read and write the plan's workspace; stage layouts are listed by
 * fastfft_describe.  Used for per-stage layout tests and per-kernel timing. */
fastfft_status fastfft_run_stages(fastfft_plan plan, const char* op, int first, int last, void* in, const void* image_to_search,
                                  void* out, fastfft_op pre_op, fastfft_op intra_op, fastfft_op post_op);

/* Runs one part of the batched FwdImageInvFFTBatch over the whole batch, for
 * per-kernel timing (no counterpart in the reference): 0 the reduction of the
 * image to the batch's frame (nothing when the batch runs on the full frame),
 * 1 the R2C rows, 2 the correlation pass, 3 the trimmed C2R rows (with the
 * argmax row keys), 4 the argmax reduction.  A part reads what the previous
 * part last left in the batch's buffers, which hold one chunk of templates:
 * the parts run in order give the call's outputs only for a batch of one
 * chunk (fastfft_describe "batch" shows the chunk).  FASTFFT_UNSUPPORTED when
 * the batch runs template by template. */
fastfft_status fastfft_run_batch_stage(fastfft_plan plan, int stage, const void* image_to_search, fastfft_op pre_op,
                                       fastfft_op intra_op, fastfft_op post_op);

/* Human-readable description of the stage sequence an operation would run
 * ("fwd", "inv" or "xcorr"), written without touching the GPU; "batch"
 * describes the batch of fastfft_set_batch (its chunk size, and the frame and
 * internal stage lists when it runs on a frame smaller than the plan's). */
fastfft_status fastfft_describe(fastfft_plan plan, const char* op, char* buf, size_t buf_len);

/* Message of the last failed call on this thread ("" if none). */
const char* fastfft_last_error_string(void);

/* Library version string. */
const char* fastfft_version(void);

#ifdef __cplusplus
}
#endif

#endif /* FASTFFT_AMD_FASTFFT_H */
