// FastFFT.h -- drop-in C++ interface of the MI355X-native FastFFT path.
//
// Keeps the reference's public class
//   FastFFT::FourierTransformer<ComputeBaseType, InputType, OtherImageType, Rank>
// (include/FastFFT.h:84-732) with the same member names, signatures, default
// arguments, buffer layouts and functor types
// (FastFFT::KernelFunction::my_functor<T, N_ARGS, IKF_t>, include/detail/functors.h:15-80),
// so code written against the reference compiles unchanged against this
// header.  Every member forwards to the C ABI of include/fastfft/fastfft.h,
// which runs hand-written HIP kernels for gfx950; this header itself is plain
// host C++ (it needs only the HIP runtime API and fp16 type headers).
//
// Instantiations provided by the library (FastFFT.cu:3371-3378):
//   <float, float,  float2,  2>  <float, __half, float2,  2>
//   <float, float,  __half2, 2>  <float, __half, __half2, 2>
//   <float, float,  float2,  3>  <float, __half, __half2, 3>
// (the header accepts any float/__half x float2/__half2 combination for Rank 2 and 3).
//
// Errors: like the reference's MyFFTRunTimeAssert* (include/detail/checks_and_debug.h:84-87),
// a failed call prints the library's message and exits; define
// FASTFFT_THROW_ON_ERROR before including this header to get a FastFFT::Error
// exception instead.
#ifndef FASTFFT_AMD_FASTFFT_HPP
#define FASTFFT_AMD_FASTFFT_HPP

#ifndef __HIP_PLATFORM_AMD__
#define __HIP_PLATFORM_AMD__ 1
#endif
#include <hip/hip_fp16.h>
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdlib>
#include <iostream>
#include <random>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "fastfft.h"

namespace FastFFT {

namespace KernelFunction {

// Intra-kernel function types (include/detail/concepts.h:21-24).
enum IKF_t { NOOP, SCALE, CONJ_MUL, CONJ_MUL_THEN_SCALE };

// Host-side descriptors of the reference's functors.  Active functors are
// `final`, as in the reference, so IS_IKF_t / HasIntraOpFunctor keep working.
template <typename T, int N_ARGS, IKF_t U, typename = void>
struct my_functor {};

template <typename T>
struct my_functor<T, 0, NOOP> {
    fastfft_op as_op( ) const { return fastfft_op{FASTFFT_OP_NOOP, 1.f}; }
};

template <typename T>
struct my_functor<T, 4, CONJ_MUL> final {
    fastfft_op as_op( ) const { return fastfft_op{FASTFFT_OP_CONJ_MUL, 1.f}; }
};

template <typename T>
struct my_functor<T, 4, CONJ_MUL_THEN_SCALE> final {
    explicit my_functor(const T& scale_factor) : scale_factor(scale_factor) {}
    fastfft_op as_op( ) const { return fastfft_op{FASTFFT_OP_CONJ_MUL_THEN_SCALE, static_cast<float>(scale_factor)}; }

  private:
    const T scale_factor;
};

template <typename T>
struct my_functor<T, 2, SCALE> final {
    const T scale_factor;
    explicit my_functor(const T& scale_factor) : scale_factor(scale_factor) {}
    fastfft_op as_op( ) const { return fastfft_op{FASTFFT_OP_SCALE, static_cast<float>(scale_factor)}; }
};

} // namespace KernelFunction

template <typename K>
constexpr bool IS_IKF_t( ) {
    return std::is_final_v<K>;
}

template <typename IntraOpType>
constexpr bool HasIntraOpFunctor = IS_IKF_t<IntraOpType>( );

class Error : public std::runtime_error {
  public:
    Error(fastfft_status s, const std::string& m) : std::runtime_error(m), status(s) {}
    fastfft_status status;
};

namespace detail {

inline void check(fastfft_status s, const char* where) {
    if ( s == FASTFFT_SUCCESS )
        return;
    std::string msg = std::string(where) + ": " + fastfft_last_error_string( );
#ifdef FASTFFT_THROW_ON_ERROR
    throw Error(s, msg);
#else
    std::cerr << "FastFFT error (status " << static_cast<int>(s) << ") in " << msg << std::endl;
    std::exit(-1);
#endif
}

template <class F>
fastfft_op to_op(const F& f) {
    if constexpr ( std::is_same_v<F, std::nullptr_t> )
        return fastfft_op{FASTFFT_OP_NOOP, 1.f};
    else
        return f.as_op( );
}

template <class T>
constexpr int input_type_id( ) {
    static_assert(std::is_same_v<T, float> || std::is_same_v<T, __half>, "Input base type must be either __half or float");
    return std::is_same_v<T, float> ? FASTFFT_INPUT_F32 : FASTFFT_INPUT_F16;
}

template <class T>
constexpr int image_type_id( ) {
    static_assert(std::is_same_v<T, float2> || std::is_same_v<T, __half2>, "OtherImageType must be float2 or __half2");
    return std::is_same_v<T, float2> ? FASTFFT_IMAGE_C64 : FASTFFT_IMAGE_C32H;
}

inline short4 to_short4(const int d[4]) {
    return make_short4(static_cast<short>(d[0]), static_cast<short>(d[1]), static_cast<short>(d[2]), static_cast<short>(d[3]));
}

} // namespace detail

template <class ComputeBaseType = float, class InputType = float, class OtherImageType = float2, int Rank = 2>
class FourierTransformer {
    static_assert(std::is_same_v<ComputeBaseType, float>, "Compute base type must be float");
    static_assert(Rank == 2 || Rank == 3, "Only 2D and 3D FFTs are supported");

  public:
    FourierTransformer( ) {
        detail::check(fastfft_create(&plan_, detail::input_type_id<InputType>( ), detail::image_type_id<OtherImageType>( ), Rank),
                      "FourierTransformer()");
    }

    ~FourierTransformer( ) { fastfft_destroy(plan_); }

    FourierTransformer(const FourierTransformer&)            = delete;
    FourierTransformer& operator=(const FourierTransformer&) = delete;
    FourierTransformer(FourierTransformer&&)                 = delete;
    FourierTransformer& operator=(FourierTransformer&&)      = delete;

    void SetForwardFFTPlan(size_t input_logical_x_dimension, size_t input_logical_y_dimension, size_t input_logical_z_dimension,
                           size_t output_logical_x_dimension, size_t output_logical_y_dimension, size_t output_logical_z_dimension,
                           bool is_padded_output = true) {
        detail::check(fastfft_set_fwd_plan(plan_, input_logical_x_dimension, input_logical_y_dimension, input_logical_z_dimension,
                                           output_logical_x_dimension, output_logical_y_dimension, output_logical_z_dimension,
                                           is_padded_output ? 1 : 0),
                      "SetForwardFFTPlan");
    }

    void SetInverseFFTPlan(size_t input_logical_x_dimension, size_t input_logical_y_dimension, size_t input_logical_z_dimension,
                           size_t output_logical_x_dimension, size_t output_logical_y_dimension, size_t output_logical_z_dimension,
                           bool is_padded_output = true) {
        detail::check(fastfft_set_inv_plan(plan_, input_logical_x_dimension, input_logical_y_dimension, input_logical_z_dimension,
                                           output_logical_x_dimension, output_logical_y_dimension, output_logical_z_dimension,
                                           is_padded_output ? 1 : 0),
                      "SetInverseFFTPlan");
    }

    // Extension: origin (wrapped lags) of the window a trimmed inverse keeps.
    void SetOutputWindow(int origin_x, int origin_y, int origin_z = 0) {
        detail::check(fastfft_set_output_window(plan_, origin_x, origin_y, origin_z), "SetOutputWindow");
    }

    // Extension: stream for all work (default: the per-thread default stream).
    void SetStream(hipStream_t stream) { detail::check(fastfft_set_stream(plan_, stream), "SetStream"); }

    // Extension: arithmetic of the correlation pass (FASTFFT_PRECISION_FP32 or
    // FASTFFT_PRECISION_SPLIT_BF16, see fastfft_set_precision).
    void SetPrecision(fastfft_precision precision) {
        detail::check(fastfft_set_precision(plan_, (int)precision), "SetPrecision");
    }

    inline void Wait( ) { detail::check(fastfft_wait(plan_), "Wait"); }

    auto inline GetDeviceBufferPointer( ) {
        void* p = nullptr;
        detail::check(fastfft_get_buffer_pointer(plan_, &p), "GetDeviceBufferPointer");
        return static_cast<float2*>(p);
    }

    // The reference removed host-to-device copies from the class (FastFFT.cu:252-261).
    void CopyHostToDeviceAndSynchronize(InputType*, int = 0) { RemovedHostToDevice( ); }

    void CopyHostToDevice(InputType*, int = 0) { RemovedHostToDevice( ); }

    void CopyDeviceToHostAndSynchronize(InputType* input_pointer, int n_elements_to_copy = 0) {
        detail::check(fastfft_copy_device_to_host_and_synchronize(plan_, input_pointer, n_elements_to_copy),
                      "CopyDeviceToHostAndSynchronize");
    }

    template <class PreOpType = std::nullptr_t, class IntraOpType = std::nullptr_t>
    void FwdFFT(InputType* input_ptr, InputType* output_ptr = nullptr, PreOpType pre_op = nullptr, IntraOpType intra_op = nullptr) {
        detail::check(fastfft_fwd(plan_, input_ptr, output_ptr, detail::to_op(pre_op), detail::to_op(intra_op)), "FwdFFT");
    }

    template <class IntraOpType = std::nullptr_t, class PostOpType = std::nullptr_t>
    void InvFFT(InputType* input_ptr, InputType* output_ptr = nullptr, IntraOpType intra_op = nullptr, PostOpType post_op = nullptr) {
        detail::check(fastfft_inv(plan_, input_ptr, output_ptr, detail::to_op(intra_op), detail::to_op(post_op)), "InvFFT");
    }

    template <class PreOpType = std::nullptr_t, class IntraOpType = std::nullptr_t, class PostOpType = std::nullptr_t>
    void FwdImageInvFFT(InputType* input_ptr, OtherImageType* image_to_search, InputType* output_ptr = nullptr,
                        PreOpType pre_op = nullptr, IntraOpType intra_op = nullptr, PostOpType post_op = nullptr) {
        static_assert(HasIntraOpFunctor<IntraOpType>, "FwdImageInvFFT needs an intra op functor (CONJ_MUL or CONJ_MUL_THEN_SCALE)");
        detail::check(fastfft_fwd_image_inv(plan_, input_ptr, image_to_search, output_ptr, detail::to_op(pre_op),
                                            detail::to_op(intra_op), detail::to_op(post_op)),
                      "FwdImageInvFFT");
    }

    // Extension: FwdImageInvFFT plus the argmax of the output window into device
    // int[2] peak_xy (x, y) and device float value, asynchronously (2-D plans).
    template <class PreOpType = std::nullptr_t, class IntraOpType = std::nullptr_t, class PostOpType = std::nullptr_t>
    void FwdImageInvFFTArgMax(InputType* input_ptr, OtherImageType* image_to_search, InputType* output_ptr, int* peak_xy,
                              float* value, PreOpType pre_op = nullptr, IntraOpType intra_op = nullptr,
                              PostOpType post_op = nullptr) {
        static_assert(HasIntraOpFunctor<IntraOpType>, "FwdImageInvFFT needs an intra op functor (CONJ_MUL or CONJ_MUL_THEN_SCALE)");
        detail::check(fastfft_fwd_image_inv_argmax(plan_, input_ptr, image_to_search, output_ptr, detail::to_op(pre_op),
                                                   detail::to_op(intra_op), detail::to_op(post_op), peak_xy, value),
                      "FwdImageInvFFTArgMax");
    }

    // Extension: register n templates (device arrays of sizes_xy[2i] x sizes_xy[2i+1],
    // FFTW padded, no larger than the forward input size) with their output buffers
    // and optional device peak / value arrays for FwdImageInvFFTBatch.
    void SetBatch(int n, const InputType* const* templates, const int* sizes_xy, InputType* const* outputs,
                  int* peaks_xy = nullptr, float* values = nullptr) {
        detail::check(fastfft_set_batch(plan_, n, reinterpret_cast<const void* const*>(templates), sizes_xy,
                                        reinterpret_cast<void* const*>(outputs), peaks_xy, values),
                      "SetBatch");
    }

    // Extension: every correlation of the batch against image_to_search, asynchronously.
    template <class PreOpType = std::nullptr_t, class IntraOpType = std::nullptr_t, class PostOpType = std::nullptr_t>
    void FwdImageInvFFTBatch(OtherImageType* image_to_search, PreOpType pre_op = nullptr, IntraOpType intra_op = nullptr,
                             PostOpType post_op = nullptr) {
        static_assert(HasIntraOpFunctor<IntraOpType>, "FwdImageInvFFT needs an intra op functor (CONJ_MUL or CONJ_MUL_THEN_SCALE)");
        detail::check(fastfft_fwd_image_inv_batch(plan_, image_to_search, detail::to_op(pre_op), detail::to_op(intra_op),
                                                  detail::to_op(post_op)),
                      "FwdImageInvFFTBatch");
    }

    // Extension (template matching): per-pixel maximum, its template id, sum and sum
    // of squares over the batch's correlations, updated by FwdImageInvFFTBatch in
    // place of the outputs (fastfft_set_batch_stats); all null switches it off.
    void SetBatchStats(float* mip, int* best, float* sum, float* sum_sq, int first_id = 0) {
        detail::check(fastfft_set_batch_stats(plan_, mip, best, sum, sum_sq, first_id), "SetBatchStats");
    }

    // Extension: mip = -inf, best = -1, sums = 0, asynchronously (fastfft_reset_batch_stats).
    void ResetBatchStats( ) { detail::check(fastfft_reset_batch_stats(plan_), "ResetBatchStats"); }

    // Extension: argmax of a real [ny][pitch] device image, (x, y) and value.
    void ArgMax(const InputType* img, int nx, int ny, int pitch, int& peak_x, int& peak_y, float& value) {
        int pk[2];
        detail::check(fastfft_argmax(plan_, img, nx, ny, pitch, pk, &value), "ArgMax");
        peak_x = pk[0];
        peak_y = pk[1];
    }

    int inline ReturnInputMemorySize( ) { return Sizes( )[0]; }

    int inline ReturnFwdOutputMemorySize( ) { return Sizes( )[1]; }

    int inline ReturnInvOutputMemorySize( ) { return Sizes( )[2]; }

    short4 inline ReturnFwdInputDimensions( ) { return Dims(FASTFFT_FWD_INPUT); }

    short4 inline ReturnFwdOutputDimensions( ) { return Dims(FASTFFT_FWD_OUTPUT); }

    short4 inline ReturnInvInputDimensions( ) { return Dims(FASTFFT_INV_INPUT); }

    short4 inline ReturnInvOutputDimensions( ) { return Dims(FASTFFT_INV_OUTPUT); }

    // Host helpers (FastFFT.h:222-249); device pointers are rejected as in the reference.
    template <typename T, bool is_on_host = true>
    void SetToConstant(T* input_pointer, int N_values, const T& wanted_value) {
        if ( is_on_host ) {
            for ( int i = 0; i < N_values; i++ )
                input_pointer[i] = wanted_value;
        }
        else {
            std::exit(-1);
        }
    }

    template <typename T, bool is_on_host = true>
    void SetToRandom(T* input_pointer, int N_values, const T& wanted_mean, const T& wanted_stddev) {
        std::random_device rd;
        std::mt19937       rng(rd( ));
        rng.seed(static_cast<std::mt19937::result_type>(std::chrono::high_resolution_clock::now( ).time_since_epoch( ).count( )));
        if ( is_on_host ) {
            std::normal_distribution<T> dist{wanted_mean, wanted_stddev};
            for ( int i = 0; i < N_values; i++ )
                input_pointer[i] = dist(rng);
        }
        else {
            std::exit(-1);
        }
    }

    void PrintState( ) {
        char buf[8192];
        for ( const char* op : {"fwd", "inv", "xcorr"} ) {
            if ( fastfft_describe(plan_, op, buf, sizeof(buf)) == FASTFFT_SUCCESS )
                std::cerr << op << ":\n" << buf;
            else
                std::cerr << op << ": " << fastfft_last_error_string( ) << "\n";
        }
    }

    fastfft_plan c_handle( ) const { return plan_; }

  private:
    struct Triple {
        int v[3];
        int operator[](int i) const { return v[i]; }
    };

    Triple Sizes( ) {
        Triple t{ };
        detail::check(fastfft_get_memory_sizes(plan_, &t.v[0], &t.v[1], &t.v[2]), "Return*MemorySize");
        return t;
    }

    short4 Dims(int which) {
        int d[4];
        detail::check(fastfft_get_dims(plan_, which, d), "Return*Dimensions");
        return detail::to_short4(d);
    }

    [[noreturn]] void RemovedHostToDevice( ) {
        std::cerr << "FastFFT: CopyHostToDevice has been removed; copy into device memory yourself" << std::endl;
        std::exit(-1);
    }

    fastfft_plan plan_ = nullptr;
};

} // namespace FastFFT

#endif // FASTFFT_AMD_FASTFFT_HPP
