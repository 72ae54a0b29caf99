// C++ drop-in test: the reference's own known-answer tests, written against
// include/fastfft/FastFFT.h exactly as src/tests/*.cu use FourierTransformer.
// Plain host C++ (compiled with g++); device memory via the HIP runtime.
//
//   round_trip      src/tests/non_cuda_compilation_unit.cu:9-85
//   unit_impulse    src/tests/unit_impulse_test.cu:4-218 (increase, in and out of place)
//   convolution     src/tests/padded_convolution_FastFFT_vs_cuFFT.cu:179-265
//   scale           src/tests/preop_functor_scale_test.cu:229-364
//   half_io         the same scale test with <float, __half, __half2, 2>
//
// Usage: test_fastfft_api [all|round_trip|unit_impulse|convolution|scale|half_io]
// Prints one PASS/FAIL line per test; exit status 0 when all selected tests pass.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fastfft/FastFFT.h"

#define HIP_OK(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if ( e_ != hipSuccess ) {                                                                    \
            std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
            std::exit(2);                                                                            \
        }                                                                                            \
    } while ( 0 )

static int n_fail = 0;

static void report(const std::string& name, bool ok, const std::string& detail = "") {
    std::printf("%s %s %s\n", ok ? "PASS" : "FAIL", name.c_str( ), detail.c_str( ));
    std::fflush(stdout);
    if ( ! ok )
        n_fail++;
}

static const std::vector<int> kSizes = {32, 64, 128, 256, 512, 1024, 2048, 4096}; // src/tests/tests.h:11

static bool test_round_trip( ) {
    const int                                               input_size = 64;
    FastFFT::FourierTransformer<float, float, float2, 2> FT;
    FT.SetForwardFFTPlan(input_size, input_size, 1, input_size, input_size, 1);
    FT.SetInverseFFTPlan(input_size, input_size, 1, input_size, input_size, 1);
    int n_in  = FT.ReturnInputMemorySize( );
    int n_out = FT.ReturnInvOutputMemorySize( );
    if ( n_in != n_out || n_in != (input_size + 2) * input_size )
        return false;
    std::vector<float> host(n_in, 0.f);
    FT.SetToConstant(host.data( ), n_in, 3.14f);
    for ( float v : host )
        if ( v != 3.14f )
            return false;
    std::fill(host.begin( ), host.end( ), 0.f);
    host[0]  = 1.0f;
    float* d = nullptr;
    HIP_OK(hipMalloc(&d, sizeof(float) * n_in));
    HIP_OK(hipMemcpy(d, host.data( ), sizeof(float) * n_in, hipMemcpyHostToDevice));
    FT.FwdFFT(d);
    FT.InvFFT(d);
    FT.Wait( );
    std::vector<float> out(n_out, -1.f);
    HIP_OK(hipMemcpy(out.data( ), d, sizeof(float) * n_out, hipMemcpyDeviceToHost));
    HIP_OK(hipFree(d));
    return out[0] == float(input_size * input_size);
}

static bool test_unit_impulse(bool inplace) {
    bool ok = true;
    for ( size_t i = 0; i + 1 < kSizes.size( ); i++ ) {
        for ( size_t o = i + 1; o < kSizes.size( ); o++ ) {
            const int                                            s = kSizes[i], b = kSizes[o];
            FastFFT::FourierTransformer<float, float, float2, 2> FT;
            FT.SetForwardFFTPlan(s, s, 1, b, b, 1);
            FT.SetInverseFFTPlan(b, b, 1, b, b, 1);
            short4 din = FT.ReturnFwdInputDimensions( );
            int    n   = std::max(FT.ReturnInputMemorySize( ), FT.ReturnFwdOutputMemorySize( ));
            n          = std::max(n, FT.ReturnInvOutputMemorySize( ));
            std::vector<float> host(n, 0.f);
            host[0]  = 1.0f;
            float *a = nullptr, *c = nullptr;
            HIP_OK(hipMalloc(&a, sizeof(float) * n));
            HIP_OK(hipMemset(a, 0, sizeof(float) * n));
            HIP_OK(hipMemcpy(a, host.data( ), sizeof(float) * FT.ReturnInputMemorySize( ), hipMemcpyHostToDevice));
            if ( ! inplace ) {
                HIP_OK(hipMalloc(&c, sizeof(float) * n));
                HIP_OK(hipMemset(c, 0, sizeof(float) * n));
            }
            if ( inplace )
                FT.FwdFFT(a);
            else
                FT.FwdFFT(a, c);
            std::vector<float> spec(FT.ReturnFwdOutputMemorySize( ));
            FT.CopyDeviceToHostAndSynchronize(spec.data( ));
            double sum = 0;
            for ( size_t k = 0; k + 1 < spec.size( ); k += 2 )
                sum += std::sqrt(double(spec[k]) * spec[k] + double(spec[k + 1]) * spec[k + 1]);
            ok = ok && sum == double(FT.ReturnFwdOutputMemorySize( ) / 2) && din.w == s / 2 + 1;
            if ( inplace )
                FT.InvFFT(a);
            else
                FT.InvFFT(c, a);
            std::vector<float> back(FT.ReturnInvOutputMemorySize( ));
            FT.CopyDeviceToHostAndSynchronize(back.data( ));
            double rs = 0;
            for ( int y = 0; y < b; y++ )
                for ( int x = 0; x < b; x++ )
                    rs += back[y * (b + 2) + x];
            ok = ok && rs == double(b) * b;
            HIP_OK(hipFree(a));
            if ( c )
                HIP_OK(hipFree(c));
        }
    }
    return ok;
}

template <class InputType, class ImageType>
static bool convolution_case(int s, int b, int mode) {
    // mode 0: no scale (expect b*b*6 at the origin, exact zeros elsewhere)
    // mode 1: SCALE pre op, mode 2: CONJ_MUL_THEN_SCALE (expect exactly 1.0 at the origin)
    using FT_t = FastFFT::FourierTransformer<float, InputType, ImageType, 2>;
    FT_t   FT, target;
    FT.SetForwardFFTPlan(s, s, 1, b, b, 1);
    FT.SetInverseFFTPlan(b, b, 1, b, b, 1);
    target.SetForwardFFTPlan(b, b, 1, b, b, 1);
    target.SetInverseFFTPlan(b, b, 1, b, b, 1);
    const int nt = target.ReturnFwdOutputMemorySize( );
    const int n  = std::max(FT.ReturnInputMemorySize( ), FT.ReturnInvOutputMemorySize( ));
    std::vector<InputType> ht(nt, InputType(0.f)), hx(n, InputType(0.f));
    ht[0] = InputType(3.0f);
    hx[0] = InputType(2.0f);
    InputType *dt = nullptr, *dx = nullptr;
    HIP_OK(hipMalloc(&dt, sizeof(InputType) * nt));
    HIP_OK(hipMalloc(&dx, sizeof(InputType) * n));
    HIP_OK(hipMemcpy(dt, ht.data( ), sizeof(InputType) * nt, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dx, hx.data( ), sizeof(InputType) * n, hipMemcpyHostToDevice));
    target.FwdFFT(dt); // spectrum of the target, in place
    if constexpr ( std::is_same_v<ImageType, __half2> && std::is_same_v<InputType, __half> ) {
        // a <float, __half, __half2> target writes its spectrum as __half2 already
    }
    const float s6 = 1.f / (float(b) * b * 6.f);
    FastFFT::KernelFunction::my_functor<float, 0, FastFFT::KernelFunction::NOOP>                noop;
    FastFFT::KernelFunction::my_functor<float, 2, FastFFT::KernelFunction::SCALE>               scale(s6);
    FastFFT::KernelFunction::my_functor<float, 4, FastFFT::KernelFunction::CONJ_MUL>            conj_mul;
    FastFFT::KernelFunction::my_functor<float, 4, FastFFT::KernelFunction::CONJ_MUL_THEN_SCALE> conj_mul_then_scale(s6);
    auto* image = reinterpret_cast<ImageType*>(dt);
    if ( mode == 0 )
        FT.FwdImageInvFFT(dx, image, dx, noop, conj_mul, noop);
    else if ( mode == 1 )
        FT.FwdImageInvFFT(dx, image, dx, scale, conj_mul, noop);
    else
        FT.FwdImageInvFFT(dx, image, dx, noop, conj_mul_then_scale, noop);
    std::vector<InputType> out(FT.ReturnInvOutputMemorySize( ));
    FT.CopyDeviceToHostAndSynchronize(out.data( ));
    HIP_OK(hipFree(dt));
    HIP_OK(hipFree(dx));
    const float want = mode == 0 ? float(b) * b * 6.f : 1.f;
    if ( float(out[0]) != want )
        return false;
    if ( mode == 0 ) {
        for ( int y = 0; y < b; y++ )
            for ( int x = 0; x < b; x++ )
                if ( (x || y) && float(out[y * (b + 2) + x]) != 0.f )
                    return false;
    }
    return true;
}

int main(int argc, char** argv) {
    std::string which = argc > 1 ? argv[1] : "all";
    auto        want  = [&](const char* n) { return which == "all" || which == n; };
    if ( want("round_trip") )
        report("round_trip_64", test_round_trip( ));
    if ( want("unit_impulse") ) {
        report("unit_impulse_in_place", test_unit_impulse(true));
        report("unit_impulse_out_of_place", test_unit_impulse(false));
    }
    if ( want("convolution") ) {
        for ( auto sb : std::vector<std::pair<int, int>>{{64, 64}, {64, 512}, {256, 4096}, {384, 4096}} )
            report("convolution_" + std::to_string(sb.first) + "_" + std::to_string(sb.second),
                   convolution_case<float, float2>(sb.first, sb.second, 0));
    }
    if ( want("scale") ) {
        report("scale_pre_op", convolution_case<float, float2>(64, 512, 1));
        report("scale_conj_mul_then_scale", convolution_case<float, float2>(64, 512, 2));
    }
    if ( want("half_io") ) {
        // preop_functor_scale_test.cu runs the same checks with __half I/O and a __half2 image (sizes <= 128)
        report("half_io_convolution", convolution_case<__half, __half2>(16, 64, 0));
        report("half_io_scale_pre_op", convolution_case<__half, __half2>(16, 64, 1));
        report("half_io_conj_mul_then_scale", convolution_case<__half, __half2>(16, 64, 2));
    }
    std::printf("%s: %d failure(s)\n", n_fail ? "FAILED" : "OK", n_fail);
    return n_fail ? 1 : 0;
}
