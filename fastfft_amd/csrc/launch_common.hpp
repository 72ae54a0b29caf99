// Shared helpers of the launch_*.hip translation units.
#pragma once

#include <cstdlib>
#include <type_traits>
#include <typeinfo>
#include <utility>

#include "launch.hpp"
#include "line_kernels.hpp"

namespace fastfft_amd {

static_assert(kTwiddleTableN == kTwiddleN + 2 * kTwiddleN, "host and device twiddle table lengths differ");

inline LineMap to_device(const HostLineMap& m) {
    LineMap d;
    d.s1 = m.s1;
    d.s2 = m.s2;
    d.es = m.es;
    d.n2 = m.n2;
    return d;
}

// Calls f(std::integral_constant<int, P>) for the runtime sub-transform size P.
template <class F>
hipError_t dispatch_P(int P, F&& f) {
    switch ( P ) {
        case 2: return f(std::integral_constant<int, 2>{ });
        case 4: return f(std::integral_constant<int, 4>{ });
        case 8: return f(std::integral_constant<int, 8>{ });
        case 16: return f(std::integral_constant<int, 16>{ });
        case 32: return f(std::integral_constant<int, 32>{ });
        case 64: return f(std::integral_constant<int, 64>{ });
        case 128: return f(std::integral_constant<int, 128>{ });
        case 256: return f(std::integral_constant<int, 256>{ });
        case 512: return f(std::integral_constant<int, 512>{ });
        case 1024: return f(std::integral_constant<int, 1024>{ });
        case 2048: return f(std::integral_constant<int, 2048>{ });
        case 4096: return f(std::integral_constant<int, 4096>{ });
        case 8192: return f(std::integral_constant<int, 8192>{ });
        default: return hipErrorInvalidValue;
    }
}

// Grid and LDS of a line stage, with the dynamic-LDS attribute raised above 64 KiB.
template <int P, class K>
hipError_t prepare_launch(K kernel, int nlines, dim3& grid, dim3& block, size_t& lds) {
    using S = LineShape<P>;
    lds     = sizeof(float2) * (size_t)S::LDS;
    grid    = dim3((unsigned)((nlines + S::G - 1) / S::G));
    block   = dim3(S::BLOCK);
    if ( lds > 65536 ) {
        hipError_t e = hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if ( e != hipSuccess )
            return e;
    }
    return hipSuccess;
}

inline int twiddle_stride(int N) { return kTwiddleN / N; }

// A type whose mangled name spells out the kernel K with its template arguments.
template <auto K>
struct KernelTag { };

// K<<<grid, block, lds, st>>>(a...), named on stderr first when FASTFFT_LOG is set.
template <auto K, class... A>
inline void launch_kernel(dim3 grid, dim3 block, size_t lds, hipStream_t st, A&&... a) {
    if ( log_enabled( ) )
        log_launch(typeid(KernelTag<K>).name( ), grid, block, lds);
    K<<<grid, block, lds, st>>>(std::forward<A>(a)...);
}

// Integer value of a tuning environment variable (dflt when unset).  Read on
// every call so that tests can switch kernels within one process; getenv is
// cheap next to a kernel launch.
inline int env_int(const char* name, int dflt) {
    const char* v = std::getenv(name);
    return (v && v[0]) ? std::atoi(v) : dflt;
}

} // namespace fastfft_amd
