// Device scratch of the allocation-free correlation entry point
// (fastfft_fwd_image_inv_argmax_noalloc): the correlation's two intermediates and
// the argmax slots live in one static device array per GPU instead of the plan's
// workspace, so that a call allocates nothing.
#include <mutex>

#include "launch.hpp"

namespace fastfft_amd {

// One copy per GPU the library runs on, held until the process ends.  In host code
// `scratch` names a host-side stand-in: kernels and copies take the device address
// kept in g_scratch_ptr, never the array's name.
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
constexpr int kScratchDevices = 64;
std::mutex    g_scratch_mutex;
void*         g_scratch_ptr[kScratchDevices] = {nullptr};
} // namespace

hipError_t static_scratch_for_current_device(void** out) {
    *out           = nullptr;
    int        dev = 0;
    hipError_t e   = hipGetDevice(&dev);
    if ( e != hipSuccess )
        return e;
    if ( dev < 0 || dev >= kScratchDevices )
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if ( ! g_scratch_ptr[dev] ) {
        void* ptr = nullptr;
        e         = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if ( e != hipSuccess )
            return e;
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

} // namespace fastfft_amd
