// Twiddle-table generation (the only FP64 step) and the argmax epilogue.
#include "launch_common.hpp"

namespace fastfft_amd {

// tw[k] = W_8192^k for k < 8192; tw[8192 + n + m] = W_n^m for powers of two n
// and m < n (entry 8192 + 0 is unused and holds 1).
__global__ void k_twiddle_init(float2* __restrict__ tw) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if ( k >= kTwiddleTableN )
        return;
    long m = k, n = kTwiddleN;
    if ( k >= kTwiddleN ) {
        const int idx = k - kTwiddleN;
        n             = 1;
        while ( 2 * n <= idx )
            n *= 2;
        m = idx - n;
        if ( idx == 0 ) {
            m = 0;
            n = 1;
        }
    }
    double s, c;
    // exp(-2 pi i m / n): the angle in units of pi is 2m/n, exact in binary
    sincospi(2.0 * (double)m / (double)n, &s, &c);
    tw[k] = make_float2((float)c, (float)(-s));
}

// The table itself: a static device array (one copy per GPU, kept until the process
// ends), so that no call allocates it.  Its device address comes from
// hipGetSymbolAddress, called with the device current.
static __device__ __align__(256) float2 twiddle_table[kTwiddleTableN];

hipError_t twiddle_table_address(float2** out) {
    void*            p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(twiddle_table));
    *out               = e == hipSuccess ? static_cast<float2*>(p) : nullptr;
    return e;
}

hipError_t launch_twiddle_init(float2* tw, hipStream_t stream) {
    launch_kernel<k_twiddle_init>(dim3((kTwiddleTableN + 255) / 256), dim3(256), 0, stream, tw);
    return hipGetLastError( );
}

// ---------------------------------------------------------------- argmax
// (value, linear index) pairs ordered by value, then by smaller index, so the
// first maximum in row-major order wins, as a sequential scan would.
struct Best {
    float v;
    int   i;
};

__device__ __forceinline__ Best better(Best a, Best b) {
    if ( b.v > a.v || (b.v == a.v && b.i < a.i) )
        return b;
    return a;
}

constexpr int kArgmaxThreads = 256;
constexpr int kArgmaxBlocks  = kArgmaxPartials;

__device__ __forceinline__ Best block_reduce(Best b) {
    __shared__ float sv[kArgmaxThreads];
    __shared__ int   si[kArgmaxThreads];
    sv[threadIdx.x] = b.v;
    si[threadIdx.x] = b.i;
    __syncthreads( );
    for ( int s = kArgmaxThreads / 2; s > 0; s >>= 1 ) {
        if ( threadIdx.x < s ) {
            Best o   = {sv[threadIdx.x + s], si[threadIdx.x + s]};
            Best m   = {sv[threadIdx.x], si[threadIdx.x]};
            Best r   = better(m, o);
            sv[threadIdx.x] = r.v;
            si[threadIdx.x] = r.i;
        }
        __syncthreads( );
    }
    Best r = {sv[0], si[0]};
    return r;
}

template <typename T>
__global__ void __launch_bounds__(kArgmaxThreads)
        k_argmax_partial(const T* __restrict__ img, int nx, int ny, int pitch, float* __restrict__ pv, int* __restrict__ pi) {
    const long n  = (long)nx * ny;
    Best       b  = {-INFINITY, 0x7fffffff};
    for ( long i = (long)blockIdx.x * kArgmaxThreads + threadIdx.x; i < n; i += (long)gridDim.x * kArgmaxThreads ) {
        const int y = (int)(i / nx), x = (int)(i % nx);
        Best      c = {load_real(img, (long)y * pitch + x), (int)i};
        b           = better(b, c);
    }
    b = block_reduce(b);
    if ( threadIdx.x == 0 ) {
        pv[blockIdx.x] = b.v;
        pi[blockIdx.x] = b.i;
    }
}

__global__ void __launch_bounds__(kArgmaxThreads)
        k_argmax_final(const float* __restrict__ pv, const int* __restrict__ pi, int nparts, int nx, float* __restrict__ rv,
                       int* __restrict__ rxy) {
    Best b = {-INFINITY, 0x7fffffff};
    for ( int i = threadIdx.x; i < nparts; i += kArgmaxThreads ) {
        Best c = {pv[i], pi[i]};
        b      = better(b, c);
    }
    b = block_reduce(b);
    if ( threadIdx.x == 0 ) {
        rv[0]  = b.v;
        rxy[0] = b.i % nx;
        rxy[1] = b.i / nx;
    }
}

// First maximum of n keys (value order, then smaller index).
__global__ void __launch_bounds__(kArgmaxThreads)
        k_argmax_keys(const unsigned long long* __restrict__ keys, int n, int nx, float* __restrict__ rv, int* __restrict__ rxy,
                      long key_stride) {
    __shared__ unsigned long long s_best;
    keys += blockIdx.x * key_stride;
    rv += blockIdx.x;
    rxy += 2 * blockIdx.x;
    unsigned long long            best = 0;
#pragma unroll 4
    for ( int i = threadIdx.x; i < n; i += kArgmaxThreads ) { // the loads of four iterations in flight together
        const unsigned long long k = keys[i];
        best                       = k > best ? k : best;
    }
#pragma unroll
    for ( int off = 32; off > 0; off >>= 1 ) {
        const unsigned long long other = __shfl_xor(best, off);
        best                           = other > best ? other : best;
    }
    if ( threadIdx.x == 0 )
        s_best = 0;
    __syncthreads( );
    if ( (threadIdx.x & 63) == 0 )
        atomicMax(&s_best, best);
    __syncthreads( );
    if ( threadIdx.x == 0 ) {
        const unsigned idx = argmax_key_index(s_best);
        rv[0]              = argmax_key_value(s_best);
        rxy[0]             = (int)(idx % (unsigned)nx);
        rxy[1]             = (int)(idx / (unsigned)nx);
    }
}

hipError_t launch_argmax_keys(const unsigned long long* keys, int n, int nx, float* result_value, int* result_xy,
                              hipStream_t stream, int batch, long key_stride) {
    launch_kernel<k_argmax_keys>(batch, kArgmaxThreads, 0, stream, keys, n, nx, result_value, result_xy, key_stride);
    return hipGetLastError( );
}

// partial_v / partial_i: kArgmaxBlocks entries of device scratch; the result goes to
// result_value[0] and result_xy[0..1] (device memory, may be caller buffers).
hipError_t launch_argmax(const void* img, RealT rt, int nx, int ny, int pitch, float* partial_v, int* partial_i,
                         float* result_value, int* result_xy, hipStream_t stream) {
    const long n      = (long)nx * ny;
    int        blocks = (int)std::min<long>(kArgmaxBlocks, (n + kArgmaxThreads * 4 - 1) / (kArgmaxThreads * 4));
    blocks            = blocks < 1 ? 1 : blocks;
    float* pv         = partial_v;
    int*   pi         = partial_i;
    if ( rt == RealT::f32 )
        launch_kernel<k_argmax_partial<float>>(blocks, kArgmaxThreads, 0, stream, static_cast<const float*>(img), nx, ny, pitch, pv, pi);
    else
        launch_kernel<k_argmax_partial<__half>>(blocks, kArgmaxThreads, 0, stream, static_cast<const __half*>(img), nx, ny, pitch, pv, pi);
    hipError_t e = hipGetLastError( );
    if ( e != hipSuccess )
        return e;
    launch_kernel<k_argmax_final>(1, kArgmaxThreads, 0, stream, pv, pi, blocks, nx, result_value, result_xy);
    return hipGetLastError( );
}

} // namespace fastfft_amd
