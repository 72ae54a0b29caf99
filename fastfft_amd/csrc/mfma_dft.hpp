// Split-BF16 DFT-32 stages on the gfx950 matrix cores.
//
// A 1024-point DFT is computed as two 32-point stages (N = 32 x 32, four-step):
//   Y1[k1][n2]         = sum_{n1} W_32^{n1 k1} x[32 n1 + n2]
//   X[k1 + 32 k2]      = sum_{n2} W_32^{n2 k2} ( W_1024^{n2 k1} Y1[k1][n2] )
// Each stage is a 32 x 32 x 32 complex matrix product on
// v_mfma_f32_32x32x16_bf16 (D = A B + C; A is 32 x 16, B is 16 x 32, two
// K-chunks per stage).  Lane layouts (checked on the device):
//   A: lane l holds A[l % 32][8 (l / 32) + e], e < 8
//   B: lane l holds B[8 (l / 32) + e][l % 32], e < 8
//   D: lane l, register r holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
// Every FP32 operand is split into hi = bf16(x) and lo = bf16(x - hi), and a
// real product is taken as hi*hi + hi*lo + lo*hi with FP32 accumulation
// (relative error about 2^-17 per operand instead of 2^-9); twiddles between
// the stages are applied in FP32.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_core.hpp"

namespace fastfft_amd {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) unsigned u32x4;

// hi / lo parts of the real and imaginary components of one operand chunk
struct CpxSplit {
    bf16x8 rh, rl, ih, il;
};

// A DFT-32 operand chunk with its imaginary part also negated, so that the
// complex products need no sign changes of the data operand
struct DftOperand {
    bf16x8 rh, rl, ih, il, nih, nil;
};

// D row held by register r of a lane in half h = lane / 32
__device__ __forceinline__ constexpr int mfma_row(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }

__device__ __forceinline__ bf16x8 bf16_neg(bf16x8 x) {
    u32x4 u = __builtin_bit_cast(u32x4, x);
    u ^= 0x80008000u;
    return __builtin_bit_cast(bf16x8, u);
}

// split of 8 complex values (one lane's share of a K-chunk)
__device__ __forceinline__ CpxSplit split8(const float2 (&z)[8]) {
    CpxSplit s;
#pragma unroll
    for ( int e = 0; e < 8; e++ ) {
        const __bf16 hr = (__bf16)z[e].x;
        const __bf16 hi = (__bf16)z[e].y;
        s.rh[e]         = hr;
        s.ih[e]         = hi;
        s.rl[e]         = (__bf16)(z[e].x - (float)hr);
        s.il[e]         = (__bf16)(z[e].y - (float)hi);
    }
    return s;
}

__device__ __forceinline__ f32x16 mfma_bf16(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// (cr, ci) += F X over one K-chunk, F the forward DFT-32 matrix (SIGN = -1) or
// its conjugate (SIGN = +1); F and X split into hi / lo parts.
template <int SIGN>
__device__ __forceinline__ void cmfma(const DftOperand& F, const CpxSplit& X, f32x16& cr, f32x16& ci) {
    // forward: cr += Fr Xr + (-Fi) Xi, ci += Fr Xi + Fi Xr
    // inverse: cr += Fr Xr + Fi Xi,    ci += Fr Xi + (-Fi) Xr
    const bf16x8 ah = SIGN < 0 ? F.nih : F.ih, al = SIGN < 0 ? F.nil : F.il; // with Xi into cr
    const bf16x8 bh = SIGN < 0 ? F.ih : F.nih, bl = SIGN < 0 ? F.il : F.nil; // with Xr into ci
    cr = mfma_bf16(F.rh, X.rh, cr);
    ci = mfma_bf16(F.rh, X.ih, ci);
    cr = mfma_bf16(F.rh, X.rl, cr);
    ci = mfma_bf16(F.rh, X.il, ci);
    cr = mfma_bf16(F.rl, X.rh, cr);
    ci = mfma_bf16(F.rl, X.ih, ci);
    cr = mfma_bf16(ah, X.ih, cr);
    ci = mfma_bf16(bh, X.rh, ci);
    cr = mfma_bf16(ah, X.il, cr);
    ci = mfma_bf16(bh, X.rl, ci);
    cr = mfma_bf16(al, X.ih, cr);
    ci = mfma_bf16(bl, X.rh, ci);
}

// DFT-32 A operands of one lane for K-chunk c: F[m][k] = W_32^{m k}, m = lane % 32,
// k = 16 c + 8 h + e (natural order) or, PERM, k = mfma_row(8 c + e, h) (the
// order in which a D layout holds its rows, so that a stage can consume the
// previous stage's result without a transpose).  twm is the multi-resolution
// twiddle table (twm[32 + j] = W_32^j).
template <bool PERM>
__device__ __forceinline__ CpxSplit dft32_operand(const float2* twm, int lane, int c) {
    const int m = lane & 31, h = lane >> 5;
    float2    z[8];
#pragma unroll
    for ( int e = 0; e < 8; e++ ) {
        const int k = PERM ? mfma_row(8 * c + e, h) : 16 * c + 8 * h + e;
        z[e]        = twm[32 + ((m * k) & 31)];
    }
    return split8(z);
}

} // namespace fastfft_amd
