// Fused correlation pass for N = 1024 on the matrix cores (split-BF16 DFT-32
// stages, mfma_dft.hpp): the second axis of FwdImageInvFFT for the 1024-point
// frames of the batched correlation, the reduced-precision counterpart of
// k_xcorr_col (the reference's block_fft_kernel_C2C_FWD_INCREASE_OP_INV_NONE
// with a trimmed store, src/fastfft/FastFFT.cu:20-56).
//
// Workgroup (column c, group g) has four wavefronts, each correlating column c
// of one template at a time:
//   forward  X = DFT_1024(a)      a[y] = A[y][c], y < L <= 512 (zeros above)
//   product  Z[k] = G[k] conj(X[k]) s           (G: image column c, staged in LDS)
//   inverse  b[j] = IDFT_1024(Z)[j], j < M      (window origin 0)
// Forward (n = 32 n1 + n2, k = k1 + 32 k2):
//   stage 1: Y1[k1][n2] = sum_{n1 < 16} F[k1][n1] a[32 n1 + n2]    (one K-chunk: L <= 512)
//   twiddle: Y1[k1][n2] *= W_1024^{n2 k1}; transpose in LDS
//   stage 2: X[k1 + 32 k2] = sum_{n2} F[k2][n2] Y1[k1][n2]
// The forward result sits in the D layout (lane: k1, register: k2), which is
// the B layout of the inverse's first stage once the K order of its operand
// matrix is permuted the same way (dft32_operand<true>), so:
//   stage 1: V[ja][k1] = sum_{k2} conj F[ja][k2] Z[k1 + 32 k2]
//   twiddle: V[ja][k1] *= W_1024^{-ja k1}; transpose in LDS
//   stage 2: b[ja + 32 jb] = sum_{k1} conj F[jb][k1] V[ja][k1]
#pragma once

#include "items.hpp"
#include "mfma_dft.hpp"
#include "rows_wave.hpp"

namespace fastfft_amd {

constexpr int kMfmaWaves = 4; // wavefronts per workgroup

// Grid: ncols x groups, 256 threads; group g takes items [g per, (g + 1) per)
// with per = ceil(nitems / groups), wave w of the group items w, w + 4, ....
// a[y * a_es] is row y of a template's column c (the item's pointer plus c),
// out[j * o_es] row j of its output column; img[c * ldg + k] is element k of
// image column c.  The DFT-32 operands (natural and permuted K order, two
// K-chunks each) and the per-lane twiddles W_1024^{(lane % 32) mfma_row(r)}
// (the same for both transforms: conjugated for the inverse) are built once per
// workgroup in LDS.
__global__ void __launch_bounds__(256)
        k_xcorr_mfma1024(XcorrItem one, const XcorrItem* __restrict__ items, int nitems, long a_es,
                         const float2* __restrict__ img, long ldg, long o_es, int M, float intra_scale,
                         const float2* __restrict__ tw) {
    constexpr int N = 1024, TP = 33; // transpose pitch (conflict-free reads of a column)
    __shared__ float2   s_img[N];
    __shared__ float2   s_tr[kMfmaWaves][32 * TP];
    __shared__ bf16x8   s_op[4][6][64]; // [natural c0, c1, permuted c0, c1][rh, rl, ih, il, -ih, -il][lane]
    __shared__ float2   s_tw[16][64];
    const int     lane = threadIdx.x & 63;
    const int     w    = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); // wave-uniform: scalar item loads
    const int     h    = lane >> 5;
    const int     n    = lane & 31;
    const long    col  = xcd_row(blockIdx.x, gridDim.x);
    const float2* twm  = multi_twiddles(tw);
    // ---- workgroup set-up: image column, operands, twiddles
    const auto* gc = as_global(img) + col * ldg;
    {
        float2 g[N / 256]; // all loads first, then the LDS writes
#pragma unroll
        for ( int i = 0; i < N / 256; i++ )
            g[i] = load_cpx(gc, threadIdx.x + 256 * i);
#pragma unroll
        for ( int i = 0; i < N / 256; i++ )
            s_img[threadIdx.x + 256 * i] = g[i];
    }
    {
        const CpxSplit f = w < 2 ? dft32_operand<false>(twm, lane, w) : dft32_operand<true>(twm, lane, w - 2);
        s_op[w][0][lane] = f.rh;
        s_op[w][1][lane] = f.rl;
        s_op[w][2][lane] = f.ih;
        s_op[w][3][lane] = f.il;
        s_op[w][4][lane] = bf16_neg(f.ih);
        s_op[w][5][lane] = bf16_neg(f.il);
    }
    {
        float2 t4[16 / kMfmaWaves];
#pragma unroll
        for ( int i = 0; i < 16 / kMfmaWaves; i++ )
            t4[i] = load_cpx(as_global(twm), N + ((n * mfma_row(w + kMfmaWaves * i, h)) & (N - 1)));
#pragma unroll
        for ( int i = 0; i < 16 / kMfmaWaves; i++ )
            s_tw[w + kMfmaWaves * i][lane] = t4[i];
    }
    __syncthreads( );
    // operand o of this lane (lane-consecutive 16-byte LDS reads)
    auto op = [&](int o) {
        return DftOperand{s_op[o][0][lane], s_op[o][1][lane], s_op[o][2][lane],
                          s_op[o][3][lane], s_op[o][4][lane], s_op[o][5][lane]};
    };
    const int total = items ? nitems : 1;
    const int per   = (total + (int)gridDim.y - 1) / (int)gridDim.y;
    const int first = (int)blockIdx.y * per, last = min(first + per, total);
    float2*   tr    = s_tr[w];
    // the template column of item ii: a[32 n1 + n2] at lane n2, element n1 - 8 h
    // (issued one item ahead, so that the strided loads overlap the transforms)
    auto gather = [&](int ii, float2 (&z)[8]) {
        const XcorrItem it = items ? load_item(items, ii) : one;
        const auto*     a  = as_global(static_cast<const float2*>(it.a)) + col;
#pragma unroll
        for ( int e = 0; e < 8; e++ ) {
            const int y = 32 * (8 * h + e) + n;
            z[e]        = y < it.L ? load_cpx(a, (long)y * a_es) : make_float2(0.f, 0.f);
        }
    };
    float2 znext[8];
    if ( first + w < last )
        gather(first + w, znext);
    for ( int ii = first + w; ii < last; ii += kMfmaWaves ) {
        auto* out = as_global(static_cast<float2*>((items ? load_item(items, ii) : one).out)) + col;
        // ---- forward stage 1: B[n1][n2] = a[32 n1 + n2], n1 = 8 h + e
        f32x16 cr = { }, ci = { };
        {
            float2 z[8];
#pragma unroll
            for ( int e = 0; e < 8; e++ )
                z[e] = znext[e];
            if ( ii + kMfmaWaves < last )
                gather(ii + kMfmaWaves, znext);
            cmfma<-1>(op(0), split8(z), cr, ci);
        }
        // twiddle W_1024^{n2 k1} (lane: n2, register: k1), then Y1[k1][n2] into LDS
#pragma unroll
        for ( int r = 0; r < 16; r++ )
            tr[mfma_row(r, h) * TP + n] = cmul(make_float2(cr[r], ci[r]), s_tw[r][lane]);
        wave_lds_fence( );
        // ---- forward stage 2: B[n2][k1] = Y1[k1][n2] (lane: k1)
        f32x16 xr = { }, xi = { };
#pragma unroll
        for ( int c = 0; c < 2; c++ ) {
            float2 z[8];
#pragma unroll
            for ( int e = 0; e < 8; e++ )
                z[e] = tr[n * TP + 16 * c + 8 * h + e];
            cmfma<-1>(op(c), split8(z), xr, xi);
        }
        wave_lds_fence( );
        // ---- product with the image column, then inverse stage 1 (K = k2 in register order)
        f32x16 vr = { }, vi = { };
#pragma unroll
        for ( int c = 0; c < 2; c++ ) {
            float2 z[8];
#pragma unroll
            for ( int e = 0; e < 8; e++ ) {
                const int    r = 8 * c + e;
                const float2 g = s_img[n + 32 * mfma_row(r, h)];
                const float2 p = cmulc(g, make_float2(xr[r], xi[r]));
                z[e]           = make_float2(p.x * intra_scale, p.y * intra_scale);
            }
            cmfma<+1>(op(2 + c), split8(z), vr, vi);
        }
        // twiddle W_1024^{-ja k1} (lane: k1, register: ja), then V[ja][k1] into LDS
#pragma unroll
        for ( int r = 0; r < 16; r++ )
            tr[mfma_row(r, h) * TP + n] = cmulc(make_float2(vr[r], vi[r]), s_tw[r][lane]);
        wave_lds_fence( );
        // ---- inverse stage 2: B[k1][ja] = V[ja][k1] (lane: ja)
        f32x16 br = { }, bi = { };
#pragma unroll
        for ( int c = 0; c < 2; c++ ) {
            float2 z[8];
#pragma unroll
            for ( int e = 0; e < 8; e++ )
                z[e] = tr[n * TP + 16 * c + 8 * h + e];
            cmfma<+1>(op(c), split8(z), br, bi);
        }
        wave_lds_fence( );
        // ---- b[ja + 32 jb] (lane: ja, register: jb), rows j < M of the window
#pragma unroll
        for ( int r = 0; r < 16; r++ ) {
            const int j = n + 32 * mfma_row(r, h);
            if ( j < M )
                store_cpx(out, (long)j * o_es, make_float2(br[r], bi[r]));
        }
    }
}

} // namespace fastfft_amd
