// Split-BF16 correlation pass on the matrix cores, as a Toeplitz matrix
// product per column: the reduced-precision counterpart of k_xcorr_col for the
// batched correlation's frame (the reference's
// block_fft_kernel_C2C_FWD_INCREASE_OP_INV_NONE with a trimmed store,
// src/fastfft/FastFFT.cu:20-56, computed without the y transforms).
//
// Per column kx of the frame, with S the frame image in mixed form (x
// frequency kx, y position: the frame's R2C x-pass output, unnormalised) and
// a_t the column of template t's R2C x-pass output, the correlation's
// second-axis pass is
//   b_t[j] = sum_{v < L} S[j + v] conj(a_t[v]),          j < M,
// which equals IFFT_y(FFT_y(S) conj(FFT_y(a_t)))[j] / Fy for j + v < Fy.  Over
// the K = 2L interleaved values x_t = (Re a_t[0], Im a_t[0], Re a_t[1], ...):
//   Re b_t[j] = sum_k R[2j + k] x_t[k],   R = (Sr[0], Si[0], Sr[1], Si[1], ...)
//   Im b_t[j] = sum_k I[2j + k] x_t[k],   I = (Si[0], -Sr[0], Si[1], -Sr[1], ...)
// two real matrix products with Toeplitz-structured left operands (rows j,
// step 2) and the templates of a group as the 32 columns of the right one.
// Each is taken on v_mfma_f32_32x32x16_bf16 as hi*hi + hi*lo + lo*hi of
// split operands (hi = bf16(x), lo = bf16(x - hi)), accumulated in FP32.
//
// The DC and Nyquist columns (real template values; the C2R pass keeps only
// the real parts of their outputs) share one workgroup: R = (Re S_0[y], 0),
// I = (0, Re S_last[y]) and x = (a_0[v], a_last[v]) give Re b_0 and Re b_last
// as the two products.
//
// Workgroup (column pair index, template group of 32), four waves; wave w owns
// output rows [128 w, 128 w + 128) as four 32-row tiles, for all 32 templates.
// The left operands are staged in LDS once per workgroup as bf16 hi / lo
// arrays in two copies shifted by 0 and 2 elements, so that every lane's
// operand (rows step the Toeplitz diagonal by 2 elements) is two 8-byte
// aligned reads; the copies start 64 bytes apart modulo 128, so the reads of
// any sixteen lanes fall on distinct banks.  The template values are staged by
// the workgroup in chunks of 64 rows (each thread its template's rows, split
// into hi / lo on the way into LDS), double buffered: the loads of the next
// chunk are in flight while the waves run the current one, and all four waves
// read the same staged values.
//
// MFMA lane layouts (v_mfma_f32_32x32x16_bf16, checked on the device):
//   A: lane l holds A[l % 32][8 (l / 32) + e], e < 8
//   B: lane l holds B[8 (l / 32) + e][l % 32], e < 8
//   D: lane l, register r holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
#pragma once

#include "items.hpp"
#include "rows_wave.hpp"

namespace fastfft_amd {

typedef __attribute__((ext_vector_type(8))) __bf16 tz_bf16x8;
typedef __attribute__((ext_vector_type(16))) float tz_f32x16;
typedef __attribute__((ext_vector_type(4))) unsigned tz_u32x4;
typedef __attribute__((ext_vector_type(2))) unsigned tz_u32x2;

constexpr int kTzWaves     = 4;    // waves per workgroup
constexpr int kTzRowsWave  = 128;  // output rows per wave (four 32-row tiles)
constexpr int kTzMaxM      = kTzWaves * kTzRowsWave;
constexpr int kTzTemplates = 32;   // templates per workgroup (the MFMA's N)
constexpr int kTzFy        = 1024; // frame length along y
constexpr int kTzLen       = 2 * kTzFy;   // interleaved operand length
constexpr int kTzCopy      = kTzLen + 32; // elements per shifted copy (copies 64 bytes apart modulo 128)
constexpr int kTzArrays    = 4;           // R hi, R lo, I hi, I lo
constexpr int kTzChunk     = 8;           // K-steps (8 template rows each) per staged chunk of templates

__device__ __forceinline__ tz_f32x16 tz_mfma(tz_bf16x8 a, tz_bf16x8 b, tz_f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// hi = bf16(x) (round to nearest even), lo = bf16(x - hi) of two values, packed
// as two bf16 each (element 0 in the low half)
struct TzPair {
    unsigned hi, lo;
};
__device__ __forceinline__ TzPair tz_split2(float x0, float x1) {
    const __bf16 h0 = (__bf16)x0, h1 = (__bf16)x1;
    const __bf16 l0 = (__bf16)(x0 - (float)h0), l1 = (__bf16)(x1 - (float)h1);
    TzPair       p;
    p.hi = (unsigned)__builtin_bit_cast(unsigned short, h0) | ((unsigned)__builtin_bit_cast(unsigned short, h1) << 16);
    p.lo = (unsigned)__builtin_bit_cast(unsigned short, l0) | ((unsigned)__builtin_bit_cast(unsigned short, l1) << 16);
    return p;
}

// 16 bytes of LDS at an 8-byte aligned element offset, as two 8-byte reads
__device__ __forceinline__ tz_bf16x8 tz_read8(const unsigned short* p) {
    const tz_u32x2 a = *reinterpret_cast<const tz_u32x2*>(p);
    const tz_u32x2 b = *reinterpret_cast<const tz_u32x2*>(p + 4);
    const tz_u32x4 v = {a[0], a[1], b[0], b[1]};
    return __builtin_bit_cast(tz_bf16x8, v);
}

// Body of k_xcorr_toeplitz for ordinary columns (ENDS = false) or the DC / Nyquist pair.
// The workgroup's LDS:
//   op  left operands: R and I, hi and lo, copies shifted by 0 and 2 elements
//   tp  right operands of a chunk of K-steps, double buffered: [buffer][K-step][lane half
//       g][template n] 8 bf16 (rows 8 ks + 4 g .. + 3, re and im interleaved), hi then lo
struct TzShared {
    __attribute__((aligned(16))) unsigned short op[kTzArrays][2][kTzCopy];
    tz_u32x4                                    tp[2][kTzChunk][2][kTzTemplates][2];
};

template <bool ENDS>
__device__ __forceinline__ void tz_body(TzShared& sh, const XcorrItem* __restrict__ items, int nitems, long a_es,
                                        const float2* __restrict__ S, long s_stride, long o_es, int M, int Lmax,
                                        float scale, int last, int col) {
    auto&     s_op = sh.op;
    auto&     s_tp = sh.tp;
    const int tid  = threadIdx.x;
    const int lane = tid & 63;
    const int w    = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g    = lane >> 5; // lane half: K rows 8 g .. 8 g + 7 of a K-step
    const int n    = lane & 31; // MFMA row (A) / column (B)
    const int nks  = (Lmax + 7) / 8;
    const int nchk = (nks + kTzChunk - 1) / kTzChunk;

    // ---- staging of the templates: thread tid takes template tn = tid % 32 and the two
    // (K-step, half) slots q = tid / 32 and q + 8 of a chunk, four rows each.  Every load
    // is issued (rows clamped to the template, threads past the batch on template 0) and
    // the values outside the template replaced by zeros, so no load waits for another.
    const int       tn = tid & 31;
    const int       tq = tid >> 5;
    const int       tt = (int)blockIdx.y * kTzTemplates + tn;
    const XcorrItem it = items[tt < nitems ? tt : 0];
    const int       Lt = tt < nitems ? it.L : 0; // rows taken
    const int       Lc = it.L > 0 ? it.L : 1;    // rows that may be read
    const auto*     ga = as_global(static_cast<const float2*>(it.a));
    auto load_chunk = [&](int ck, float2 (&z)[2][4]) {
#pragma unroll
        for ( int sl = 0; sl < 2; sl++ ) {
            const int q  = tq + 8 * sl; // 0 .. 15: K-step q / 2, half q % 2
            const int v0 = 8 * (ck * kTzChunk + (q >> 1)) + 4 * (q & 1);
#pragma unroll
            for ( int i = 0; i < 4; i++ ) {
                const long r = (long)min(v0 + i, Lc - 1) * a_es;
                if constexpr ( ENDS )
                    z[sl][i] = make_float2(load_cpx(ga, r).x, load_cpx(ga, r + last).x);
                else
                    z[sl][i] = load_cpx(ga, r + col);
            }
        }
    };
    auto store_chunk = [&](int buf, int ck, const float2 (&z)[2][4]) {
#pragma unroll
        for ( int sl = 0; sl < 2; sl++ ) {
            const int q  = tq + 8 * sl;
            const int v0 = 8 * (ck * kTzChunk + (q >> 1)) + 4 * (q & 1);
            tz_u32x4  h, l;
#pragma unroll
            for ( int i = 0; i < 4; i++ ) {
                const bool   in = v0 + i < Lt;
                const TzPair p  = tz_split2(in ? z[sl][i].x : 0.f, in ? z[sl][i].y : 0.f);
                h[i]            = p.hi;
                l[i]            = p.lo;
            }
            s_tp[buf][q >> 1][q & 1][tn][0] = h;
            s_tp[buf][q >> 1][q & 1][tn][1] = l;
        }
    };
    float2 zc[2][4];
    load_chunk(0, zc);

    // ---- left operands
    {
        const auto* s0 = as_global(S) + (long)col * s_stride;
        const auto* s1 = as_global(S) + (long)last * s_stride;
#pragma unroll
        for ( int yy = 0; yy < kTzFy / (kTzWaves * 64); yy++ ) {
            const int y = tid + kTzWaves * 64 * yy;
            float     r0, r1, i0, i1; // R[2y], R[2y + 1], I[2y], I[2y + 1]
            if constexpr ( ! ENDS ) {
                const float2 v = load_cpx(s0, y);
                r0 = v.x, r1 = v.y, i0 = v.y, i1 = -v.x;
            }
            else {
                r0 = load_cpx(s0, y).x, r1 = 0.f, i0 = 0.f, i1 = load_cpx(s1, y).x;
            }
            const TzPair pr = tz_split2(r0, r1), pi = tz_split2(i0, i1);
#pragma unroll
            for ( int c = 0; c < 2; c++ ) {
                const int at = 2 * y - 2 * c; // copy c holds element at + 2 c at index at
                if ( at >= 0 ) {
                    *reinterpret_cast<unsigned*>(&s_op[0][c][at]) = pr.hi;
                    *reinterpret_cast<unsigned*>(&s_op[1][c][at]) = pr.lo;
                    *reinterpret_cast<unsigned*>(&s_op[2][c][at]) = pi.hi;
                    *reinterpret_cast<unsigned*>(&s_op[3][c][at]) = pi.lo;
                }
            }
        }
    }
    store_chunk(0, 0, zc);
    __syncthreads( );

    // rows of this wave: all four tiles are computed (rows < 512 read valid operands),
    // tiles at or past M are not stored
    const int row0 = w * kTzRowsWave;
    tz_f32x16 acc_r[4], acc_i[4];
#pragma unroll
    for ( int a = 0; a < 4; a++ ) {
        acc_r[a] = tz_f32x16{ };
        acc_i[a] = tz_f32x16{ };
    }
    // LDS element of this lane's left operand: lane row j = row0 + 32 a + n, operand offset
    // 2 j + 16 ks + 8 g, in copy c = j % 2 (= n % 2) at index offset - 2 c (a multiple of 4)
    const int cpy  = n & 1;
    const int base = 2 * (row0 + n) + 8 * g - 2 * cpy;
    for ( int ck = 0; ck < nchk; ck++ ) {
        const int buf = ck & 1;
        if ( ck + 1 < nchk )
            load_chunk(ck + 1, zc);
        const int kend = min(kTzChunk, nks - ck * kTzChunk);
        for ( int kk = 0; kk < kend; kk++ ) {
            const int       ks = ck * kTzChunk + kk;
            const tz_bf16x8 bh = __builtin_bit_cast(tz_bf16x8, s_tp[buf][kk][g][n][0]);
            const tz_bf16x8 bl = __builtin_bit_cast(tz_bf16x8, s_tp[buf][kk][g][n][1]);
            // left operands of the four tiles, then the products in three rounds over the
            // eight accumulators, so that dependent matrix instructions are eight apart
            tz_bf16x8 rh[4], rl[4], ih[4], il[4];
#pragma unroll
            for ( int a = 0; a < 4; a++ ) {
                const int at = base + 64 * a + 16 * ks;
                rh[a]        = tz_read8(&s_op[0][cpy][at]);
                rl[a]        = tz_read8(&s_op[1][cpy][at]);
                ih[a]        = tz_read8(&s_op[2][cpy][at]);
                il[a]        = tz_read8(&s_op[3][cpy][at]);
            }
#pragma unroll
            for ( int a = 0; a < 4; a++ ) {
                acc_r[a] = tz_mfma(rh[a], bh, acc_r[a]);
                acc_i[a] = tz_mfma(ih[a], bh, acc_i[a]);
            }
#pragma unroll
            for ( int a = 0; a < 4; a++ ) {
                acc_r[a] = tz_mfma(rh[a], bl, acc_r[a]);
                acc_i[a] = tz_mfma(ih[a], bl, acc_i[a]);
            }
#pragma unroll
            for ( int a = 0; a < 4; a++ ) {
                acc_r[a] = tz_mfma(rl[a], bh, acc_r[a]);
                acc_i[a] = tz_mfma(il[a], bh, acc_i[a]);
            }
        }
        if ( ck + 1 < nchk ) {
            // buffer buf ^ 1 was last read in chunk ck - 1, before the barrier that ended it
            store_chunk(buf ^ 1, ck + 1, zc);
            __syncthreads( );
        }
    }
    // ---- store: register r of tile a holds row row0 + 32 a + 8 (r / 4) + 4 g + r % 4,
    // template n of the group
    const int tout = (int)blockIdx.y * kTzTemplates + n;
    if ( tout >= nitems )
        return;
    auto* go = as_global(static_cast<float2*>(items[tout].out));
#pragma unroll
    for ( int a = 0; a < 4; a++ ) {
        if ( row0 + 32 * a < M ) {
#pragma unroll
            for ( int r = 0; r < 16; r++ ) {
                const int j = row0 + 32 * a + 8 * (r >> 2) + 4 * g + (r & 3);
                if ( j < M ) {
                    if constexpr ( ! ENDS ) {
                        store_cpx(go, (long)j * o_es + col, make_float2(acc_r[a][r] * scale, acc_i[a][r] * scale));
                    }
                    else {
                        store_cpx(go, (long)j * o_es, make_float2(acc_r[a][r] * scale, 0.f));
                        store_cpx(go, (long)j * o_es + last, make_float2(acc_i[a][r] * scale, 0.f));
                    }
                }
            }
        }
    }
}

// Grid: x = last workgroups when pack_ends (columns 1 .. last - 1, and workgroup 0
// for the pair 0 / last), else last + 1; y = ceil(nitems / 32) template groups;
// 256 threads.
//   S       frame image in mixed form, column c at S + c * s_stride (Fy = 1024 values)
//   items   the templates: a[v * a_es + c] element v of column c of template t
//           (1 <= L_t rows), out[j * o_es + c] element j of its output column
//   M       output rows (<= 512), Lmax the largest template's rows (L_t <= Lmax,
//           Lmax rounded up to 8 <= 512), last the Nyquist column index (ncols - 1)
__global__ void __launch_bounds__(kTzWaves * 64, 2)
        k_xcorr_toeplitz(const XcorrItem* __restrict__ items, int nitems, long a_es, const float2* __restrict__ S,
                         long s_stride, long o_es, int M, int Lmax, float scale, int last, int pack_ends) {
    __shared__ TzShared sh;
    // each XCD takes a contiguous range of columns (template and output lines shared in one L2)
    const int col = __builtin_amdgcn_readfirstlane(xcd_row(blockIdx.x, gridDim.x));
    if ( pack_ends && col == 0 ) // workgroup 0 takes columns 0 and last
        tz_body<true>(sh, items, nitems, a_es, S, s_stride, o_es, M, Lmax, scale, last, col);
    else
        tz_body<false>(sh, items, nitems, a_es, S, s_stride, o_es, M, Lmax, scale, last, col);
}

} // namespace fastfft_amd
