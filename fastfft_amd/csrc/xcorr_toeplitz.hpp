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
// which equals IFFT_y(FFT_y(S) conj(FFT_y(a_t)))[j] / Fy for j + v < Fy.  The
// complex products are taken in Gauss's three-product form: with the real
// sequences U0 = Sr + Si, U1 = Sr, U2 = Si and the template vectors
// X0 = ar, X1 = -(ar + ai), X2 = ar - ai,
//   k_p[j] = sum_v U_p[j + v] X_p[v],   Re b = k0 - k2,   Im b = k0 + k1,
// three real matrix products with Toeplitz left operands (rows j) and the 32
// templates of a group as the columns of the right one, K = L.  Each is taken on
// v_mfma_f32_32x32x16_bf16 as hi*hi + hi*lo + lo*hi of split operands
// (hi = bf16(x), lo = bf16(x - hi)), accumulated in FP32.
//
// The DC and Nyquist columns (real template values; the C2R pass keeps only
// the real parts of their outputs) share one workgroup: U0 = Re S_0,
// U1 = Re S_last, X0 = a_0, X1 = a_last (U2 = X2 = 0) give them as k0 and k1.
//
// Workgroup: one column (or the DC / Nyquist pair) and every template group of
// the launch in turn; eight waves, two per SIMD, wave w owns output rows
// [64 w, 64 w + 64) as two 32-row tiles, for all 32 templates of a group.  The
// templates are the matrix instruction's A operand and the Toeplitz rows its B
// operand, so that a result register holds 32 consecutive output rows of one
// template: every store covers whole column segments of the column-major outputs.
//   * Left operands (the Toeplitz B operand): U_p split into hi / lo bf16, staged
//     once per workgroup in LDS in four copies shifted by 0..3 elements, so that
//     every lane's operand (8 consecutive elements from row j + 16 k + 8 g) is one
//     16-byte read at an 8-byte aligned index of copy j % 4; the copies start 32
//     bytes apart modulo 128, so the reads of sixteen lanes cover 128 distinct bytes.
//   * Toeplitz reuse: tile a's operand at K-step k is tile 0's at K-step
//     k + 2 a.  A chunk's even K-steps run first, then its odd ones; in each
//     pass a wave keeps tile 0's operands of three K-steps of one parity in
//     registers and reads one new operand per K-step and (p, hi / lo), one
//     K-step ahead, after two at the start of the pass.
//   * Right operands (the templates, A operand): the templates' X_p, read as
//     contiguous column segments and split on the way into LDS, in chunks of
//     eight K-steps, double buffered: the global loads of the next chunk are in
//     flight while the waves run the current one.  The operands of the next
//     K-step are read from LDS while the current K-step's 18 matrix instructions
//     run.
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

constexpr int kTzWaves     = 8;                    // waves per workgroup (two per SIMD)
constexpr int kTzTiles     = 2;                    // 32-row tiles per wave
constexpr int kTzRowsWave  = 32 * kTzTiles;        // output rows per wave
constexpr int kTzMaxM      = kTzWaves * kTzRowsWave;
constexpr int kTzTemplates = 32;                   // templates per group (the MFMA's N)
constexpr int kTzFy        = 1024;                 // frame length along y
constexpr int kTzMaxL      = 512;                  // template rows (K) at most
constexpr int kTzCopyLen   = kTzFy + 16;           // elements per shifted copy (copies 32 bytes apart modulo 128)
constexpr int kTzChunk     = 8;                    // K-steps (16 template rows each) per staged chunk

__device__ __forceinline__ tz_f32x16 tz_mfma(tz_u32x4 a, tz_u32x4 b, tz_f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(tz_bf16x8, a), __builtin_bit_cast(tz_bf16x8, b), c, 0,
                                                  0, 0);
}

// hi = bf16(x) (round to nearest even) and lo = bf16(x - hi), as bf16 bit patterns
__device__ __forceinline__ void tz_split(float x, unsigned& hi, unsigned& lo) {
    const __bf16 h = (__bf16)x;
    const __bf16 l = (__bf16)(x - (float)h);
    hi             = __builtin_bit_cast(unsigned short, h);
    lo             = __builtin_bit_cast(unsigned short, l);
}

// Workgroup LDS:
//   u  left operands: u[p][h][c][i] = U_p^h[i + c] (h = 0 hi, 1 lo; c the copy's shift)
//   x  right operands of a chunk of K-steps, double buffered: x[buf][k][p][h][g][n] holds
//      X_p^h rows 16 k + 8 g .. + 7 of template n (8 bf16)
struct TzShared {
    __attribute__((aligned(16))) unsigned short u[3][2][4][kTzCopyLen];
    tz_u32x4                                    x[2][kTzChunk][3][2][2][kTzTemplates];
};

template <bool ENDS>
__device__ __forceinline__ void tz_body(TzShared& sh, const XcorrItem* __restrict__ items, int nitems, long lda,
                                        const float2* __restrict__ S, long s_stride, long ldo, long out_step, int M,
                                        int Lmax, float scale, int last, int col) {
    const int tid     = threadIdx.x;
    const int lane    = tid & 63;
    const int w       = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g       = lane >> 5; // lane half: K rows 8 g .. 8 g + 7 of a K-step
    const int n       = lane & 31; // MFMA row (A) / column (B)
    const int nks     = (Lmax + 15) / 16;
    const int nchk    = (nks + kTzChunk - 1) / kTzChunk;
    const int ngroups = (nitems + kTzTemplates - 1) / kTzTemplates;
    const int nq      = ngroups * nchk; // staged chunks over all groups

    // ---- staging of the right operands: thread tid takes template tn = tid % 32 of the
    // group and ZR rows of the chunk (ZR = 16: K-step tk = tid / 32, both halves; ZR = 8:
    // K-step tk / 2, half tk % 2).  Every load is issued (threads past the batch on template
    // 0) and the values outside the template replaced by zeros, so no load waits for another.
    constexpr int NT = kTzWaves * 64;
    constexpr int ZR = kTzChunk * 16 * kTzTemplates / NT; // template rows per thread and chunk
    static_assert(ZR == 16 || ZR == 8, "a thread stages one K-step or one half of it");
    const int tn = tid & 31;
    const int tk = tid >> 5;
    const int zk = ZR == 16 ? tk : tk >> 1;           // K-step of the chunk
    const int zh = ZR == 16 ? 0 : tk & 1;             // first half
    float2    z[ZR];
    auto      load_chunk = [&](int q) {
        const int       grp = q / nchk, ck = q % nchk;
        const int       tt  = grp * kTzTemplates + tn;
        const XcorrItem it  = items[tt < nitems ? tt : 0];
        const auto*     ga  = as_global(static_cast<const float2*>(it.a));
        // ZR contiguous rows of the template's column (the column holds lda rows, a multiple
        // of 16; K-steps past it read its last 16 rows, which the staging zeroes); 32-bit
        // offsets, formed here so that none is kept across the chunk
        int ld = (int)lda;
        asm volatile("" : "+s"(ld));
        const int v0 = min(16 * (ck * kTzChunk + zk), ld - 16) + 8 * zh;
#pragma unroll
        for ( int i = 0; i < ZR / 2; i++ ) {
            if constexpr ( ENDS ) {
                const float4 a = load_cpx_pair(ga + v0, i);
                const float4 b = load_cpx_pair(ga + last * ld + v0, i);
                z[2 * i]       = make_float2(a.x, b.x);
                z[2 * i + 1]   = make_float2(a.z, b.z);
            }
            else {
                const float4 a = load_cpx_pair(ga + col * ld + v0, i);
                z[2 * i]       = make_float2(a.x, a.y);
                z[2 * i + 1]   = make_float2(a.z, a.w);
            }
        }
    };
    auto store_chunk = [&](int q, int into) {
        const int grp = q / nchk, ck = q % nchk;
        const int tt  = grp * kTzTemplates + tn;
        const int Lt  = tt < nitems ? items[tt].L : 0; // rows taken
        const int v0  = 16 * (ck * kTzChunk + zk);
        auto&     dst = sh.x[into][zk];
#pragma unroll
        for ( int hh = 0; hh < ZR / 8; hh++ ) {
            const int gg = zh + hh;
            unsigned  hw[3][8], lw[3][8];
#pragma unroll
            for ( int i = 0; i < 8; i++ ) {
                const bool   in = v0 + 8 * gg + i < Lt;
                const float2 a  = in ? z[8 * hh + i] : make_float2(0.f, 0.f);
                float        x0, x1, x2;
                if constexpr ( ENDS ) {
                    x0 = a.x, x1 = a.y, x2 = 0.f;
                }
                else {
                    x0 = a.x, x1 = -(a.x + a.y), x2 = a.x - a.y;
                }
                tz_split(x0, hw[0][i], lw[0][i]);
                tz_split(x1, hw[1][i], lw[1][i]);
                tz_split(x2, hw[2][i], lw[2][i]);
            }
#pragma unroll
            for ( int p = 0; p < 3; p++ ) {
                tz_u32x4 h, l;
#pragma unroll
                for ( int i = 0; i < 4; i++ ) {
                    h[i] = hw[p][2 * i] | (hw[p][2 * i + 1] << 16);
                    l[i] = lw[p][2 * i] | (lw[p][2 * i + 1] << 16);
                }
                dst[p][0][gg][tn] = h;
                dst[p][1][gg][tn] = l;
            }
        }
    };
    load_chunk(0);

    // ---- left operands: thread tid writes indices YT tid .. YT tid + YT - 1 of every copy,
    // from U_p[YT tid .. YT tid + YT + 2] (zero past the frame)
    {
        constexpr int YT = kTzFy / NT; // y values per thread
        static_assert(YT == 4 || YT == 2, "8- or 4-byte copy writes");
        const auto* s0 = as_global(S) + (long)col * s_stride;
        const auto* s1 = as_global(S) + (long)last * s_stride;
        const int   y0 = YT * tid;
        unsigned    uh[3][YT + 3], ul[3][YT + 3];
#pragma unroll
        for ( int i = 0; i < YT + 3; i++ ) {
            const int y = y0 + i;
            float     u0 = 0.f, u1 = 0.f, u2 = 0.f;
            if ( y < kTzFy ) {
                if constexpr ( ENDS ) {
                    u0 = load_cpx(s0, y).x;
                    u1 = load_cpx(s1, y).x;
                }
                else {
                    const float2 v = load_cpx(s0, y);
                    u0 = v.x + v.y, u1 = v.x, u2 = v.y;
                }
            }
            tz_split(u0, uh[0][i], ul[0][i]);
            tz_split(u1, uh[1][i], ul[1][i]);
            tz_split(u2, uh[2][i], ul[2][i]);
        }
#pragma unroll
        for ( int p = 0; p < 3; p++ )
#pragma unroll
            for ( int c = 0; c < 4; c++ ) {
                if constexpr ( YT == 4 ) {
                    const tz_u32x2 h = {uh[p][c] | (uh[p][c + 1] << 16), uh[p][c + 2] | (uh[p][c + 3] << 16)};
                    const tz_u32x2 l = {ul[p][c] | (ul[p][c + 1] << 16), ul[p][c + 2] | (ul[p][c + 3] << 16)};
                    *reinterpret_cast<tz_u32x2*>(&sh.u[p][0][c][y0]) = h;
                    *reinterpret_cast<tz_u32x2*>(&sh.u[p][1][c][y0]) = l;
                }
                else {
                    *reinterpret_cast<unsigned*>(&sh.u[p][0][c][y0]) = uh[p][c] | (uh[p][c + 1] << 16);
                    *reinterpret_cast<unsigned*>(&sh.u[p][1][c][y0]) = ul[p][c] | (ul[p][c + 1] << 16);
                }
            }
    }
    store_chunk(0, 0);
    __syncthreads( );

    // lane's left operand of tile 0 at K-step k, part (p, h): copy c = n % 4 at index
    // 128 w + n - c + 8 g + 16 k (a multiple of 4)
    const int                    cpy  = n & 3;
    const unsigned short* const  ubase = &sh.u[0][0][cpy][kTzRowsWave * w + n - cpy + 8 * g];
    constexpr int                kPart = 4 * kTzCopyLen; // elements between (p, h) parts
    auto read_a = [&](int k, int p, int h) -> tz_u32x4 {
        const unsigned short* e = ubase + kPart * (2 * p + h) + 16 * k;
        const tz_u32x2        a = *reinterpret_cast<const tz_u32x2*>(e);
        const tz_u32x2        b = *reinterpret_cast<const tz_u32x2*>(e + 4);
        return tz_u32x4{a[0], a[1], b[0], b[1]};
    };
    // A chunk's K-steps run even ones first, then odd ones (the sum does not depend on the
    // order), so that only one parity's operands are live: within a parity pass, K-step
    // ks = 8 ck + par + 2 i uses slot (i + a) % NS for tile a, which holds tile 0's operand of
    // K-step ks + 2 a.  A pass reads kTzTiles operands ahead, then during K-step i the one
    // that the last tile takes at K-step i + 1 (into the slot tile 0 of K-step i - 1 left).
    constexpr int NS = kTzTiles + 1; // operand slots of a pass
    tz_u32x4      aslot[NS][3][2];
    tz_u32x4  bop[2][3][2]; // right operands of the current and the next K-step
    tz_f32x16 acc[kTzTiles][3];

    for ( int grp = 0; grp < ngroups; grp++ ) {
#pragma unroll
      for ( int a = 0; a < kTzTiles; a++ )
#pragma unroll
        for ( int p = 0; p < 3; p++ )
            acc[a][p] = tz_f32x16{ };
      for ( int ck = 0; ck < nchk; ck++ ) {
        const int q   = grp * nchk + ck;
        const int buf = q & 1;
        // the next chunk's templates (the last chunk stages itself again into the idle
        // buffer, so that no branch depends on the chunk count)
        const int qn = min(q + 1, nq - 1);
        load_chunk(qn);
#pragma unroll
        for ( int par = 0; par < 2; par++ ) {
            // every K-step of a chunk runs: those past nks meet zero template rows (the staging
            // zero-fills rows past each template), and their left operands are read at K-steps
            // clamped to nks + 2 kTzTiles - 3, inside the staged column, so no branch splits the pass
            const int k0 = ck * kTzChunk + par; // first K-step of the pass
            const int kc = nks + 2 * kTzTiles - 3; // largest K-step whose left operand is read
#pragma unroll
            for ( int sl = 0; sl < kTzTiles; sl++ )
#pragma unroll
                for ( int p = 0; p < 3; p++ )
#pragma unroll
                    for ( int h = 0; h < 2; h++ )
                        aslot[sl][p][h] = read_a(min(k0 + 2 * sl, kc), p, h);
#pragma unroll
            for ( int p = 0; p < 3; p++ )
#pragma unroll
                for ( int h = 0; h < 2; h++ )
                    bop[0][p][h] = sh.x[buf][par][p][h][g][n];
#pragma unroll
            for ( int i = 0; i < kTzChunk / 2; i++ ) {
                const int ks = k0 + 2 * i;
                // the next K-step's operands: the last tile's (tile 0's of K-step ks + 2 kTzTiles)
                // and the right ones
                if ( i + 1 < kTzChunk / 2 ) {
#pragma unroll
                    for ( int p = 0; p < 3; p++ )
#pragma unroll
                        for ( int h = 0; h < 2; h++ )
                            aslot[(i + kTzTiles) % NS][p][h] = read_a(min(ks + 2 * kTzTiles, kc), p, h);
#pragma unroll
                    for ( int p = 0; p < 3; p++ )
#pragma unroll
                        for ( int h = 0; h < 2; h++ )
                            bop[(i + 1) & 1][p][h] = sh.x[buf][par + 2 * (i + 1)][p][h][g][n];
                }
                // hi*hi, hi*lo, lo*hi over the accumulators (3 per tile)
#pragma unroll
                for ( int t = 0; t < 3; t++ ) {
                    const int ha = t == 2 ? 1 : 0, hb = t == 1 ? 1 : 0;
#pragma unroll
                    for ( int p = 0; p < 3; p++ )
#pragma unroll
                        for ( int a = 0; a < kTzTiles; a++ )
                            acc[a][p] = tz_mfma(bop[i & 1][p][hb], aslot[(i + a) % NS][p][ha], acc[a][p]);
                }
                // issue order of the K-step: the next K-step's twelve LDS reads two at a time
                // between the first six matrix instructions (read a full K-step ahead, not
                // pushed back to where the current operands die), then the other thirty
                if ( i + 1 < kTzChunk / 2 ) {
#pragma unroll
                    for ( int j = 0; j < 6; j++ ) {
                        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0); // one MFMA
                        __builtin_amdgcn_sched_group_barrier(0x100, 2, 0); // two LDS reads
                    }
                    __builtin_amdgcn_sched_group_barrier(0x008, 30, 0);
                }
                // K-steps do not interleave (that would hold more operands live)
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // the next chunk's templates into the other buffer (last read in chunk q - 1, before
        // the barrier that ended it): their loads had the whole chunk's matrix work to arrive
        store_chunk(qn, buf ^ 1);
        if ( q + 1 < nq )
            __syncthreads( );
      }
      {
            // ---- the group's outputs: register r of tile a holds template 8 (r / 4) + 4 g + r % 4
            // of the group at output row 128 w + 32 a + n, so each store covers 32 consecutive rows
            // of a template's column (the templates' outputs are out_step elements apart)
            int oldo = (int)ldo, mm = M, ostep = (int)out_step;
            asm volatile("" : "+s"(oldo), "+s"(mm), "+s"(ostep));
            auto* go = as_global(static_cast<float2*>(items[0].out)) + (grp * kTzTemplates + 4 * g) * (long)ostep;
#pragma unroll
            for ( int a = 0; a < kTzTiles; a++ ) {
                const int j = kTzRowsWave * w + 32 * a + n;
                if ( j < mm ) {
#pragma unroll
                    for ( int r = 0; r < 16; r++ ) {
                        const int tl = 8 * (r >> 2) + (r & 3); // template of the group, less 4 g
                        if ( grp * kTzTemplates + 4 * g + tl < nitems ) {
                            const float k0 = acc[a][0][r], k1 = acc[a][1][r], k2 = acc[a][2][r];
                            auto*       to = go + tl * ostep;
                            if constexpr ( ! ENDS ) {
                                store_cpx(to, col * oldo + j, make_float2((k0 - k2) * scale, (k0 + k1) * scale));
                            }
                            else {
                                store_cpx(to, j, make_float2(k0 * scale, 0.f));
                                store_cpx(to, last * oldo + j, make_float2(k1 * scale, 0.f));
                            }
                        }
                    }
                }
                // one tile's accumulators at a time out of the matrix registers
                __builtin_amdgcn_sched_barrier(0);
            }
      }
    }
}

// Grid: x = last workgroups (columns 1 .. last - 1, and workgroup 0 for the pair
// 0 / last); 512 threads; every workgroup runs all ceil(nitems / 32) template groups.
//   S        frame image in mixed form, column c at S + c * s_stride (Fy = 1024 values)
//   items    the templates, column-major: a[c * lda + v] element v of column c of template
//            t (1 <= L_t <= Lmax rows; lda a multiple of 16, at least Lmax), and
//            out[c * ldo + j] element j of its output column; the outputs of the launch's
//            templates are out_step elements apart (items[t].out = items[0].out + t out_step)
//   M        output rows (<= 512), Lmax the largest template's rows (<= 512, with
//            Lmax + M - 1 <= 1024), last the Nyquist column index (ncols - 1)
__global__ void __launch_bounds__(kTzWaves * 64) __attribute__((amdgpu_waves_per_eu(kTzWaves / 4)))
        k_xcorr_toeplitz(const XcorrItem* __restrict__ items, int nitems, long lda, const float2* __restrict__ S,
                         long s_stride, long ldo, long out_step, int M, int Lmax, float scale, int last) {
    __shared__ TzShared sh;
    // each XCD takes a contiguous range of columns (template and output lines shared in one L2)
    const int col = __builtin_amdgcn_readfirstlane(xcd_row(blockIdx.x, gridDim.x));
    if ( col == 0 ) // workgroup 0 takes columns 0 and last
        tz_body<true>(sh, items, nitems, lda, S, s_stride, ldo, out_step, M, Lmax, scale, last, col);
    else
        tz_body<false>(sh, items, nitems, lda, S, s_stride, ldo, out_step, M, Lmax, scale, last, col);
}

} // namespace fastfft_amd
