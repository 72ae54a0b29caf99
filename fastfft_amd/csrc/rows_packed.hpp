// Row kernels for rows of N = 2P real values with one wavefront per row: the
// even and odd samples of a row form one complex P-point sequence, so a row
// costs one P-point transform, where the wave kernels of rows_wave.hpp spend
// one per decimated half (two for N = 2P).  These are the x-passes of the
// 2-D correlation, FwdFFT and InvFFT for 512- and 1024-point rows (the
// reference's block_fft_kernel_R2C_INCREASE_XY / _NONE_XY and
// C2R_DECREASE_XY / _NONE_XY, src/fastfft/FastFFT.cu:1055-1166, :1641-1692).
//
//   k_r2c_rows_packed  z[m] = x[2m] + i x[2m+1] (zero at and past L), Z = FFT_P(z),
//                      X[k] = (Z[k] + conj Z[P-k]) / 2 - i W_N^k (Z[k] - conj Z[P-k]) / 2
//                      for k = 0 .. P (Z[P] = Z[0]); the k < nout values are stored.
//   k_c2r_rows_packed  Z[k] = (X[k] + conj X[P-k]) + i W_N^{-k} (X[k] - conj X[P-k]),
//                      z = IFFT_P(Z), x[2m] = Re z[m], x[2m+1] = Im z[m] (the unnormalised
//                      N-point inverse of the Hermitian X[0 .. P], with the imaginary
//                      parts of X[0] and X[P] taken as zero, as the real part of the
//                      inverse of the Hermitian extension does); output j < M is
//                      x[(o + j) mod N].
//
// A workgroup holds kPackedRows rows, one per wave; the Stockham pass table is
// staged once per workgroup.  After the only workgroup barrier every wave works
// on its own row and its own LDS region.
#pragma once

#include "rows_wave.hpp"

namespace fastfft_amd {

constexpr int kPackedRows = 4;

// Grid: ceil(nrows / kPackedRows) x templates.  items != nullptr: template
// items[blockIdx.y] supplies in, out, the input pitch, L and its row count.
template <int P, typename TIn>
__global__ void __launch_bounds__(64 * kPackedRows)
        k_r2c_rows_packed(const TIn* __restrict__ in, LineMap im, int L, float2* __restrict__ out, LineMap om, int nout,
                          int nrows, float pre_scale, const float2* __restrict__ tw, const RowBatchItem* __restrict__ items) {
    constexpr int T = 64, E = P / T, N = 2 * P, NT = T * kPackedRows;
    using FFT       = WaveFFT<P>;
    __shared__ float2 lds[kPackedRows * FFT::SLOTS];
    __shared__ float2 s_twp[P];
    const int     w   = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    const int     t   = threadIdx.x % T;
    const float2* twm = multi_twiddles(tw);
    if ( items ) {
        const RowBatchItem it = load_item(items, blockIdx.y);
        in                    = static_cast<const TIn*>(it.in);
        out                   = static_cast<float2*>(it.out);
        im.s1                 = it.in_ld;
        L                     = it.L;
        nrows                 = it.rows;
    }
    if ( (int)blockIdx.x * kPackedRows >= nrows ) // the whole workgroup, before any barrier
        return;
    const int  row  = (int)blockIdx.x * kPackedRows + w;
    const bool live = row < nrows;
    // ---- every load of the thread issued before its LDS writes
    float2 ptv[FFT::template pass_table_rows<NT>];
    FFT::template load_pass_table<NT>(ptv, twm, threadIdx.x);
    const float2 wt = load_cpx(as_global(twm), N + t); // W_N^t
    float2       v[E];
    if ( live ) {
        const auto* rp = as_global(in) + im.base(row);
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int n = 2 * (t + T * i);
            v[i]        = make_float2(n < L ? load_real(rp, (long)n * im.es) : 0.f,
                                      n + 1 < L ? load_real(rp, (long)(n + 1) * im.es) : 0.f);
        }
    }
    FFT::template store_pass_table<NT>(s_twp, ptv, threadIdx.x);
    __syncthreads( );
    if ( ! live )
        return;
#pragma unroll
    for ( int i = 0; i < E; i++ )
        v[i] = make_float2(v[i].x * pre_scale, v[i].y * pre_scale);
    FFT fft;
    fft.init_pt(t, s_twp);
    float2* wl = lds + w * FFT::SLOTS;
    fft.template run<-1>(v, wl, t); // v[i] = Z[t + 64 i]
    // Z[P - k] from the wave's region (its last exchange was read before these writes)
#pragma unroll
    for ( int i = 0; i < E; i++ )
        wl[FFT::Layout::slot(t + T * i)] = v[i];
    wave_lds_fence( );
    auto* op = as_global(out) + om.base(row);
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const int    k  = t + T * i;
        const float2 zp = wl[FFT::Layout::slot((P - k) & (P - 1))];
        const float2 a  = make_float2(0.5f * (v[i].x + zp.x), 0.5f * (v[i].y - zp.y)); // (Z + conj Zp) / 2
        const float2 b  = make_float2(0.5f * (v[i].x - zp.x), 0.5f * (v[i].y + zp.y)); // (Z - conj Zp) / 2
        const float2 c  = cmul(cmul(twm[N + T * i], wt), b);                           // W_N^k b
        if ( k < nout )
            store_cpx(op, (long)k * om.es, make_float2(a.x + c.y, a.y - c.x)); // a - i c
    }
    // X[P] = Re Z[0] - Im Z[0]
    if ( t == 0 && P < nout )
        store_cpx(op, (long)P * om.es, make_float2(v[0].x - v[0].y, 0.f));
}

// Grid: ceil(nrows / kPackedRows) x templates.  row_keys == nullptr: no argmax;
// else the largest argmax key (fft_core.hpp) of row r, index r * M + j, goes to
// row_keys[r].  items != nullptr: template items[blockIdx.y] supplies in, out
// and row_keys.
template <int P, typename TOut>
__global__ void __launch_bounds__(64 * kPackedRows)
        k_c2r_rows_packed(const float2* __restrict__ in, LineMap im, TOut* __restrict__ out, LineMap om, int M, int o,
                          int nrows, float post_scale, const float2* __restrict__ tw,
                          unsigned long long* __restrict__ row_keys, const RowBatchItem* __restrict__ items) {
    constexpr int T = 64, E = P / T, N = 2 * P, NT = T * kPackedRows;
    using FFT       = WaveFFT<P>;
    __shared__ float2 lds[kPackedRows * FFT::SLOTS];
    __shared__ float2 s_twp[P];
    const int     w   = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    const int     t   = threadIdx.x % T;
    const float2* twm = multi_twiddles(tw);
    if ( items ) {
        const RowBatchItem it = load_item(items, blockIdx.y);
        in                    = static_cast<const float2*>(it.in);
        out                   = static_cast<TOut*>(it.out);
        row_keys              = it.keys;
    }
    if ( (int)blockIdx.x * kPackedRows >= nrows )
        return;
    const int  row  = (int)blockIdx.x * kPackedRows + w;
    const bool live = row < nrows;
    float2     ptv[FFT::template pass_table_rows<NT>];
    FFT::template load_pass_table<NT>(ptv, twm, threadIdx.x);
    const float2 wt = load_cpx(as_global(twm), N + t); // W_N^t
    // X[k] and X[P - k], k = t + 64 i (the row holds X[0 .. P])
    float2 x[E], xp[E];
    if ( live ) {
        const auto* ip = as_global(in) + im.base(row);
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int k = t + T * i;
            x[i]        = load_cpx(ip, (long)k * im.es);
            xp[i]       = load_cpx(ip, (long)(P - k) * im.es);
        }
    }
    FFT::template store_pass_table<NT>(s_twp, ptv, threadIdx.x);
    __syncthreads( );
    if ( ! live )
        return;
    // the real part of the Hermitian extension's inverse ignores Im X[0] and Im X[P]
    if ( t == 0 ) {
        x[0].y  = 0.f;
        xp[0].y = 0.f;
    }
    FFT fft;
    fft.init_pt(t, s_twp);
    float2 v[E];
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const float2 e = make_float2(x[i].x + xp[i].x, x[i].y - xp[i].y); // X + conj Xp
        const float2 d = make_float2(x[i].x - xp[i].x, x[i].y + xp[i].y); // X - conj Xp
        const float2 c = cmulc(d, cmul(twm[N + T * i], wt));              // W_N^{-k} d
        v[i]           = make_float2(e.x - c.y, e.y + c.x);               // e + i c
    }
    fft.template run<+1>(v, lds + w * FFT::SLOTS, t); // v[i] = z[t + 64 i]
    auto*              op   = as_global(out) + om.base(row);
    const int          oN   = ((o % N) + N) % N;
    unsigned long long best = 0;
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const int   n = 2 * (t + T * i);
        const float y[2] = {v[i].x * post_scale, v[i].y * post_scale};
#pragma unroll
        for ( int h = 0; h < 2; h++ ) {
            const int j = (n + h - oN + N) & (N - 1);
            if ( j < M ) {
                store_real(op, (long)j * om.es, y[h]);
                if ( row_keys ) {
                    // the value as stored (rounded when the output is half precision)
                    const auto k = argmax_key(convert_real<TOut>(y[h]), (unsigned)(row * M + j));
                    best         = k > best ? k : best;
                }
            }
        }
    }
    if ( row_keys ) {
#pragma unroll
        for ( int off = 32; off > 0; off >>= 1 ) {
            const unsigned long long other = __shfl_xor(best, off);
            best                           = other > best ? other : best;
        }
        if ( t == 0 )
            as_global(row_keys)[row] = best;
    }
}

// Column-major batch slots.  The batched correlation on a reduced frame keeps its
// intermediates column-major, A[kx][v] and B[kx][j] (column kx of template t is
// contiguous), so that the correlation passes read and write whole columns.  The
// row kernels below take kCmRows rows per workgroup (each wave kCmRows / kPackedRows
// rows in turn) and move them through an LDS tile of kCmRows rows of FFT::SLOTS
// values (pitch FFT::SLOTS + 2), so that every global access of a column is kCmRows consecutive values
// (64 bytes) of it, taken by four lanes of 16 bytes: out[k * ldc + row0 + r] / in[k * ldc + row0 + r].  Row r of the
// tile doubles as the exchange region of the wave that transforms it, so the tile
// costs no LDS beyond the waves' own regions.
constexpr int kCmRows = 8;
constexpr int kCmWaves = 8; // waves per workgroup of the column-major row kernels (one row each)

// k_r2c_rows_packed for batches with column-major output (ldc: elements between
// columns, even).  Grid: ceil(max rows / kCmRows) x templates.  Rows past the
// template's row count are stored as zeros up to the end of the workgroup's rows.
template <int P, typename TIn, int NZ = P / 64>
__global__ void __launch_bounds__(64 * kCmWaves)
        k_r2c_rows_packed_cm(long ldc, int nout, float pre_scale, const float2* __restrict__ tw,
                             const RowBatchItem* __restrict__ items) {
    constexpr int T = 64, E = P / T, N = 2 * P, NT = T * kCmWaves, RPW = kCmRows / kCmWaves;
    using FFT       = WaveFFT<P>;
    // tile row pitch: a row holds X[0 .. P] and the wave's exchange region, and rows start
    // 16 bytes apart modulo 128 (the column accesses of kCmRows rows hit distinct banks)
    constexpr int TP = FFT::SLOTS > P + 2 ? FFT::SLOTS : P + 2;
    static_assert(TP * 8 % 128 == 16, "tile rows 16 bytes apart modulo 128");
    __shared__ float2 tile[kCmRows * TP];
    __shared__ float2 s_twp[P];
    const int          w    = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    const int          t    = threadIdx.x % T;
    const float2*      twm  = multi_twiddles(tw);
    const RowBatchItem it   = load_item(items, blockIdx.y);
    const int          row0 = xcd_row(blockIdx.x, gridDim.x) * kCmRows; // neighbouring row blocks share one L2
    if ( row0 >= it.rows ) // the whole workgroup, before any barrier
        return;
    const auto* in = as_global(static_cast<const TIn*>(it.in));
    // ---- every load of the thread issued before its LDS writes: the pass table and the
    // wave's rows row0 + w + kCmWaves r
    float2 ptv[FFT::template pass_table_rows<NT>];
    FFT::template load_pass_table<NT>(ptv, twm, threadIdx.x);
    const float2 wt = load_cpx(as_global(twm), N + t); // W_N^t
    float2       v[RPW][E];
    // even row pitch on an aligned buffer: each lane's two values as one load
    const bool pairs = (it.in_ld & 1) == 0 && (reinterpret_cast<unsigned long>(it.in) & (2 * sizeof(TIn) - 1)) == 0;
#pragma unroll
    for ( int r = 0; r < RPW; r++ ) {
        const int   row = row0 + w + kCmWaves * r;
        const auto* rp  = in + (long)min(row, it.rows - 1) * it.in_ld;
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int n = 2 * (t + T * i);
            if ( pairs && n + 1 < it.L )
                v[r][i] = load_real_pair(rp, (long)n);
            else
                v[r][i] = make_float2(n < it.L ? load_real(rp, (long)n) : 0.f, n + 1 < it.L ? load_real(rp, (long)(n + 1)) : 0.f);
        }
    }
    FFT::template store_pass_table<NT>(s_twp, ptv, threadIdx.x);
    __syncthreads( );
    FFT fft;
    fft.init_pt(t, s_twp);
#pragma unroll
    for ( int r = 0; r < RPW; r++ ) {
        const int  rr   = w + kCmWaves * r; // row within the tile, and the wave's region
        const bool live = row0 + rr < it.rows;
        float2*    wl   = tile + rr * TP;
#pragma unroll
        for ( int i = 0; i < E; i++ )
            v[r][i] = make_float2(v[r][i].x * pre_scale, v[r][i].y * pre_scale);
        // registers NZ .. E - 1 hold zeros (no template row is longer than 128 NZ values): the
        // first pass drops their terms
        fft.template run<-1, NZ>(v[r], wl, t); // v[r][i] = Z[t + 64 i]
        // Z[P - k] through the region (its last exchange was read before these writes)
#pragma unroll
        for ( int i = 0; i < E; i++ )
            wl[FFT::Layout::slot(t + T * i)] = v[r][i];
        wave_lds_fence( );
        float2 zp[E];
#pragma unroll
        for ( int i = 0; i < E; i++ )
            zp[i] = wl[FFT::Layout::slot((P - t - T * i) & (P - 1))];
        wave_lds_fence( ); // every lane's Z reads are done before X overwrites the region
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int    k = t + T * i;
            const float2 a = make_float2(0.5f * (v[r][i].x + zp[i].x), 0.5f * (v[r][i].y - zp[i].y));
            const float2 b = make_float2(0.5f * (v[r][i].x - zp[i].x), 0.5f * (v[r][i].y + zp[i].y));
            const float2 c = cmul(cmul(twm[N + T * i], wt), b);
            wl[k]          = live ? make_float2(a.x + c.y, a.y - c.x) : make_float2(0.f, 0.f);
        }
        if ( t == 0 )
            wl[P] = live ? make_float2(v[r][0].x - v[r][0].y, 0.f) : make_float2(0.f, 0.f);
    }
    __syncthreads( );
    // ---- columns k < nout: thread pairs store the tile's kCmRows values of a column
    auto* op = as_global(static_cast<float2*>(it.out)) + row0;
    constexpr int HQ = kCmRows / 2; // lanes per column, 2 rows (16 bytes) each
    for ( int idx = threadIdx.x; idx < HQ * nout; idx += NT ) {
        const int     k  = idx / HQ, h = idx % HQ;
        const float2* tp = tile + 2 * h * TP + k;
        store_cpx_pair(op + (long)k * ldc + 2 * h, 0, tp[0], tp[TP]);
    }
}

// k_c2r_rows_packed for batches with column-major input (ldc: elements between
// columns, even; the input column holds at least ceil(nrows / kCmRows) kCmRows
// values).  Grid: ceil(nrows / kCmRows) x templates.  KEEP < E: the window lies in
// outputs 0 .. 128 KEEP - 1 (origin 0), so registers KEEP .. E - 1 of the inverse are not
// needed and the last pass does not form them.
template <int P, typename TOut, int KEEP = P / 64>
__global__ void __launch_bounds__(64 * kCmWaves)
        k_c2r_rows_packed_cm(long ldc, LineMap om, int M, int o, int nrows, float post_scale, const float2* __restrict__ tw,
                             const RowBatchItem* __restrict__ items) {
    constexpr int T = 64, E = P / T, N = 2 * P, NT = T * kCmWaves, RPW = kCmRows / kCmWaves;
    using FFT       = WaveFFT<P>;
    // tile row pitch: a row holds X[0 .. P] and the wave's exchange region, and rows start
    // 16 bytes apart modulo 128 (the column accesses of kCmRows rows hit distinct banks)
    constexpr int TP = FFT::SLOTS > P + 2 ? FFT::SLOTS : P + 2;
    static_assert(TP * 8 % 128 == 16, "tile rows 16 bytes apart modulo 128");
    __shared__ float2 tile[kCmRows * TP];
    __shared__ float2 s_twp[P];
    const int          w    = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    const int          t    = threadIdx.x % T;
    const float2*      twm  = multi_twiddles(tw);
    const RowBatchItem it   = load_item(items, blockIdx.y);
    const int          row0 = xcd_row(blockIdx.x, gridDim.x) * kCmRows; // neighbouring row blocks share one L2
    if ( row0 >= nrows )
        return;
    // ---- the tile: columns k <= P, kCmRows values each, then the pass table
    {
        const auto* ip = as_global(static_cast<const float2*>(it.in)) + row0;
        float2      ptv[FFT::template pass_table_rows<NT>];
        FFT::template load_pass_table<NT>(ptv, twm, threadIdx.x);
        // lanes per column HQ, 2 rows (16 bytes) each, so that four lanes take a column's 64
        // bytes; in two rounds of loads (fewer registers live)
        constexpr int HQ = kCmRows / 2;
        constexpr int NL = (HQ * (P + 1) + NT - 1) / NT, NL1 = (NL + 1) / 2;
#pragma unroll
        for ( int l0 = 0; l0 < NL; l0 += NL1 ) {
            float4 a[NL1];
#pragma unroll
            for ( int l = 0; l < NL1; l++ ) {
                const int idx = threadIdx.x + NT * (l0 + l);
                const int k   = min(idx / HQ, P), h = idx % HQ;
                a[l]          = load_cpx_pair(ip + (long)k * ldc + 2 * h, 0);
            }
#pragma unroll
            for ( int l = 0; l < NL1; l++ ) {
                const int idx = threadIdx.x + NT * (l0 + l);
                if ( l0 + l < NL && idx < HQ * (P + 1) ) {
                    float2* tp = tile + 2 * (idx % HQ) * TP + idx / HQ;
                    tp[0]      = make_float2(a[l].x, a[l].y);
                    tp[TP]     = make_float2(a[l].z, a[l].w);
                }
            }
        }
        FFT::template store_pass_table<NT>(s_twp, ptv, threadIdx.x);
    }
    __syncthreads( );
    const float2 wt = load_cpx(as_global(twm), N + t); // W_N^t
    FFT fft;
    fft.init_pt(t, s_twp);
    const int oN = ((o % N) + N) % N;
#pragma unroll
    for ( int r = 0; r < RPW; r++ ) {
        const int rr  = w + kCmWaves * r;
        const int row = row0 + rr;
        if ( row >= nrows ) // wave-uniform; no barrier follows
            break;
        // the row into registers, then its tile row serves as the transform's exchange region
        // (only this wave reads or writes it)
        float2        vr[E];
        float2*       tr = tile + rr * TP;
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int k  = t + T * i;
            float2    x  = tr[k];
            float2    xp = tr[P - k];
            if ( k == 0 ) { // the real part of the Hermitian extension's inverse ignores Im X[0], Im X[P]
                x.y  = 0.f;
                xp.y = 0.f;
            }
            const float2 e = make_float2(x.x + xp.x, x.y - xp.y);
            const float2 d = make_float2(x.x - xp.x, x.y + xp.y);
            const float2 c = cmulc(d, cmul(twm[N + T * i], wt));
            vr[i]          = make_float2(e.x - c.y, e.y + c.x);
        }
        wave_lds_fence( ); // every lane's reads of the row are done before the transform writes it
        fft.template run<+1>(vr, tr, t); // vr[i] = z[t + 64 i]
        auto*              op   = as_global(static_cast<TOut*>(it.out)) + om.base(row);
        unsigned long long best = 0;
        // every kept value inside the window, at unit stride on an aligned row: one store per pair
        const bool whole = oN == 0 && M >= 2 * T * KEEP && om.es == 1 &&
                           (reinterpret_cast<unsigned long>(op) & (2 * sizeof(TOut) - 1)) == 0;
        if ( whole ) {
#pragma unroll
            for ( int i = 0; i < KEEP; i++ ) {
                const int   n  = 2 * (t + T * i);
                const float y0 = vr[i].x * post_scale, y1 = vr[i].y * post_scale;
                store_real_pair(op, n, y0, y1);
                if ( it.keys ) {
                    const auto k0 = argmax_key(convert_real<TOut>(y0), (unsigned)(row * M + n));
                    const auto k1 = argmax_key(convert_real<TOut>(y1), (unsigned)(row * M + n + 1));
                    best          = k0 > best ? k0 : best;
                    best          = k1 > best ? k1 : best;
                }
            }
        }
        else {
#pragma unroll
            for ( int i = 0; i < KEEP; i++ ) {
                const int   n    = 2 * (t + T * i);
                const float y[2] = {vr[i].x * post_scale, vr[i].y * post_scale};
#pragma unroll
                for ( int h = 0; h < 2; h++ ) {
                    const int j = (n + h - oN + N) & (N - 1);
                    if ( j < M ) {
                        store_real(op, (long)j * om.es, y[h]);
                        if ( it.keys ) {
                            const auto k = argmax_key(convert_real<TOut>(y[h]), (unsigned)(row * M + j));
                            best         = k > best ? k : best;
                        }
                    }
                }
            }
        }
        if ( it.keys ) {
#pragma unroll
            for ( int off = 32; off > 0; off >>= 1 ) {
                const unsigned long long other = __shfl_xor(best, off);
                best                           = other > best ? other : best;
            }
            if ( t == 0 )
                as_global(it.keys)[row] = best;
        }
    }
}

// Template-matching statistics over a chunk of templates (no counterpart in the
// reference; the per-pixel consumer of a 2DTM search): the wave of row r runs
// the C2R of that row for items[0 .. nitems) in order and keeps, for each
// window position p = r * M + j, the largest value and the first item reaching
// it, the sum and the sum of squares, then updates
//   mip[p] = max(mip[p], local max)  (best[p] = first_id + item when it rises),
//   sum[p] += local sum, sum_sq[p] += local sum of squares.
// Each transformed row goes through the wave's LDS region, so that lane t keeps
// only window outputs j = 2 (t + 64 r) + h (HALF: M <= N/2, half the registers).
// Every position belongs to one wave, so the update needs no atomics; chunks
// launched in order on one stream see each other's results.  Input rows are
// unit-stride (im.es == 1).  With gridDim.y = G > 1 template groups, group g
// takes items [g per, (g + 1) per) and writes its statistics to its own
// partial maps part + 4 g npix (mip, best, sum, sum_sq as 32-bit words) for
// k_stats_merge.
// CM: column-major input (element k of row r at k * im.es + r, im.es even): per item the
// workgroup first copies its kPackedRows rows' columns into an LDS tile.
template <int P, bool HALF, bool CM = false>
__global__ void __launch_bounds__(64 * kPackedRows)
        k_c2r_stats_packed(LineMap im, int M, int o, int nrows, float post_scale, const float2* __restrict__ tw,
                           const RowBatchItem* __restrict__ items, int nitems, int first_id, float* __restrict__ mip,
                           int* __restrict__ best, float* __restrict__ sum, float* __restrict__ sum_sq,
                           float* __restrict__ part) {
    constexpr int T = 64, E = P / T, N = 2 * P, NT = T * kPackedRows, R = HALF ? E / 2 : E;
    if ( gridDim.y > 1 ) {
        const int  per  = (nitems + (int)gridDim.y - 1) / (int)gridDim.y;
        const int  g0   = (int)blockIdx.y * per;
        const long npix = (long)nrows * M;
        items += g0;
        first_id += g0;
        nitems = min(per, nitems - g0);
        mip    = part + 4 * blockIdx.y * npix;
        best   = reinterpret_cast<int*>(mip + npix);
        sum    = mip + 2 * npix;
        sum_sq = mip + 3 * npix;
    }
    using FFT       = WaveFFT<P>;
    constexpr int CP = kPackedRows + 1; // tile pitch (CM)
    __shared__ float2 lds[kPackedRows * FFT::SLOTS];
    __shared__ float2 s_twp[P];
    __shared__ float2 tile[CM ? (P + 1) * CP : 1];
    const int     w   = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    const int     t   = threadIdx.x % T;
    const float2* twm = multi_twiddles(tw);
    const int rblk = CM ? xcd_row(blockIdx.x, gridDim.x) : (int)blockIdx.x; // (CM: neighbouring row blocks share one L2)
    if ( rblk * kPackedRows >= nrows )
        return;
    const int  row  = rblk * kPackedRows + w;
    const bool live = row < nrows;
    FFT::template fill_pass_table<NT>(s_twp, twm, threadIdx.x);
    __syncthreads( );
    if ( ! CM && ! live ) // (CM: every wave takes part in the tile barriers)
        return;
    const float2 wt = load_cpx(as_global(twm), N + t); // W_N^t
    FFT          fft;
    fft.init_pt(t, s_twp);
    float2*   wl = lds + w * FFT::SLOTS;
    const int oN = ((o % N) + N) % N;
    // per window position j = 2 (t + 64 r) + h: local max, its item, sum, sum of squares
    float lmax[R][2], lsum[R][2], lsq[R][2];
    int   litem[R][2];
#pragma unroll
    for ( int r = 0; r < R; r++ )
#pragma unroll
        for ( int h = 0; h < 2; h++ ) {
            lmax[r][h]  = -__builtin_huge_valf( );
            litem[r][h] = 0;
            lsum[r][h] = lsq[r][h] = 0.f;
        }
    for ( int it = 0; it < nitems; it++ ) {
        const RowBatchItem item = load_item(items, it);
        float2             x[E], xp[E];
        if constexpr ( CM ) {
            // the workgroup's rows: kPackedRows consecutive values of every column k <= P
            const int   r0 = rblk * kPackedRows;
            const auto* ip = as_global(static_cast<const float2*>(item.in)) + r0;
            __syncthreads( ); // the previous item's tile reads are done
            for ( int k = threadIdx.x; k < P + 1; k += NT ) {
                const float4 a  = load_cpx_pair(ip + (long)k * im.es, 0);
                const float4 b  = load_cpx_pair(ip + (long)k * im.es, 1);
                float2*      tp = tile + k * CP;
                tp[0]           = make_float2(a.x, a.y);
                tp[1]           = make_float2(a.z, a.w);
                tp[2]           = make_float2(b.x, b.y);
                tp[3]           = make_float2(b.z, b.w);
            }
            __syncthreads( );
            if ( ! live )
                continue;
#pragma unroll
            for ( int i = 0; i < E; i++ ) {
                const int k = t + T * i;
                x[i]        = tile[k * CP + w];
                xp[i]       = tile[(P - k) * CP + w];
            }
        }
        else {
            // unit-stride rows (im.es == 1): X[t + 64 i] and X[P - t - 64 i] at constant
            // offsets from two lane pointers
            const auto* ip = as_global(static_cast<const float2*>(item.in)) + im.base(row);
            const auto* lo = ip + t;
            const auto* hi = ip + (P - t);
#pragma unroll
            for ( int i = 0; i < E; i++ ) {
                x[i]  = load_cpx(lo, T * i);
                xp[i] = load_cpx(hi, -T * i);
            }
        }
        if ( t == 0 ) {
            x[0].y  = 0.f;
            xp[0].y = 0.f;
        }
        float2 v[E];
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const float2 e = make_float2(x[i].x + xp[i].x, x[i].y - xp[i].y);
            const float2 d = make_float2(x[i].x - xp[i].x, x[i].y + xp[i].y);
            const float2 c = cmulc(d, cmul(twm[N + T * i], wt));
            v[i]           = make_float2(e.x - c.y, e.y + c.x);
        }
        fft.template run<+1>(v, wl, t); // v[i] = z[t + 64 i]: x[2m] = Re, x[2m+1] = Im
#pragma unroll
        for ( int i = 0; i < E; i++ )
            wl[FFT::Layout::slot(t + T * i)] = v[i];
        wave_lds_fence( );
#pragma unroll
        for ( int r = 0; r < R; r++ ) {
#pragma unroll
            for ( int h = 0; h < 2; h++ ) {
                const int    j = 2 * (t + T * r) + h;
                const int    n = (oN + j) & (N - 1);
                const float2 z = wl[FFT::Layout::slot(n >> 1)];
                const float  y = ((n & 1) ? z.y : z.x) * post_scale;
                if ( j < M ) {
                    if ( y > lmax[r][h] ) {
                        lmax[r][h]  = y;
                        litem[r][h] = it;
                    }
                    lsum[r][h] += y;
                    lsq[r][h] += y * y;
                }
            }
        }
        wave_lds_fence( ); // the window reads are done before the next transform writes the region
    }
    if ( ! live ) // (CM: the waves past the last row stayed for the tile barriers)
        return;
    auto* gm = as_global(mip);
    auto* gb = as_global(best);
    auto* gs = as_global(sum);
    auto* gq = as_global(sum_sq);
#pragma unroll
    for ( int r = 0; r < R; r++ ) {
#pragma unroll
        for ( int h = 0; h < 2; h++ ) {
            const int j = 2 * (t + T * r) + h;
            if ( j < M ) {
                const long p = (long)row * M + j;
                if ( gridDim.y > 1 ) {
                    // a group's partial maps: written, not merged (an empty group leaves -inf and zeros)
                    gm[p] = lmax[r][h];
                    gb[p] = first_id + litem[r][h];
                    gs[p] = lsum[r][h];
                    gq[p] = lsq[r][h];
                }
                else {
                    if ( lmax[r][h] > gm[p] ) {
                        gm[p] = lmax[r][h];
                        gb[p] = first_id + litem[r][h];
                    }
                    gs[p] = gs[p] + lsum[r][h];
                    gq[p] = gq[p] + lsq[r][h];
                }
            }
        }
    }
}

// C2R of rows of N = 2 PP real values, PP = P Q complex points (N = 2048, 4096,
// 8192 with P = 512), any window of M <= N outputs, one workgroup of Q waves per
// row (the trimmed C2R x-pass of InvFFT and of the batch's frame reduction; the
// reference's block_fft_kernel_C2R_DECREASE_XY / _NONE_XY, FastFFT.cu:1641-1692).
// The row's even and odd outputs come from one complex PP-point inverse:
//   Z[k] = (X[k] + conj X[PP-k]) + i W_N^{-k} (X[k] - conj X[PP-k]),  k < PP
//   z = IFFT_PP(Z), x[2m] = Re z[m], x[2m+1] = Im z[m],
// with Im X[0] and Im X[PP] taken as zero.  Wave q computes U_q = IFFT_P of
// Z[Q k' + q] (decimation in time) and each output evaluates
//   z[m] = sum_q W_PP^{-m q} U_q[m mod P]
// by Horner's rule in W_PP^{-m}; output j < M is x[(o + j) mod N].
template <int P, int Q, typename TOut>
__global__ void __launch_bounds__(64 * Q)
        k_c2r_window_packed(const float2* __restrict__ in, LineMap im, TOut* __restrict__ out, LineMap om, int M, int o,
                            int nlines, float post_scale, const float2* __restrict__ tw,
                            unsigned long long* __restrict__ row_keys) {
    constexpr int T = 64, E = P / T, PP = P * Q, N = 2 * PP, NT = T * Q;
    using FFT       = WaveFFT<P>;
    constexpr int LDS = (Q * FFT::SLOTS > PP + 1 + PP / 32 ? Q * FFT::SLOTS : PP + 1 + PP / 32);
    __shared__ float2 lds[LDS]; // the input row (padded), then the exchange buffers and U_q
    __shared__ float2 s_twp[P];
    const int     q    = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    const int     t    = threadIdx.x % T;
    const int     line = xcd_row(blockIdx.x, nlines);
    const float2* twm  = multi_twiddles(tw);
    auto          pad  = [](int k) { return k + (k >> 5); }; // spreads the stride-Q reads over the banks
    // ---- the row X[0 .. PP] into LDS, every load of the thread before its LDS writes
    {
        float2 ptv[FFT::template pass_table_rows<NT>];
        FFT::template load_pass_table<NT>(ptv, twm, threadIdx.x);
        constexpr int R  = (PP + 1 + NT - 1) / NT;
        const auto*   lp = as_global(in) + im.base(line);
        float2        xr[R];
#pragma unroll
        for ( int r = 0; r < R; r++ ) {
            const int k = threadIdx.x + NT * r;
            xr[r]       = k <= PP ? load_cpx(lp, (long)k * im.es) : make_float2(0.f, 0.f);
        }
        FFT::template store_pass_table<NT>(s_twp, ptv, threadIdx.x);
#pragma unroll
        for ( int r = 0; r < R; r++ ) {
            const int k = threadIdx.x + NT * r;
            if ( k <= PP )
                lds[pad(k)] = xr[r];
        }
    }
    __syncthreads( );
    FFT fft;
    fft.init_pt(t, s_twp);
    // ---- Z[Q k' + q], k' = t + 64 i
    float2 v[E];
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const int k  = Q * (t + T * i) + q;
        float2    xk = lds[pad(k)], xp = lds[pad(PP - k)];
        if ( k == 0 ) {
            xk.y = 0.f;
            xp.y = 0.f;
        }
        const float2 e = make_float2(xk.x + xp.x, xk.y - xp.y);                  // X + conj Xp
        const float2 d = make_float2(xk.x - xp.x, xk.y + xp.y);                  // X - conj Xp
        const float2 c = cmulc(d, load_cpx(as_global(twm), N + k));              // W_N^{-k} d
        v[i]           = make_float2(e.x - c.y, e.y + c.x);                      // e + i c
    }
    __syncthreads( ); // the row is read; its LDS becomes the exchange buffers
    float2* wl = lds + q * FFT::SLOTS;
    fft.template run<+1>(v, wl, t);
    // U_q[m], m = t + 64 i, in natural order in this wave's region
#pragma unroll
    for ( int i = 0; i < E; i++ )
        wl[t + T * i] = v[i];
    __syncthreads( );
    // ---- the window
    const int          oN   = ((o % N) + N) % N;
    auto*              op   = as_global(out) + om.base(line);
    unsigned long long best = 0;
    for ( int j = threadIdx.x; j < M; j += NT ) {
        const int    n  = (oN + j) & (N - 1);
        const int    m  = n >> 1;
        const int    mm = m & (P - 1);
        const float2 w  = load_cpx(as_global(twm), PP + m); // W_PP^{m}; the inverse takes its conjugate
        float2       z  = lds[(Q - 1) * FFT::SLOTS + mm];
#pragma unroll
        for ( int qq = Q - 2; qq >= 0; qq-- )
            z = cadd(cmulc(z, w), lds[qq * FFT::SLOTS + mm]);
        const float y = ((n & 1) ? z.y : z.x) * post_scale;
        store_real(op, (long)j * om.es, y);
        if ( row_keys ) {
            // the value as stored (rounded when the output is half precision)
            const auto k = argmax_key(convert_real<TOut>(y), (unsigned)(line * M + j));
            best         = k > best ? k : best;
        }
    }
    if ( row_keys ) { // the same for every thread of the block
        __shared__ unsigned long long s_key[Q];
#pragma unroll
        for ( int off = 32; off > 0; off >>= 1 ) {
            const unsigned long long other = __shfl_xor(best, off);
            best                           = other > best ? other : best;
        }
        if ( t == 0 )
            s_key[q] = best;
        __syncthreads( );
        if ( threadIdx.x == 0 ) {
#pragma unroll
            for ( int qq = 1; qq < Q; qq++ )
                best = s_key[qq] > best ? s_key[qq] : best;
            as_global(row_keys)[line] = best;
        }
    }
}

// A trimmed C2R row followed by the R2C of the same row (no counterpart in the
// reference; the image's reduction to a 1024-point-wide correlation frame and
// that frame's forward rows): one workgroup of Q waves per line computes the
// M <= 1024 window outputs of the N = 1024 Q-point C2R exactly as
// k_c2r_window_packed does (times post_scale), keeps them in LDS instead of
// storing them, and wave 0 then runs the 1024-point packed R2C of
// k_r2c_rows_packed on them (times pre_scale, zero past M), storing the k < nout
// spectrum values at om2.  Outputs equal those of the two kernels run in turn,
// without the real row's round trip through HBM.
template <int Q>
__global__ void __launch_bounds__(64 * Q)
        k_c2r_r2c_rows(const float2* __restrict__ in, LineMap im, int M, int o, int nlines, float post_scale,
                       float2* __restrict__ out, LineMap om2, int nout, float pre_scale, const float2* __restrict__ tw) {
    constexpr int P = 512, T = 64, E = P / T, PP = P * Q, N = 2 * PP, NT = T * Q, N2 = 2 * P;
    using FFT       = WaveFFT<P>;
    constexpr int LDS = (Q * FFT::SLOTS > PP + 1 + PP / 32 ? Q * FFT::SLOTS : PP + 1 + PP / 32);
    __shared__ float2 lds[LDS]; // the input row (padded), then the exchange buffers and U_q
    __shared__ float2 s_twp[P];
    __shared__ float2 s_row[P]; // the real window row, x[2m] + i x[2m+1]
    const int     q    = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    const int     t    = threadIdx.x % T;
    const int     line = xcd_row(blockIdx.x, nlines);
    const float2* twm  = multi_twiddles(tw);
    auto          pad  = [](int k) { return k + (k >> 5); };
    {
        float2 ptv[FFT::template pass_table_rows<NT>];
        FFT::template load_pass_table<NT>(ptv, twm, threadIdx.x);
        constexpr int R  = (PP + 1 + NT - 1) / NT;
        const auto*   lp = as_global(in) + im.base(line);
        float2        xr[R];
#pragma unroll
        for ( int r = 0; r < R; r++ ) {
            const int k = threadIdx.x + NT * r;
            xr[r]       = k <= PP ? load_cpx(lp, (long)k * im.es) : make_float2(0.f, 0.f);
        }
        FFT::template store_pass_table<NT>(s_twp, ptv, threadIdx.x);
#pragma unroll
        for ( int r = 0; r < R; r++ ) {
            const int k = threadIdx.x + NT * r;
            if ( k <= PP )
                lds[pad(k)] = xr[r];
        }
    }
    __syncthreads( );
    FFT fft;
    fft.init_pt(t, s_twp);
    float2 v[E];
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const int k  = Q * (t + T * i) + q;
        float2    xk = lds[pad(k)], xp = lds[pad(PP - k)];
        if ( k == 0 ) {
            xk.y = 0.f;
            xp.y = 0.f;
        }
        const float2 e = make_float2(xk.x + xp.x, xk.y - xp.y);     // X + conj Xp
        const float2 d = make_float2(xk.x - xp.x, xk.y + xp.y);     // X - conj Xp
        const float2 c = cmulc(d, load_cpx(as_global(twm), N + k)); // W_N^{-k} d
        v[i]           = make_float2(e.x - c.y, e.y + c.x);         // e + i c
    }
    __syncthreads( );
    float2* wl = lds + q * FFT::SLOTS;
    fft.template run<+1>(v, wl, t);
#pragma unroll
    for ( int i = 0; i < E; i++ )
        wl[t + T * i] = v[i];
    __syncthreads( );
    // ---- the window row, zero past M, into s_row
    const int oN = ((o % N) + N) % N;
    float*    xr = reinterpret_cast<float*>(s_row);
    for ( int j = threadIdx.x; j < N2; j += NT ) {
        float y = 0.f;
        if ( j < M ) {
            const int    n  = (oN + j) & (N - 1);
            const int    m  = n >> 1;
            const int    mm = m & (P - 1);
            const float2 w  = load_cpx(as_global(twm), PP + m);
            float2       z  = lds[(Q - 1) * FFT::SLOTS + mm];
#pragma unroll
            for ( int qq = Q - 2; qq >= 0; qq-- )
                z = cadd(cmulc(z, w), lds[qq * FFT::SLOTS + mm]);
            y = ((n & 1) ? z.y : z.x) * post_scale;
        }
        xr[j] = y;
    }
    __syncthreads( ); // U_q are read; wave 0 reuses its region for the R2C
    if ( q != 0 )
        return;
    const float2 wt = load_cpx(as_global(twm), N2 + t); // W_N2^t
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const float2 a = s_row[t + T * i];
        v[i]           = make_float2(a.x * pre_scale, a.y * pre_scale);
    }
    fft.template run<-1>(v, lds, t); // v[i] = Z[t + 64 i]
#pragma unroll
    for ( int i = 0; i < E; i++ )
        lds[FFT::Layout::slot(t + T * i)] = v[i];
    wave_lds_fence( );
    auto* op = as_global(out) + om2.base(line);
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const int    k  = t + T * i;
        const float2 zp = lds[FFT::Layout::slot((P - k) & (P - 1))];
        const float2 a  = make_float2(0.5f * (v[i].x + zp.x), 0.5f * (v[i].y - zp.y)); // (Z + conj Zp) / 2
        const float2 b  = make_float2(0.5f * (v[i].x - zp.x), 0.5f * (v[i].y + zp.y)); // (Z - conj Zp) / 2
        const float2 c  = cmul(cmul(twm[N2 + T * i], wt), b);                          // W_N2^k b
        if ( k < nout )
            store_cpx(op, (long)k * om2.es, make_float2(a.x + c.y, a.y - c.x)); // a - i c
    }
    if ( t == 0 && P < nout )
        store_cpx(op, (long)P * om2.es, make_float2(v[0].x - v[0].y, 0.f));
}

// k_c2r_r2c_rows on RB neighbouring rows per workgroup (RB Q waves, Q per row).
// Rows y0 .. y0 + RB - 1 of the input and of the output map lie next to each
// other (line stride one element, element stride even), so element k of the
// RB rows is one contiguous segment of 8 RB bytes, moved as RB / 2 16-byte
// accesses instead of RB scattered 8-byte ones (the frame's rows are columns
// of the y-pass output, 8 KB apart per element).  Each row is computed as in
// k_c2r_r2c_rows by its own Q waves; the RB rows' spectrum values are gathered
// in LDS and the whole workgroup stores them.
template <int Q, int RB>
__global__ void __launch_bounds__(64 * Q * RB)
        k_c2r_r2c_rows_tile(const float2* __restrict__ in, LineMap im, int M, int o, int nlines, float post_scale,
                            float2* __restrict__ out, LineMap om2, int nout, float pre_scale,
                            const float2* __restrict__ tw) {
    static_assert(RB == 2 || RB == 4, "two or four rows per workgroup");
    constexpr int P = 512, T = 64, E = P / T, PP = P * Q, N = 2 * PP, NT = T * Q, N2 = 2 * P, NB = NT * RB;
    constexpr int RP = RB / 2; // 16-byte row pairs per element
    using FFT       = WaveFFT<P>;
    constexpr int LDS = (Q * FFT::SLOTS > PP + 1 + PP / 32 ? Q * FFT::SLOTS : PP + 1 + PP / 32);
    constexpr int ROW = P + 2; // s_row pitch: the real row as P complex values, later the P + 1 spectrum values
    __shared__ float2 lds[RB][LDS];
    __shared__ float2 s_twp[P];
    __shared__ float2 s_row[RB][ROW];
    const int     w   = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    const int     r   = w / Q; // the workgroup's row of this wave
    const int     q   = w % Q; // its sub-transform
    const int     t   = threadIdx.x % T;
    const int     tr  = threadIdx.x % NT; // thread index within the row
    const int     y0  = xcd_row(blockIdx.x, nlines / RB) * RB;
    const float2* twm = multi_twiddles(tw);
    auto          pad = [](int k) { return k + (k >> 5); };
    {
        float2 ptv[FFT::template pass_table_rows<NB>];
        FFT::template load_pass_table<NB>(ptv, twm, threadIdx.x);
        constexpr int R  = ((PP + 1) * RP + NB - 1) / NB;
        const auto*   lp = as_global(in) + im.base(y0);
        float4        xr[R];
#pragma unroll
        for ( int rr = 0; rr < R; rr++ ) {
            const int g = threadIdx.x + NB * rr;
            const int k = g / RP, p = g % RP;
            xr[rr]      = k <= PP ? load_cpx_pair(lp, ((long)k * im.es + 2 * p) >> 1) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        FFT::template store_pass_table<NB>(s_twp, ptv, threadIdx.x);
#pragma unroll
        for ( int rr = 0; rr < R; rr++ ) {
            const int g = threadIdx.x + NB * rr;
            const int k = g / RP, p = g % RP;
            if ( k <= PP ) {
                lds[2 * p][pad(k)]     = make_float2(xr[rr].x, xr[rr].y);
                lds[2 * p + 1][pad(k)] = make_float2(xr[rr].z, xr[rr].w);
            }
        }
    }
    __syncthreads( );
    float2* rl = lds[r];
    FFT     fft;
    fft.init_pt(t, s_twp);
    float2 v[E];
#pragma unroll
    for ( int i = 0; i < E; i++ ) {
        const int k  = Q * (t + T * i) + q;
        float2    xk = rl[pad(k)], xp = rl[pad(PP - k)];
        if ( k == 0 ) {
            xk.y = 0.f;
            xp.y = 0.f;
        }
        const float2 e = make_float2(xk.x + xp.x, xk.y - xp.y);     // X + conj Xp
        const float2 d = make_float2(xk.x - xp.x, xk.y + xp.y);     // X - conj Xp
        const float2 c = cmulc(d, load_cpx(as_global(twm), N + k)); // W_N^{-k} d
        v[i]           = make_float2(e.x - c.y, e.y + c.x);         // e + i c
    }
    __syncthreads( );
    float2* wl = rl + q * FFT::SLOTS;
    fft.template run<+1>(v, wl, t);
#pragma unroll
    for ( int i = 0; i < E; i++ )
        wl[t + T * i] = v[i];
    __syncthreads( );
    // ---- the window row, zero past M, into s_row[r]
    const int oN = ((o % N) + N) % N;
    float*    xw = reinterpret_cast<float*>(s_row[r]);
    for ( int j = tr; j < N2; j += NT ) {
        float y = 0.f;
        if ( j < M ) {
            const int    n  = (oN + j) & (N - 1);
            const int    m  = n >> 1;
            const int    mm = m & (P - 1);
            const float2 wm = load_cpx(as_global(twm), PP + m);
            float2       z  = rl[(Q - 1) * FFT::SLOTS + mm];
#pragma unroll
            for ( int qq = Q - 2; qq >= 0; qq-- )
                z = cadd(cmulc(z, wm), rl[qq * FFT::SLOTS + mm]);
            y = ((n & 1) ? z.y : z.x) * post_scale;
        }
        xw[j] = y;
    }
    __syncthreads( ); // U_q are read; wave 0 of each row reuses the row's region for the R2C
    if ( q == 0 ) {
        const float2 wt = load_cpx(as_global(twm), N2 + t); // W_N2^t
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const float2 a = s_row[r][t + T * i];
            v[i]           = make_float2(a.x * pre_scale, a.y * pre_scale);
        }
        fft.template run<-1>(v, rl, t); // v[i] = Z[t + 64 i]
#pragma unroll
        for ( int i = 0; i < E; i++ )
            rl[FFT::Layout::slot(t + T * i)] = v[i];
        wave_lds_fence( );
        // the spectrum X[0 .. P] of the row into s_row[r] (the real row was read into v)
#pragma unroll
        for ( int i = 0; i < E; i++ ) {
            const int    k  = t + T * i;
            const float2 zp = rl[FFT::Layout::slot((P - k) & (P - 1))];
            const float2 a  = make_float2(0.5f * (v[i].x + zp.x), 0.5f * (v[i].y - zp.y)); // (Z + conj Zp) / 2
            const float2 b  = make_float2(0.5f * (v[i].x - zp.x), 0.5f * (v[i].y + zp.y)); // (Z - conj Zp) / 2
            const float2 c  = cmul(cmul(twm[N2 + T * i], wt), b);                          // W_N2^k b
            s_row[r][k]     = make_float2(a.x + c.y, a.y - c.x);                            // a - i c
        }
        if ( t == 0 )
            s_row[r][P] = make_float2(v[0].x - v[0].y, 0.f);
    }
    __syncthreads( );
    // ---- the RB rows' spectra, element k of the rows as RB / 2 16-byte stores
    auto* op = as_global(out) + om2.base(y0);
    for ( int g = threadIdx.x; g < nout * RP; g += NB ) {
        const int k = g / RP, p = g % RP;
        store_cpx_pair(op, ((long)k * om2.es + 2 * p) >> 1, s_row[2 * p][k], s_row[2 * p + 1][k]);
    }
}

} // namespace fastfft_amd
