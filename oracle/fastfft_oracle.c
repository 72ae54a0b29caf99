/*
 * FP64 CPU oracle for the FastFFT size-changing transform path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file: it is the checker that the GPU path is compared against
 * (tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py).
 *
 * It restates, in plain C and double precision, what the reference computes
 * on the FourierTransformer path:
 *
 *   FwdFFT          pad (or not) -> R2C along x -> C2C along y (2-D) / z, y (3-D)
 *                   reference: src/fastfft/FastFFT.cu:315-459 (Generic_Fwd),
 *                   kernels :1055-1166 (R2C_*_XY), :1348-1367 / :1433-1475 (C2C)
 *   InvFFT          C2C inverse along y (and z) -> C2R along x, optional trim
 *                   reference: FastFFT.cu:461-607 (Generic_Inv), :1641-1692 (C2R)
 *   FwdImageInvFFT  FwdFFT(template) -> image * conj(template) -> InvFFT
 *                   reference: FastFFT.cu:609-798, :20-56; functor
 *                   include/detail/functors.h:31-40 (CONJ_MUL), :42-80 (SCALE)
 *
 * Conventions fixed by the reference and kept here:
 *   - real buffers are FFTW padded: row pitch 2*(x/2+1) floats
 *     (include/FastFFT.h:363-404, ReturnPaddedMemorySize)
 *   - 2-D spectra are transposed, C[kx][ky] with ky contiguous and
 *     kx = 0..Nx/2 (FastFFT.cuh:750-767 store_r2c_transposed_xy;
 *     src/tests/helper_functions.cuh:149-153 debug stage 3 layout)
 *   - 3-D spectra are C[kz][kx][ky], ky contiguous
 *     (helper_functions.cuh:153-155 debug stage 3, Rank 3)
 *   - transforms are unnormalised: Inv(Fwd(x)) = Nx*Ny*Nz * x
 *     (src/tests/non_cuda_compilation_unit.cu:71-83)
 *   - C2R ignores the imaginary parts of the kx=0 and kx=Nx/2 planes
 *     (standard Hermitian C2R; the reference's cuFFTDx c2r and FFTW agree)
 *   - an output window of out_x*out_y(*out_z) samples starting at wrapped lag
 *     (ox, oy, oz) is kept.  Origin 0 is the reference's trimmed inverse
 *     (first out_y rows, first out_x columns, FastFFT.cu:1342, :1675);
 *     a negative origin (e.g. -256) gives a window centred on lag 0.
 *
 * Accuracy: iterative radix-2 FFT with twiddles from cos()/sin() in double;
 * relative error ~1e-15, far below the 2e-6 (FP32) / 2e-5 (split-BF16)
 * tolerances the GPU path is held to.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_OK 0
#define ORACLE_EINVAL 1
#define ORACLE_ENOMEM 2

static int is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }

static long wrap(long a, long n) {
    long r = a % n;
    return r < 0 ? r + n : r;
}

/* In-place unnormalised complex FFT on interleaved (re, im) doubles.
 * sign = -1: X[k] = sum x[n] exp(-2 pi i n k / n_pts); sign = +1: inverse. */
void oracle_fft1d(double* z, int n, int sign) {
    if ( n <= 1 )
        return;
    /* bit reversal */
    for ( int i = 1, j = 0; i < n; i++ ) {
        int bit = n >> 1;
        for ( ; j & bit; bit >>= 1 )
            j ^= bit;
        j ^= bit;
        if ( i < j ) {
            double tr = z[2 * i], ti = z[2 * i + 1];
            z[2 * i]     = z[2 * j];
            z[2 * i + 1] = z[2 * j + 1];
            z[2 * j]     = tr;
            z[2 * j + 1] = ti;
        }
    }
    for ( int len = 2; len <= n; len <<= 1 ) {
        int half = len >> 1;
        for ( int k = 0; k < half; k++ ) {
            double ang = sign * 2.0 * M_PI * (double)k / (double)len;
            double wr = cos(ang), wi = sin(ang);
            for ( int s = 0; s < n; s += len ) {
                double* a  = z + 2 * (s + k);
                double* b  = z + 2 * (s + k + half);
                double  br = b[0] * wr - b[1] * wi;
                double  bi = b[0] * wi + b[1] * wr;
                b[0]       = a[0] - br;
                b[1]       = a[1] - bi;
                a[0] += br;
                a[1] += bi;
            }
        }
    }
}

/* FFT along a strided line: gathers into a scratch line, transforms, scatters. */
static void fft_strided(double* base, long stride, int n, int sign, double* tmp) {
    for ( int i = 0; i < n; i++ ) {
        tmp[2 * i]     = base[2 * stride * i];
        tmp[2 * i + 1] = base[2 * stride * i + 1];
    }
    oracle_fft1d(tmp, n, sign);
    for ( int i = 0; i < n; i++ ) {
        base[2 * stride * i]     = tmp[2 * i];
        base[2 * stride * i + 1] = tmp[2 * i + 1];
    }
}

/*
 * 2-D forward.  in: real [in_y][pitch] floats holding an in_x * in_y image,
 * placed at the origin of an n_x * n_y zero frame (size increase, reference
 * FastFFT.cu:413-418).  When in_x > n_x (size decrease, Fourier cropping,
 * FastFFT.cu:420-424) the full in_x * in_y transform is taken and the first
 * n_x/2+1 kx and first n_y ky are kept (output-pruned DIT of
 * FastFFT.cu:1232-1261 keeps k < P).
 * spec: C[kx][ky], (n_x/2+1) rows of n_y complex (interleaved doubles).
 */
int oracle_fwd_2d(const float* in, int in_x, int in_y, int pitch, int n_x, int n_y, double pre_scale, double* spec) {
    if ( in_x <= 0 || in_y <= 0 || n_x <= 0 || n_y <= 0 || pitch < in_x )
        return ORACLE_EINVAL;
    int fx = in_x > n_x ? in_x : n_x; /* full transform size along x */
    int fy = in_y > n_y ? in_y : n_y;
    int wx = fx / 2 + 1;
    /* rows: full complex FFT of each (zero padded) real row, keep kx <= fx/2 */
    double* rows = (double*)calloc((size_t)2 * wx * fy, sizeof(double)); /* [kx][y] */
    if ( ! rows )
        return ORACLE_ENOMEM;
    int fail = 0;
#pragma omp parallel
    {
        double* line = (double*)malloc(sizeof(double) * 2 * (size_t)(fx > fy ? fx : fy));
        if ( ! line ) {
#pragma omp atomic write
            fail = 1;
        }
        else {
#pragma omp for schedule(static)
            for ( int y = 0; y < in_y; y++ ) {
                memset(line, 0, sizeof(double) * 2 * (size_t)fx);
                for ( int x = 0; x < in_x; x++ )
                    line[2 * x] = pre_scale * (double)in[(size_t)y * pitch + x];
                oracle_fft1d(line, fx, -1);
                for ( int k = 0; k < wx; k++ ) {
                    rows[2 * ((size_t)k * fy + y)]     = line[2 * k];
                    rows[2 * ((size_t)k * fy + y) + 1] = line[2 * k + 1];
                }
            }
#pragma omp for schedule(static)
            for ( int k = 0; k < wx; k++ )
                oracle_fft1d(rows + 2 * (size_t)k * fy, fy, -1);
            free(line);
        }
    }
    if ( fail ) {
        free(rows);
        return ORACLE_ENOMEM;
    }
    int ow = n_x / 2 + 1;
    for ( int k = 0; k < ow; k++ )
        for ( int ky = 0; ky < n_y; ky++ ) {
            spec[2 * ((size_t)k * n_y + ky)]     = rows[2 * ((size_t)k * fy + ky)];
            spec[2 * ((size_t)k * n_y + ky) + 1] = rows[2 * ((size_t)k * fy + ky) + 1];
        }
    free(rows);
    return ORACLE_OK;
}

/* C2R of one half-spectrum line (n/2+1 complex) to n reals, unnormalised. */
static void c2r_line(const double* half, int n, double* full, double* out) {
    int h = n / 2;
    for ( int k = 0; k <= h; k++ ) {
        full[2 * k]     = half[2 * k];
        full[2 * k + 1] = half[2 * k + 1];
    }
    /* Hermitian extension; imaginary parts of k=0 and k=n/2 drop out of Re() */
    for ( int k = h + 1; k < n; k++ ) {
        full[2 * k]     = half[2 * (n - k)];
        full[2 * k + 1] = -half[2 * (n - k) + 1];
    }
    oracle_fft1d(full, n, +1);
    for ( int i = 0; i < n; i++ )
        out[i] = full[2 * i];
}

/*
 * 2-D inverse.  spec: C[kx][ky], (n_x/2+1) rows of n_y complex doubles.
 * Computes the full unnormalised C2C(y)+C2R(x) inverse and keeps the
 * out_x * out_y window starting at wrapped lag (ox, oy).  out: [out_y][out_pitch].
 */
int oracle_inv_2d(const double* spec, int n_x, int n_y, int out_x, int out_y, int ox, int oy, double post_scale, double* out,
                  int out_pitch) {
    if ( n_x <= 0 || n_y <= 0 || out_x <= 0 || out_y <= 0 || out_x > n_x || out_y > n_y || out_pitch < out_x )
        return ORACLE_EINVAL;
    int     wx   = n_x / 2 + 1;
    double* cols = (double*)malloc(sizeof(double) * 2 * (size_t)wx * n_y);
    if ( ! cols )
        return ORACLE_ENOMEM;
    memcpy(cols, spec, sizeof(double) * 2 * (size_t)wx * n_y);
#pragma omp parallel for schedule(static)
    for ( int k = 0; k < wx; k++ )
        oracle_fft1d(cols + 2 * (size_t)k * n_y, n_y, +1);
    int fail = 0;
#pragma omp parallel
    {
        double* half = (double*)malloc(sizeof(double) * 2 * (size_t)wx);
        double* full = (double*)malloc(sizeof(double) * 2 * (size_t)n_x);
        double* row  = (double*)malloc(sizeof(double) * (size_t)n_x);
        if ( ! half || ! full || ! row ) {
#pragma omp atomic write
            fail = 1;
        }
        else {
#pragma omp for schedule(static)
            for ( int j = 0; j < out_y; j++ ) {
                long y = wrap((long)oy + j, n_y);
                for ( int k = 0; k < wx; k++ ) {
                    half[2 * k]     = cols[2 * ((size_t)k * n_y + y)];
                    half[2 * k + 1] = cols[2 * ((size_t)k * n_y + y) + 1];
                }
                c2r_line(half, n_x, full, row);
                for ( int i = 0; i < out_x; i++ )
                    out[(size_t)j * out_pitch + i] = post_scale * row[wrap((long)ox + i, n_x)];
            }
        }
        free(half);
        free(full);
        free(row);
    }
    free(cols);
    return fail ? ORACLE_ENOMEM : ORACLE_OK;
}

/*
 * Cross-correlation, FwdImageInvFFT with intra op CONJ_MUL or
 * CONJ_MUL_THEN_SCALE (include/detail/functors.h:31-60):
 *     Z = image_spectrum * conj(FwdFFT(pre_scale * template)) * intra_scale
 *     out = post_scale * window(InvFFT(Z))
 * The template (tx * ty, pitch tpitch floats) is zero padded to n_x * n_y.
 */
int oracle_xcorr_2d(const float* tmpl, int tx, int ty, int tpitch, const double* image_spec, int n_x, int n_y, double pre_scale,
                    double intra_scale, int out_x, int out_y, int ox, int oy, double post_scale, double* out, int out_pitch) {
    if ( tx > n_x || ty > n_y )
        return ORACLE_EINVAL;
    size_t  nc   = (size_t)(n_x / 2 + 1) * n_y;
    double* spec = (double*)malloc(sizeof(double) * 2 * nc);
    if ( ! spec )
        return ORACLE_ENOMEM;
    int rc = oracle_fwd_2d(tmpl, tx, ty, tpitch, n_x, n_y, pre_scale, spec);
    if ( rc ) {
        free(spec);
        return rc;
    }
#pragma omp parallel for schedule(static)
    for ( long i = 0; i < (long)nc; i++ ) {
        double tr = spec[2 * i], ti = spec[2 * i + 1];
        double gr = image_spec[2 * i], gi = image_spec[2 * i + 1];
        /* g * conj(t) */
        spec[2 * i]     = (gr * tr + gi * ti) * intra_scale;
        spec[2 * i + 1] = (gi * tr - gr * ti) * intra_scale;
    }
    rc = oracle_inv_2d(spec, n_x, n_y, out_x, out_y, ox, oy, post_scale, out, out_pitch);
    free(spec);
    return rc;
}

/*
 * 3-D forward.  in: real [in_z][in_y][pitch]; zero padded to n_x*n_y*n_z.
 * spec: C[kz][kx][ky] with ky contiguous, kx = 0..n_x/2.
 * Only size increase or no change (the reference aborts on 3-D decrease,
 * FastFFT.cu:449-453).
 */
int oracle_fwd_3d(const float* in, int in_x, int in_y, int in_z, int pitch, int n_x, int n_y, int n_z, double pre_scale,
                  double* spec) {
    if ( in_x > n_x || in_y > n_y || in_z > n_z || in_x <= 0 || in_y <= 0 || in_z <= 0 || pitch < in_x )
        return ORACLE_EINVAL;
    int    wx   = n_x / 2 + 1;
    size_t nall = (size_t)wx * n_y * n_z;
    memset(spec, 0, sizeof(double) * 2 * nall);
    int fail = 0;
#pragma omp parallel
    {
        int     nmax = n_x > n_y ? n_x : n_y;
        double* line = (double*)malloc(sizeof(double) * 2 * (size_t)(nmax > n_z ? nmax : n_z));
        if ( ! line ) {
#pragma omp atomic write
            fail = 1;
        }
        else {
            /* x pass on the nonzero rows, written into C[z][kx][y] */
#pragma omp for collapse(2) schedule(static)
            for ( int z = 0; z < in_z; z++ )
                for ( int y = 0; y < in_y; y++ ) {
                    memset(line, 0, sizeof(double) * 2 * (size_t)n_x);
                    for ( int x = 0; x < in_x; x++ )
                        line[2 * x] = pre_scale * (double)in[((size_t)z * in_y + y) * pitch + x];
                    oracle_fft1d(line, n_x, -1);
                    for ( int k = 0; k < wx; k++ ) {
                        size_t o          = ((size_t)z * wx + k) * n_y + y;
                        spec[2 * o]       = line[2 * k];
                        spec[2 * o + 1]   = line[2 * k + 1];
                    }
                }
            /* y pass (contiguous) on the in_z nonzero planes */
#pragma omp for collapse(2) schedule(static)
            for ( int z = 0; z < in_z; z++ )
                for ( int k = 0; k < wx; k++ )
                    oracle_fft1d(spec + 2 * (((size_t)z * wx + k) * n_y), n_y, -1);
            /* z pass, stride wx*n_y */
#pragma omp for collapse(2) schedule(static)
            for ( int k = 0; k < wx; k++ )
                for ( int ky = 0; ky < n_y; ky++ )
                    fft_strided(spec + 2 * ((size_t)k * n_y + ky), (long)wx * n_y, n_z, -1, line);
            free(line);
        }
    }
    return fail ? ORACLE_ENOMEM : ORACLE_OK;
}

/*
 * 3-D inverse.  spec: C[kz][kx][ky].  Full unnormalised inverse, then the
 * out_x*out_y*out_z window starting at wrapped (ox, oy, oz) is kept.
 * out: [out_z][out_y][out_pitch].
 */
int oracle_inv_3d(const double* spec, int n_x, int n_y, int n_z, int out_x, int out_y, int out_z, int ox, int oy, int oz,
                  double post_scale, double* out, int out_pitch) {
    if ( out_x > n_x || out_y > n_y || out_z > n_z || out_x <= 0 || out_y <= 0 || out_z <= 0 || out_pitch < out_x )
        return ORACLE_EINVAL;
    int     wx   = n_x / 2 + 1;
    size_t  nall = (size_t)wx * n_y * n_z;
    double* w    = (double*)malloc(sizeof(double) * 2 * nall);
    if ( ! w )
        return ORACLE_ENOMEM;
    memcpy(w, spec, sizeof(double) * 2 * nall);
    int fail = 0;
#pragma omp parallel
    {
        int     nmax = n_x > n_y ? n_x : n_y;
        nmax         = nmax > n_z ? nmax : n_z;
        double* line = (double*)malloc(sizeof(double) * 2 * (size_t)nmax);
        double* half = (double*)malloc(sizeof(double) * 2 * (size_t)wx);
        double* row  = (double*)malloc(sizeof(double) * (size_t)n_x);
        if ( ! line || ! half || ! row ) {
#pragma omp atomic write
            fail = 1;
        }
        else {
#pragma omp for collapse(2) schedule(static)
            for ( int k = 0; k < wx; k++ )
                for ( int ky = 0; ky < n_y; ky++ )
                    fft_strided(w + 2 * ((size_t)k * n_y + ky), (long)wx * n_y, n_z, +1, line);
#pragma omp for collapse(2) schedule(static)
            for ( int z = 0; z < n_z; z++ )
                for ( int k = 0; k < wx; k++ )
                    oracle_fft1d(w + 2 * (((size_t)z * wx + k) * n_y), n_y, +1);
#pragma omp for collapse(2) schedule(static)
            for ( int jz = 0; jz < out_z; jz++ )
                for ( int jy = 0; jy < out_y; jy++ ) {
                    long z = wrap((long)oz + jz, n_z), y = wrap((long)oy + jy, n_y);
                    for ( int k = 0; k < wx; k++ ) {
                        size_t o        = ((size_t)z * wx + k) * n_y + y;
                        half[2 * k]     = w[2 * o];
                        half[2 * k + 1] = w[2 * o + 1];
                    }
                    /* line doubles as the full-length complex scratch */
                    c2r_line(half, n_x, line, row);
                    for ( int i = 0; i < out_x; i++ )
                        out[((size_t)jz * out_y + jy) * out_pitch + i] = post_scale * row[wrap((long)ox + i, n_x)];
                }
        }
        free(line);
        free(half);
        free(row);
    }
    free(w);
    return fail ? ORACLE_ENOMEM : ORACLE_OK;
}

/* Position of the maximum of a [ny][pitch] real image (first hit in row-major
 * order wins ties).  Used for the peak-search epilogue of the correlation. */
void oracle_argmax_2d(const double* img, int nx, int ny, int pitch, int* ix, int* iy, double* val) {
    double best = -INFINITY;
    int    bx = 0, by = 0;
    for ( int y = 0; y < ny; y++ )
        for ( int x = 0; x < nx; x++ ) {
            double v = img[(size_t)y * pitch + x];
            if ( v > best ) {
                best = v;
                bx   = x;
                by   = y;
            }
        }
    *ix  = bx;
    *iy  = by;
    *val = best;
}

/* Element counts (in floats) of an FFTW-padded real buffer, as the reference's
 * ReturnPaddedMemorySize computes them (include/FastFFT.h:363-404). */
long oracle_padded_memory_size(int x, int y, int z) {
    long w = (x % 2 == 0) ? x / 2 + 1 : (x - 1) / 2 + 1;
    return 2 * w * (long)y * (long)z;
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads( );
#else
    return 1;
#endif
}

int oracle_is_pow2(long n) { return is_pow2(n); }
