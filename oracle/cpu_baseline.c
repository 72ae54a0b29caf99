/*
 * cpu_baseline.c -- the CPU counterpart of the padded and trimmed
 * cross-correlation, timed beside the GPU path by the cpu_baseline leg of
 * bench.py.  BENCHMARK INFRASTRUCTURE: it is neither the product nor the
 * oracle, and nothing under fastfft_amd/ loads it.
 *
 * The reference's own CPU control is FFTW (src/fastfft/Image.cu:116-170:
 * fftwf_plan_dft_r2c / c2r on FFTW-padded rows).  This file runs the same
 * pipeline through the FFTW3 single-precision API of whichever library the
 * host has, resolved at run time with dlopen (no FFTW headers or libraries
 * are needed to build it):
 *   1. true FFTW: libfftw3f.so.3 with libfftw3f_threads.so.3;
 *   2. Intel MKL's FFTW3 interface in libmkl_rt (the directory is passed in).
 * Per template, after the micrograph's r2c transform (once, timed apart):
 *   zero-pad the t x t template into an n x n FFTW-padded buffer, r2c,
 *   multiply image * conj(template) (OpenMP over rows), c2r in place, then
 *   crop the m x m window of wrapped lags -m/2 .. m/2-1 and take its argmax
 *   (first maximum in row-major order).
 * All arithmetic is FP32, in the natural (untransposed) FFTW layout; the
 * outputs are unnormalised like FastFFT's.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef float              cpub_complex[2];
typedef struct cpub_plan_s* cpub_plan;

/* the FFTW3 API subset used here (public interface of fftw3.h) */
typedef cpub_plan (*fn_plan_r2c_2d)(int, int, float*, cpub_complex*, unsigned);
typedef cpub_plan (*fn_plan_c2r_2d)(int, int, cpub_complex*, float*, unsigned);
typedef void (*fn_exec_r2c)(const cpub_plan, float*, cpub_complex*);
typedef void (*fn_exec_c2r)(const cpub_plan, cpub_complex*, float*);
typedef void (*fn_destroy)(cpub_plan);
typedef int (*fn_init_threads)(void);
typedef void (*fn_plan_nthreads)(int);
typedef void* (*fn_malloc)(size_t);
typedef void (*fn_free)(void*);

#define CPUB_FFTW_MEASURE 0u
#define CPUB_FFTW_ESTIMATE (1u << 6)

static struct {
    int              kind; /* 0 none, 1 FFTW, 2 MKL FFTW3 interface */
    char             desc[256];
    fn_plan_r2c_2d   plan_r2c;
    fn_plan_c2r_2d   plan_c2r;
    fn_exec_r2c      exec_r2c;
    fn_exec_c2r      exec_c2r;
    fn_destroy       destroy;
    fn_init_threads  init_threads;
    fn_plan_nthreads plan_nthreads;
    fn_malloc        fmalloc;
    fn_free          ffree;
} g;

static int bind_api(void* h, void* hthreads) {
    g.plan_r2c      = (fn_plan_r2c_2d)dlsym(h, "fftwf_plan_dft_r2c_2d");
    g.plan_c2r      = (fn_plan_c2r_2d)dlsym(h, "fftwf_plan_dft_c2r_2d");
    g.exec_r2c      = (fn_exec_r2c)dlsym(h, "fftwf_execute_dft_r2c");
    g.exec_c2r      = (fn_exec_c2r)dlsym(h, "fftwf_execute_dft_c2r");
    g.destroy       = (fn_destroy)dlsym(h, "fftwf_destroy_plan");
    g.fmalloc       = (fn_malloc)dlsym(h, "fftwf_malloc");
    g.ffree         = (fn_free)dlsym(h, "fftwf_free");
    g.init_threads  = (fn_init_threads)dlsym(hthreads, "fftwf_init_threads");
    g.plan_nthreads = (fn_plan_nthreads)dlsym(hthreads, "fftwf_plan_with_nthreads");
    return g.plan_r2c && g.plan_c2r && g.exec_r2c && g.exec_c2r && g.destroy && g.fmalloc && g.ffree &&
           g.init_threads && g.plan_nthreads;
}

/* Resolves the library (FFTW first, then MKL in mkl_dir).  Returns the kind
 * (1 FFTW, 2 MKL, 0 none) and writes a description with the version. */
int cpub_open(const char* mkl_dir, char* desc, int desc_len) {
    if ( g.kind )
        goto done;
    void* h = dlopen("libfftw3f.so.3", RTLD_NOW | RTLD_GLOBAL);
    if ( h ) {
        void* ht = dlopen("libfftw3f_threads.so.3", RTLD_NOW | RTLD_GLOBAL);
        if ( ht && bind_api(h, ht) ) {
            const char* v = (const char*)dlsym(h, "fftwf_version");
            snprintf(g.desc, sizeof(g.desc), "FFTW (%s)", v ? v : "version unknown");
            g.kind = 1;
            goto done;
        }
    }
    if ( mkl_dir && mkl_dir[0] ) {
        char path[1024];
        snprintf(path, sizeof(path), "%s/libmkl_rt.so.1", mkl_dir);
        h = dlopen(path, RTLD_NOW | RTLD_GLOBAL);
        if ( ! h ) {
            snprintf(path, sizeof(path), "%s/libmkl_rt.so", mkl_dir);
            h = dlopen(path, RTLD_NOW | RTLD_GLOBAL);
        }
        if ( h && bind_api(h, h) ) {
            char ver[200] = "version unknown";
            void (*getv)(char*, int) = (void (*)(char*, int))dlsym(h, "mkl_get_version_string");
            if ( getv ) {
                /* MKL fills the buffer with blanks and does not terminate it */
                memset(ver, 0, sizeof(ver));
                getv(ver, (int)sizeof(ver) - 1);
                ver[sizeof(ver) - 1] = 0;
            }
            for ( int i = (int)strlen(ver) - 1; i >= 0 && (ver[i] == ' ' || ver[i] == 0); i-- )
                ver[i] = 0;
            snprintf(g.desc, sizeof(g.desc), "MKL FFTW3 API (%s)", ver);
            g.kind = 2;
        }
    }
done:
    if ( desc && desc_len > 0 )
        snprintf(desc, (size_t)desc_len, "%s", g.kind ? g.desc : "none");
    return g.kind;
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/*
 * Correlates the t x t template (row-major floats) with the n x n micrograph
 * (row-major floats): warmup untimed correlations, then at least min_corr and
 * until max_seconds have passed.  Results: mic_fwd_s (the micrograph's r2c),
 * corr_s (time of the timed correlations), n_done, the window of the last
 * correlation (m x m floats, row-major: out[y][x] = lag (x - m/2, y - m/2)) and
 * its argmax (x, y) in the window.  Returns 0 on success.
 */
int cpub_xcorr(const float* mic, const float* tmpl, int n, int t, int m, int nthreads, int warmup, int min_corr,
               double max_seconds, double* mic_fwd_s, double* corr_s, int* n_done, float* window, int* peak_xy) {
    if ( ! g.kind || n < 2 || (n & 1) || t < 1 || t > n || m < 1 || m > n || nthreads < 1 )
        return 1;
    const int    w     = n / 2 + 1;
    const int    pitch = 2 * w;
    const size_t nreal = (size_t)n * pitch;
    float*       img   = (float*)g.fmalloc(sizeof(float) * nreal);
    float*       buf   = (float*)g.fmalloc(sizeof(float) * nreal);
    if ( ! img || ! buf )
        return 2;
    g.init_threads( );
    g.plan_nthreads(nthreads);
    omp_set_num_threads(nthreads);
    const unsigned flags = g.kind == 1 ? CPUB_FFTW_MEASURE : CPUB_FFTW_ESTIMATE;
    /* plans first (MEASURE overwrites its arrays), then the data */
    cpub_plan fwd = g.plan_r2c(n, n, buf, (cpub_complex*)buf, flags);
    cpub_plan inv = g.plan_c2r(n, n, (cpub_complex*)buf, buf, flags);
    if ( ! fwd || ! inv )
        return 3;
    /* micrograph spectrum, once */
    memset(img, 0, sizeof(float) * nreal);
    for ( int y = 0; y < n; y++ )
        memcpy(img + (size_t)y * pitch, mic + (size_t)y * n, sizeof(float) * n);
    double t0 = now_s( );
    g.exec_r2c(fwd, img, (cpub_complex*)img);
    *mic_fwd_s = now_s( ) - t0;

    int    done = 0;
    double tc   = 0.0;
    for ( int it = 0;; it++ ) {
        const int timed = it >= warmup;
        if ( timed && done >= min_corr && tc >= max_seconds )
            break;
        double ts = now_s( );
        /* zero-padded template */
#pragma omp parallel for schedule(static)
        for ( int y = 0; y < n; y++ ) {
            float* row = buf + (size_t)y * pitch;
            if ( y < t ) {
                memcpy(row, tmpl + (size_t)y * t, sizeof(float) * t);
                memset(row + t, 0, sizeof(float) * (pitch - t));
            }
            else {
                memset(row, 0, sizeof(float) * pitch);
            }
        }
        g.exec_r2c(fwd, buf, (cpub_complex*)buf);
        /* image * conj(template) */
#pragma omp parallel for schedule(static)
        for ( int y = 0; y < n; y++ ) {
            const float* a = img + (size_t)y * pitch;
            float*       b = buf + (size_t)y * pitch;
            for ( int k = 0; k < w; k++ ) {
                const float ar = a[2 * k], ai = a[2 * k + 1], br = b[2 * k], bi = b[2 * k + 1];
                b[2 * k]     = ar * br + ai * bi;
                b[2 * k + 1] = ai * br - ar * bi;
            }
        }
        g.exec_c2r(inv, (cpub_complex*)buf, buf);
        /* window of wrapped lags and its first maximum */
        float best  = -3.402823466e38f;
        int   bx    = 0, by = 0;
        for ( int y = 0; y < m; y++ ) {
            const int    sy  = ((y - m / 2) % n + n) % n;
            const float* row = buf + (size_t)sy * pitch;
            for ( int x = 0; x < m; x++ ) {
                const float v = row[((x - m / 2) % n + n) % n];
                window[(size_t)y * m + x] = v;
                if ( v > best ) {
                    best = v;
                    bx   = x;
                    by   = y;
                }
            }
        }
        peak_xy[0] = bx;
        peak_xy[1] = by;
        if ( timed ) {
            tc += now_s( ) - ts;
            done++;
        }
    }
    *corr_s = tc;
    *n_done = done;
    g.destroy(fwd);
    g.destroy(inv);
    g.ffree(img);
    g.ffree(buf);
    return 0;
}
