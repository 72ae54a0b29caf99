/* Host-only driver of the FP64 oracle (oracle/fastfft_oracle.c) for the
 * AddressSanitizer / UndefinedBehaviorSanitizer build in tests/test_asan.py:
 * every 2-D and 3-D entry point on small padded, trimmed and windowed shapes,
 * with the round trip inv(fwd(x)) = N x checked, so that every buffer the
 * oracle indexes is exercised under the sanitizers. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

void oracle_fft1d(double* z, int n, int sign);
int  oracle_fwd_2d(const float* in, int in_x, int in_y, int pitch, int n_x, int n_y, double pre_scale, double* spec);
int  oracle_inv_2d(const double* spec, int n_x, int n_y, int out_x, int out_y, int ox, int oy, double post_scale, double* out,
                   int out_pitch);
int  oracle_xcorr_2d(const float* tmpl, int tx, int ty, int tpitch, const double* image_spec, int n_x, int n_y, double pre_scale,
                     double intra_scale, int out_x, int out_y, int ox, int oy, double post_scale, double* out, int out_pitch);
int  oracle_fwd_3d(const float* in, int in_x, int in_y, int in_z, int pitch, int n_x, int n_y, int n_z, double pre_scale,
                   double* spec);
int  oracle_inv_3d(const double* spec, int n_x, int n_y, int n_z, int out_x, int out_y, int out_z, int ox, int oy, int oz,
                   double post_scale, double* out, int out_pitch);
void oracle_argmax_2d(const double* img, int nx, int ny, int pitch, int* ix, int* iy, double* val);

int main(void) {
    int bad = 0;
    for ( int n = 4; n <= 64; n *= 2 ) {
        const int t = n / 2 + 1, pitch = 2 * (t / 2 + 1), opitch = 2 * (n / 2 + 1);
        float*    in = calloc((size_t)pitch * t, sizeof(float));
        for ( int i = 0; i < pitch * t; i++ )
            in[i] = (float)((i * 7919) % 17) - 8.f;
        double* spec = calloc((size_t)2 * (n / 2 + 1) * n, sizeof(double));
        double* out  = calloc((size_t)opitch * n, sizeof(double));
        bad |= oracle_fwd_2d(in, t, t, pitch, n, n, 1.0, spec) != 0;
        bad |= oracle_inv_2d(spec, n, n, n, n, 0, 0, 1.0 / ((double)n * n), out, opitch) != 0;
        for ( int y = 0; y < t; y++ )
            for ( int x = 0; x < t; x++ )
                if ( fabs(out[y * opitch + x] - in[y * pitch + x]) > 1e-9 ) {
                    bad = 1;
                }
        const int m = n / 2;
        double*   win = calloc((size_t)2 * (m / 2 + 1) * m, sizeof(double));
        bad |= oracle_xcorr_2d(in, t, t, pitch, spec, n, n, 1.0, 1.0, m, m, -m / 2, -m / 2, 1.0, win, 2 * (m / 2 + 1)) != 0;
        int    ix, iy;
        double v;
        oracle_argmax_2d(win, m, m, 2 * (m / 2 + 1), &ix, &iy, &v);
        bad |= ix < 0 || ix >= m || iy < 0 || iy >= m;
        free(in);
        free(spec);
        free(out);
        free(win);
    }
    {
        const int n = 8, t = 4, pitch = 2 * (t / 2 + 1), m = 4;
        float*    in = calloc((size_t)pitch * t * t, sizeof(float));
        for ( int i = 0; i < pitch * t * t; i++ )
            in[i] = (float)(i % 5);
        double* spec = calloc((size_t)2 * (n / 2 + 1) * n * n, sizeof(double));
        double* out  = calloc((size_t)2 * (m / 2 + 1) * m * m, sizeof(double));
        bad |= oracle_fwd_3d(in, t, t, t, pitch, n, n, n, 1.0, spec) != 0;
        bad |= oracle_inv_3d(spec, n, n, n, m, m, m, -m / 2, -m / 2, -m / 2, 1.0, out, 2 * (m / 2 + 1)) != 0;
        free(in);
        free(spec);
        free(out);
    }
    double z[2 * 16];
    for ( int i = 0; i < 32; i++ )
        z[i] = i % 3;
    oracle_fft1d(z, 16, -1);
    oracle_fft1d(z, 16, +1);
    printf("oracle driver: %s\n", bad ? "problems" : "ok");
    return bad;
}
