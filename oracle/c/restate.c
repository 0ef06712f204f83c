/* C/OpenMP restatement of the reference's Numba CPU kernels.
 *
 * TEST INFRASTRUCTURE ONLY (the CPU baseline and a parity checker; the
 * product never links it).
 *
 * rs_gather follows src/pyfourier/_subroutines/_grid/_cart2noncart/_3D.py:13-51
 *   (prange over frames*batch*points, (z, y, x) tap loops, table-driven);
 * rs_spread follows src/pyfourier/_subroutines/_grid/_noncart2cart/_3D.py:13-51
 *   (prange over frames*batch, serial over points: race free because each
 *   thread owns one (frame, batch) grid).
 * 1-D and 2-D problems are passed as 3-D with unit leading axes (width 1,
 * index 0, value 1).  Complex data are interleaved float32 pairs; tables are
 * per axis (P, W) int32 / float32 with P = frames*points.
 */
#include <stddef.h>

void rs_gather(int nframes, int batch, long npts, const float* grid, int gz, int gy, int gx,
               const int* zi, const int* yi, const int* xi, const float* zv, const float* yv,
               const float* xv, int wz, int wy, int wx, float* out) {
    const long total = (long)nframes * batch * npts;
    const long gsz = (long)gz * gy * gx;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < total; ++i) {
        long frame = i / ((long)batch * npts);
        long tmp = i % ((long)batch * npts);
        long b = tmp / npts;
        long p = tmp % npts;
        long row = frame * npts + p;
        const float* g = grid + 2 * (frame * batch + b) * gsz;
        float re = 0.f, im = 0.f;
        for (int a = 0; a < wz; ++a) {
            long idz = zi[row * wz + a];
            float vz = zv[row * wz + a];
            for (int c = 0; c < wy; ++c) {
                long idy = yi[row * wy + c];
                float vy = vz * yv[row * wy + c];
                for (int e = 0; e < wx; ++e) {
                    long idx = xi[row * wx + e];
                    float v = vy * xv[row * wx + e];
                    long o = 2 * ((idz * gy + idy) * gx + idx);
                    re += v * g[o];
                    im += v * g[o + 1];
                }
            }
        }
        out[2 * i] += re;
        out[2 * i + 1] += im;
    }
}

void rs_spread(int nframes, int batch, long npts, const float* ksp, int gz, int gy, int gx,
               const int* zi, const int* yi, const int* xi, const float* zv, const float* yv,
               const float* xv, int wz, int wy, int wx, float* grid) {
    const long gsz = (long)gz * gy * gx;
#pragma omp parallel for schedule(static)
    for (long fb = 0; fb < (long)nframes * batch; ++fb) {
        long frame = fb / batch;
        float* g = grid + 2 * fb * gsz;
        const float* k = ksp + 2 * fb * npts;
        for (long p = 0; p < npts; ++p) {
            long row = frame * npts + p;
            float kr = k[2 * p], ki = k[2 * p + 1];
            for (int a = 0; a < wz; ++a) {
                long idz = zi[row * wz + a];
                float vz = zv[row * wz + a];
                for (int c = 0; c < wy; ++c) {
                    long idy = yi[row * wy + c];
                    float vy = vz * yv[row * wy + c];
                    for (int e = 0; e < wx; ++e) {
                        long idx = xi[row * wx + e];
                        float v = vy * xv[row * wx + e];
                        long o = 2 * ((idz * gy + idy) * gx + idx);
                        g[o] += v * kr;
                        g[o + 1] += v * ki;
                    }
                }
            }
        }
    }
}

#include <omp.h>
void rs_set_threads(int n) { omp_set_num_threads(n); }
int rs_get_threads(void) { return omp_get_max_threads(); }
