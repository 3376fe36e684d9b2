/*
 * Oracle (test infrastructure only): Fourier helpers restating
 * redux::image::FourierTransform (include/redux/image/fouriertransform.hpp,
 * src/lib/image/fouriertransform.cpp) with FFTW's conventions.
 */
#include "mfbd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* 1-D unnormalised DFT of length n, sign -1 = FFTW_FORWARD, +1 = FFTW_BACKWARD.
 * Powers of two use an iterative radix-2 Cooley-Tukey with exact-angle twiddles;
 * other lengths (e.g. an 88-pixel noise patch) use the direct O(n^2) sum. */
static void fft1d(ocplx* x, int n, int sign, ocplx* tmp)
{
    if (n <= 1) return;
    if (is_pow2(n)) {
        for (int i = 1, j = 0; i < n; ++i) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) { ocplx t = x[i]; x[i] = x[j]; x[j] = t; }
        }
        /* twiddles of the last (n, sign) per thread, cached: identical values, computed
         * once per transform size instead of once per line */
        static _Thread_local ocplx* tw[2] = {NULL, NULL};
        static _Thread_local int twn[2] = {0, 0};
        const int slot = sign < 0 ? 0 : 1;
        if (twn[slot] != n) {
            free(tw[slot]);
            tw[slot] = (ocplx*)malloc(sizeof(ocplx) * (size_t)(n / 2));
            for (int k = 0; k < n / 2; ++k) {
                double a = sign * 2.0 * M_PI * (double)k / (double)n;
                tw[slot][k] = cos(a) + I * sin(a);
            }
            twn[slot] = n;
        }
        const ocplx* twk = tw[slot];
        (void)tmp;
        for (int len = 2; len <= n; len <<= 1) {
            int half = len >> 1, step = n / len;
            for (int i = 0; i < n; i += len) {
                for (int j = 0; j < half; ++j) {
                    ocplx u = x[i + j];
                    ocplx v = x[i + j + half] * twk[j * step];
                    x[i + j] = u + v;
                    x[i + j + half] = u - v;
                }
            }
        }
        return;
    }
    ocplx* w = tmp;          /* n twiddles */
    ocplx* out = tmp + n;    /* n outputs   */
    for (int k = 0; k < n; ++k) {
        double a = sign * 2.0 * M_PI * (double)k / (double)n;
        w[k] = cos(a) + I * sin(a);
    }
    for (int k = 0; k < n; ++k) {
        ocplx acc = 0;
        for (int j = 0; j < n; ++j) acc += x[j] * w[((long)j * k) % n];
        out[k] = acc;
    }
    memcpy(x, out, sizeof(ocplx) * n);
}

/* 2-D in-place, unnormalised (fftw_plan_dft_2d, fouriertransform.cpp:85-86). */
void o_fft2(ocplx* a, int ny, int nx, int sign)
{
    int m = ny > nx ? ny : nx;
    ocplx* line = (ocplx*)malloc(sizeof(ocplx) * m);
    ocplx* tmp = (ocplx*)malloc(sizeof(ocplx) * 2 * m);
    for (int y = 0; y < ny; ++y) fft1d(a + (long)y * nx, nx, sign, tmp);
    for (int x = 0; x < nx; ++x) {
        for (int y = 0; y < ny; ++y) line[y] = a[(long)y * nx + x];
        fft1d(line, ny, sign, tmp);
        for (int y = 0; y < ny; ++y) a[(long)y * nx + x] = line[y];
    }
    free(line);
    free(tmp);
}

/* FourierTransform::reorderInto (fouriertransform.hpp:111-171): quadrant swap of a
 * (small) array into a (larger) one; equal sizes give fftshift for even sides. */
#define REORDER_BODY(T)                                                                  \
    if (!ony && !onx) { ony = iny; onx = inx; }                                          \
    if (iny == ony && inx == onx) center = 0;                                            \
    const long inMidY = iny / 2, inMidX = inx / 2, oddY = iny % 2, oddX = inx % 2;       \
    const long diffY = ony - iny, diffX = onx - inx, diffYh = diffY / 2, diffXh = diffX / 2; \
    const long wDx = center ? diffXh : 0;                                                \
    const long eDx = center ? (diffXh - diffX) : 0;                                      \
    const long sDy = center ? diffYh * onx : 0;                                          \
    const long nDy = center ? (diffYh - diffY) * onx : 0;                                \
    const T* seIn = in + inMidX + oddX;                                                  \
    const T* nwIn = in + (inMidY + oddY) * inx;                                          \
    const T* neIn = seIn + (inMidY + oddY) * inx;                                        \
    T* swOut = out + sDy + wDx;                                                          \
    T* seOut = out + onx - inMidX - oddX + sDy + eDx;                                    \
    T* nwOut = out + (ony - inMidY - oddY) * onx + nDy + wDx;                            \
    T* neOut = out + (ony - inMidY - oddY + 1) * onx - inMidX - oddX + nDy + eDx;        \
    for (long y = 0; y < inMidY; ++y) {                                                  \
        memcpy(swOut + y * onx, neIn + y * inx, sizeof(T) * inMidX);                     \
        memcpy(seOut + y * onx, nwIn + y * inx, sizeof(T) * (inMidX + oddX));            \
    }                                                                                    \
    for (long y = inMidY + oddY - 1; y >= 0; --y) {                                      \
        memcpy(neOut + y * onx, in + y * inx, sizeof(T) * (inMidX + oddX));              \
        memcpy(nwOut + y * onx, seIn + y * inx, sizeof(T) * inMidX);                     \
    }

void o_reorder_into_c(const ocplx* in, int iny, int inx, ocplx* out, int ony, int onx, int center)
{
    REORDER_BODY(ocplx)
}

void o_reorder_into_d(const double* in, int iny, int inx, double* out, int ony, int onx, int center)
{
    REORDER_BODY(double)
}

/* FourierTransform::reorder (fouriertransform.hpp:174-180) */
void o_reorder_c(ocplx* a, int ny, int nx)
{
    ocplx* buf = (ocplx*)malloc(sizeof(ocplx) * (long)ny * nx);
    o_reorder_into_c(a, ny, nx, buf, 0, 0, 0);
    memcpy(a, buf, sizeof(ocplx) * (long)ny * nx);
    free(buf);
}

void o_reorder_d(double* a, int ny, int nx)
{
    double* buf = (double*)malloc(sizeof(double) * (long)ny * nx);
    o_reorder_into_d(a, ny, nx, buf, 0, 0, 0);
    memcpy(a, buf, sizeof(double) * (long)ny * nx);
    free(buf);
}

/* FourierTransform::autocorrelate for complex data (fouriertransform.hpp:280-286):
 * out = reorder( IFFT(|FFT(in)|^2) / (ny*nx) ). */
void o_autocorrelate_c(ocplx* a, int ny, int nx, int center)
{
    long n = (long)ny * nx;
    o_fft2(a, ny, nx, -1);
    for (long i = 0; i < n; ++i) a[i] = creal(a[i]) * creal(a[i]) + cimag(a[i]) * cimag(a[i]);
    o_fft2(a, ny, nx, +1);
    const double nrm = 1.0 / (double)n;
    for (long i = 0; i < n; ++i) a[i] *= nrm;
    if (center) o_reorder_c(a, ny, nx);
}

/* static FourierTransform::autocorrelate(Array<double>&) (fouriertransform.hpp:291-297):
 * real part of the same, used by Pupil::generateSupport (pupil.cpp:268). */
void o_autocorrelate_d(double* a, int ny, int nx, int center)
{
    long n = (long)ny * nx;
    ocplx* c = (ocplx*)malloc(sizeof(ocplx) * n);
    for (long i = 0; i < n; ++i) c[i] = a[i];
    o_fft2(c, ny, nx, -1);
    for (long i = 0; i < n; ++i) c[i] = creal(c[i]) * creal(c[i]) + cimag(c[i]) * cimag(c[i]);
    o_fft2(c, ny, nx, +1);
    const double nrm = 1.0 / (double)n;
    for (long i = 0; i < n; ++i) a[i] = creal(c[i]) * nrm;
    free(c);
    if (center) o_reorder_d(a, ny, nx);
}

/* FourierTransform::noise (fouriertransform.cpp:448-495) for a FULLCOMPLEX,
 * un-reordered transform of an ny x nx real image. */
double o_noise(const ocplx* ft, int ny, int nx, int mask, double limit)
{
    double noise_power = 0.0;
    long cnt = 0;
    long inPixels = (long)ny * nx;
    if (mask < 0) mask = (ny > nx ? ny : nx) / 6 + 1;
    double limits = limit * limit;
    if (limit < 0) limits = inPixels / 4.0;
    int endX = nx, endY = ny;
    if (mask > 0) {
        endY = ny - mask + 1;
        endX = nx - mask + 1;
    }
    for (int x = mask; x < endX; ++x) {
        double xx = (x < nx - x) ? x : nx - x;
        xx *= xx;
        for (int y = mask; y < endY; ++y) {
            double yy = (y < ny - y) ? y : ny - y;
            if ((xx + yy * yy) < limits) continue;
            ocplx v = ft[(long)y * nx + x];
            noise_power += creal(v) * creal(v) + cimag(v) * cimag(v);
            cnt++;
        }
    }
    if (cnt && inPixels) noise_power = sqrt(noise_power / ((double)cnt * inPixels));
    return noise_power;
}
