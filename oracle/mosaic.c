/*
 * Oracle (test infrastructure only): mosaic of restored patches, Object::doMozaic
 * (src/lib/momfbd/object.cpp:1160-1207) with redux::image::mozaic / insertPatch
 * (include/redux/image/utils.hpp:389-502), apodizeInPlace (src/lib/image/utils.cpp:584-613),
 * math::apodize (src/lib/math/functions.cpp:39-50) and img_trim (utils.hpp:521-553).
 *
 * The reference sums the patches on a thread pool, each thread over the patches it pops,
 * so its summation order is not fixed; here the patches are summed in their given order
 * (one thread), products and sums rounded separately (no fused multiply-add).
 */
#include "mfbd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* functions.cpp:39-50 */
static void o_apodize(double* data, int n, double target)
{
    const double mean = 0.5 * (data[0] + target), amplitude = 0.5 * (data[0] - target);
    const double step = M_PI / (n - 1);
    for (int i = 1; i < n; ++i) data[i] = mean + amplitude * cos(i * step);
}

/* utils.cpp:584-613 applied to an all-ones array (utils.hpp:454-455) */
void o_blend_window(double* w, int nRows, int nCols, int blend, int margin)
{
    int rowBlend = blend, colBlend = blend;
    if (rowBlend + margin > nCols / 2) rowBlend = nCols / 2 - margin;
    if (colBlend + margin > nRows / 2) colBlend = nRows / 2 - margin;
    for (int i = 0; i < nRows * nCols; ++i) w[i] = 1.0;
    const int sz = (rowBlend > colBlend ? rowBlend : colBlend) + margin + 2;
    double* tmp = calloc(sz, sizeof(double));
    o_apodize(tmp + margin, rowBlend + 2, 1.0);
    for (int c = 0; c < nCols; ++c)
        for (int r = 0; r < rowBlend + margin; ++r) {
            w[r * nCols + c] *= tmp[r + 1];
            w[(nRows - r - 1) * nCols + c] *= tmp[r + 1];
        }
    memset(tmp, 0, sz * sizeof(double));
    o_apodize(tmp + margin, colBlend + 2, 1.0);
    for (int r = 0; r < nRows; ++r)
        for (int c = 0; c < colBlend + margin; ++c) {
            w[r * nCols + c] *= tmp[c + 1];
            w[r * nCols + nCols - c - 1] *= tmp[c + 1];
        }
    free(tmp);
}

/* object.cpp:1185-1205: img holds (maxPosY+s+1) x (maxPosX+s+1) floats on entry; on return
 * the first rows*cols of it hold the (trimmed, when clip) mosaic starting at (r0, c0). */
int o_mosaic(int n, int s, const float* patches, const int* posY, const int* posX, int clip,
             float* img, int* rows, int* cols, int* r0, int* c0)
{
    int maxY = 0, maxX = 0;
    for (int i = 0; i < n; ++i) {
        if (posY[i] > maxY) maxY = posY[i];
        if (posX[i] > maxX) maxX = posX[i];
    }
    const int R = maxY + s + 1, Cc = maxX + s + 1;
    const int margin = s / 8, blend = (s - 2 * margin) / 3;
    double* w = malloc(sizeof(double) * s * s);
    double* acc = calloc((size_t)R * Cc, sizeof(double));
    double* sum = calloc((size_t)R * Cc, sizeof(double));
    if (!w || !acc || !sum) { free(w); free(acc); free(sum); return 1; }
    o_blend_window(w, s, s, blend, margin);
    for (int i = 0; i < n; ++i) {                       /* insertPatch, utils.hpp:390-433 */
        const int pr = posY[i] + s >= R ? R - posY[i] : s;
        const int pc = posX[i] + s >= Cc ? Cc - posX[i] : s;
        for (int r = 0; r < pr; ++r)
            for (int c = 0; c < pc; ++c) {
                const size_t o = (size_t)(posY[i] + r) * Cc + posX[i] + c;
                sum[o] += w[r * s + c];
                const double t = w[r * s + c] * (double)patches[((size_t)i * s + r) * s + c];
                acc[o] += t;
            }
    }
    for (size_t o = 0; o < (size_t)R * Cc; ++o)            /* utils.hpp:486-491 */
        img[o] = (float)(sum[o] > 1E-6 ? acc[o] / sum[o] : acc[o]);
    int fr = 0, lr = R - 1, fc = 0, lc = Cc - 1;
    if (clip) {                                             /* img_trim, utils.hpp:521-553 */
        double* cs = calloc(Cc, sizeof(double));
        double* rs = calloc(R, sizeof(double));
        for (int c = 0; c < Cc; ++c)
            for (int r = 0; r < R; ++r) {
                const double v = fabs(img[(size_t)r * Cc + c]);
                cs[c] += v;
                rs[r] += v;
            }
        const float thr = 1E-15f;
        while (fc < lc && cs[fc] < thr) ++fc;
        while (lc > fc && cs[lc] < thr) --lc;
        while (fr < lr && rs[fr] < thr) ++fr;
        while (lr > fr && rs[lr] < thr) --lr;
        free(cs);
        free(rs);
    }
    const int oR = lr - fr + 1, oC = lc - fc + 1;
    for (int r = 0; r < oR; ++r)        /* rows move toward the front only: memmove is safe */
        memmove(img + (size_t)r * oC, img + (size_t)(r + fr) * Cc + fc, sizeof(float) * oC);
    *rows = oR;
    *cols = oC;
    *r0 = fr;
    *c0 = fc;
    free(w);
    free(acc);
    free(sum);
    return 0;
}
