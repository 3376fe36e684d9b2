/*
 * Oracle (test infrastructure only): FP64 restatement of the MOMFBD patch
 * solver.  Follows, function by function:
 *   src/lib/momfbd/solver.cpp      Solver::run/metric/calcPQ/gradient/shiftAndInit
 *   src/lib/momfbd/subimage.cpp    SubImage::initialize/calcPhi/calcPFOTF/addPQ/
 *                                  calcVogelWeight/gradientVogel/
 *                                  gradientFiniteDifference/metricChange/adjustShifts
 *   src/lib/momfbd/object.cpp      initPQ/addRegGamma/addToFT/calcHelpers/
 *                                  calcMetric/reInitialize/restorePatch
 *   src/lib/momfbd/channel.cpp     initPatch (phi_channel, otfNormalization)
 *   src/lib/util/gsl/conjugate_gradient.cpp   conjugate_rdx set/iterate/take_step
 *   src/lib/math/helpers.cpp       bracket, brent
 *   src/lib/image/utils.cpp        apodizeInPlace, smooth, ScharmerFilter,
 *                                  connectedRegion;  src/lib/math/functions.cpp apodize
 *   src/lib/util/arraystats.cpp    getMinMaxMean/getRmsStddev
 * One deliberate omission: gsl_multimin_fdfminimizer_set evaluates f and the
 * gradient, but conjugate_rdx_iterate overwrites that gradient before reading it
 * (conjugate_gradient.cpp:83,144), so only f is evaluated here.
 */
#include "mfbd_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define GOLDEN 1.6180339887498949025   /* constants.hpp:26, (1+sqrt 5)/2 */
#define TINY 1.0E-32                   /* helpers.cpp:56 */
#define GSL_SUCCESS 0
#define GSL_FAILURE -1
#define GSL_ENOPROG 27

struct OState {
    const OProblem* pr;
    int s, P, N, K, P2, N2, s2;
    /* per object */
    double* area; int* nPS; int** ps; int** pio; int* nOS; int** os;
    double* regGamma; double* ftSum; ocplx* Pc; double* Qc; ocplx* PQ; double* PS; double* QS;
    double* curMetric; int* nImgObj;
    /* per channel */
    double* phiCh; double* otfNorm;
    /* per image */
    double* phi; ocplx* PF; ocplx* OTF; ocplx* imgFT; double* vogel;
    int* curShift; double* tiltOff; double* noise; double* stddev; double* mean;
    double* window; double* noiseWindow;
    /* solver */
    double* alpha; double* alphaOffset; double* gradAlpha; double* regW;
    int* enabled; int gradMethod; /* 1 = finite difference, 2 = Vogel */
    int nF, nDF;
    unsigned long long noiseState;   /* metricNoise stream */
    /* scratch */
    ocplx* tc1; ocplx* tc2; ocplx* tc3; double* td1; double* td2;
};

static int img_obj(const OState* st, int i) { return st->pr->chObj[st->pr->imgCh[i]]; }
static double sqr(double x) { return x * x; }
static double cnorm(ocplx z) { return creal(z) * creal(z) + cimag(z) * cimag(z); }

/* ---------- utilities (arraystats.cpp:17-81, functions.cpp:39-50, utils.cpp:583-613) ---------- */

static double stats_mean(const double* d, long n)
{
    double sum = 0; long count = 0;
    for (const double* p = d + n; p-- > d;) {
        double t = *p;
        if (isfinite(t)) { sum += t; count++; }
    }
    return count ? sum / count : sum;
}

static double stats_stddev(const double* d, long n, double mean)
{
    double rms = 0; long count = 0;
    for (const double* p = d + n; p-- > d;) {
        double t = *p;
        if (t == t) { rms += t * t; count++; }
    }
    if (!count) return 0;
    rms /= count;
    return sqrt(rms - mean * mean);
}

static void apodize1(double* data, int n, double target)
{
    double mean = 0.5 * (*data + target);
    double amplitude = 0.5 * (*data - target);
    double step = M_PI / (n - 1);
    for (int i = 1; i < n; ++i) data[i] = mean + amplitude * cos(i * step);
}

static void apodize_in_place(double* data, int nRows, int nCols, int blend)
{
    int rowBlend = blend, colBlend = blend;
    if (rowBlend > nCols / 2) rowBlend = nCols / 2;
    if (colBlend > nRows / 2) colBlend = nRows / 2;
    int sz = (rowBlend > colBlend ? rowBlend : colBlend) + 2;
    double* tmp = (double*)calloc(sz, sizeof(double));
    apodize1(tmp, rowBlend + 2, 1.0);
    for (int c = 0; c < nCols; ++c)
        for (int r = 0; r < rowBlend; ++r) {
            data[r * nCols + c] *= tmp[r + 1];
            data[(nRows - r - 1) * nCols + c] *= tmp[r + 1];
        }
    memset(tmp, 0, sz * sizeof(double));
    apodize1(tmp, colBlend + 2, 1.0);
    for (int rr = 0; rr < nRows; ++rr)
        for (int c = 0; c < colBlend; ++c) {
            data[rr * nCols + c] *= tmp[c + 1];
            data[rr * nCols + nCols - c - 1] *= tmp[c + 1];
        }
    free(tmp);
}

/* ---------- state construction ---------- */

OState* o_state_new(const OProblem* pr)
{
    OState* st = (OState*)calloc(1, sizeof(OState));
    st->pr = pr;
    st->noiseState = pr->noiseSeed;
    st->s = pr->s; st->P = pr->P; st->N = pr->N; st->K = pr->K;
    st->P2 = pr->P * pr->P; st->N2 = pr->N * pr->N; st->s2 = pr->s * pr->s;
    int nO = pr->nObj, nC = pr->nCh, nI = pr->nImg, P = pr->P, N = pr->N, P2 = st->P2, N2 = st->N2;

    st->area = calloc(nO, sizeof(double));
    st->nPS = calloc(nO, sizeof(int)); st->ps = calloc(nO, sizeof(int*)); st->pio = calloc(nO, sizeof(int*));
    st->nOS = calloc(nO, sizeof(int)); st->os = calloc(nO, sizeof(int*));
    double* otf = malloc(sizeof(double) * N2);
    for (int o = 0; o < nO; ++o) {
        /* Pupil::generateSupport (pupil.cpp:245-277), threshold 1e-9 */
        const double* pup = pr->pupil + (long)o * P2;
        st->ps[o] = malloc(sizeof(int) * P2); st->pio[o] = malloc(sizeof(int) * P2);
        memset(otf, 0, sizeof(double) * N2);
        double area = 0; int n = 0;
        for (int y = 0; y < P; ++y)
            for (int x = 0; x < P; ++x) {
                int cnt = y * P + x;
                double v = pup[cnt];
                otf[y * N + x] = v;
                if (v > 1e-9) {
                    st->ps[o][n] = cnt;
                    st->pio[o][n] = y * N + x + (P / 2) + (P / 2) * N;
                    n++;
                    area += v;
                }
            }
        st->nPS[o] = n; st->area[o] = area;
        o_autocorrelate_d(otf, N, N, 1);
        st->os[o] = malloc(sizeof(int) * N2);
        n = 0;
        for (int i = 0; i < N2; ++i) if (fabs(otf[i]) > 1e-9) st->os[o][n++] = i;
        st->nOS[o] = n;
    }
    free(otf);

    st->regGamma = calloc(nO, sizeof(double));
    st->ftSum = calloc((long)nO * N2, sizeof(double));
    st->Pc = calloc((long)nO * N2, sizeof(ocplx)); st->Qc = calloc((long)nO * N2, sizeof(double));
    st->PQ = calloc((long)nO * N2, sizeof(ocplx)); st->PS = calloc((long)nO * N2, sizeof(double));
    st->QS = calloc((long)nO * N2, sizeof(double));
    st->curMetric = calloc(nO, sizeof(double));
    st->nImgObj = calloc(nO, sizeof(int));
    for (int i = 0; i < nI; ++i) st->nImgObj[img_obj(st, i)]++;

    st->phiCh = calloc((long)nC * P2, sizeof(double));
    st->otfNorm = calloc(nC, sizeof(double));
    st->phi = calloc((long)nI * P2, sizeof(double));
    st->PF = calloc((long)nI * P2, sizeof(ocplx));
    st->OTF = calloc((long)nI * N2, sizeof(ocplx));
    st->imgFT = calloc((long)nI * N2, sizeof(ocplx));
    st->vogel = calloc((long)nI * P2, sizeof(double));
    st->curShift = calloc((long)nI * 2, sizeof(int));
    st->tiltOff = calloc((long)nI * 2, sizeof(double));
    st->noise = calloc(nI, sizeof(double)); st->stddev = calloc(nI, sizeof(double));
    st->mean = calloc(nI, sizeof(double));

    int m = N > pr->s ? N : pr->s;
    st->tc1 = calloc((long)m * m, sizeof(ocplx)); st->tc2 = calloc((long)m * m, sizeof(ocplx));
    st->tc3 = calloc((long)m * m, sizeof(ocplx));
    st->td1 = calloc((long)m * m, sizeof(double)); st->td2 = calloc((long)m * m, sizeof(double));

    /* Solver::init windows (solver.cpp:112-126) */
    st->window = malloc(sizeof(double) * st->s2); st->noiseWindow = malloc(sizeof(double) * st->s2);
    for (int i = 0; i < st->s2; ++i) st->window[i] = st->noiseWindow[i] = 1.0;
    int apod = pr->apodSize < 0 ? pr->s / 8 : pr->apodSize;
    if (apod) {
        apodize_in_place(st->window, pr->s, pr->s, apod);
        apodize_in_place(st->noiseWindow, pr->s, pr->s, apod / 2);
    }

    st->alpha = calloc(pr->nPar, sizeof(double)); st->alphaOffset = calloc(pr->nPar, sizeof(double));
    st->gradAlpha = calloc(pr->nPar, sizeof(double)); st->regW = calloc(pr->nPar, sizeof(double));
    st->enabled = calloc(pr->K, sizeof(int));
    st->gradMethod = 1;
    return st;
}

void o_state_free(OState* st)
{
    if (!st) return;
    for (int o = 0; o < st->pr->nObj; ++o) { free(st->ps[o]); free(st->pio[o]); free(st->os[o]); }
    free(st->area); free(st->nPS); free(st->ps); free(st->pio); free(st->nOS); free(st->os);
    free(st->regGamma); free(st->ftSum); free(st->Pc); free(st->Qc); free(st->PQ); free(st->PS); free(st->QS);
    free(st->curMetric); free(st->nImgObj); free(st->phiCh); free(st->otfNorm);
    free(st->phi); free(st->PF); free(st->OTF); free(st->imgFT); free(st->vogel);
    free(st->curShift); free(st->tiltOff); free(st->noise); free(st->stddev); free(st->mean);
    free(st->tc1); free(st->tc2); free(st->tc3); free(st->td1); free(st->td2);
    free(st->window); free(st->noiseWindow);
    free(st->alpha); free(st->alphaOffset); free(st->gradAlpha); free(st->regW); free(st->enabled);
    free(st);
}

/* ---------- per-image physics ---------- */

/* SubImage::calcPhi / addToPhi (subimage.cpp:509-544) */
static void calc_phi(OState* st, int i, const double* a, double* phiOut)
{
    const OProblem* pr = st->pr;
    int o = img_obj(st, i), c = pr->imgCh[i];
    memcpy(phiOut, st->phiCh + (long)c * st->P2, sizeof(double) * st->P2);
    for (int mi = 0; mi < st->K; ++mi) {
        double al = a[mi];
        if (mi == pr->tilt[2 * o]) al += st->tiltOff[2 * i];
        else if (mi == pr->tilt[2 * o + 1]) al += st->tiltOff[2 * i + 1];
        if (fabs(al) > 0.0) {
            const double* mode = pr->modes + ((long)o * st->K + mi) * st->P2;
            for (int p = 0; p < st->P2; ++p) phiOut[p] = phiOut[p] + al * mode[p];
        }
    }
}

/* SubImage::calcPFOTF (subimage.cpp:586-612) */
static void calc_pfotf(OState* st, int i)
{
    const OProblem* pr = st->pr;
    int o = img_obj(st, i), c = pr->imgCh[i];
    ocplx* pf = st->PF + (long)i * st->P2;
    ocplx* otf = st->OTF + (long)i * st->N2;
    const double* phi = st->phi + (long)i * st->P2;
    const double* pup = pr->pupil + (long)o * st->P2;
    memset(pf, 0, sizeof(ocplx) * st->P2);
    memset(otf, 0, sizeof(ocplx) * st->N2);
    for (int k = 0; k < st->nPS[o]; ++k) {
        int p = st->ps[o][k];
        pf[p] = pup[p] * cos(phi[p]) + I * (pup[p] * sin(phi[p]));
        otf[st->pio[o][k]] = st->otfNorm[c] * pf[p];
    }
    o_autocorrelate_c(otf, st->N, st->N, 1);
}

/* SubImage::calcOTF(otf, phi) (subimage.cpp:567-583) */
static void calc_otf(OState* st, int i, ocplx* otf, const double* phi)
{
    const OProblem* pr = st->pr;
    int o = img_obj(st, i), c = pr->imgCh[i];
    const double* pup = pr->pupil + (long)o * st->P2;
    memset(otf, 0, sizeof(ocplx) * st->N2);
    for (int k = 0; k < st->nPS[o]; ++k) {
        int p = st->ps[o][k];
        double mag = pup[p] * st->otfNorm[c];
        otf[st->pio[o][k]] = mag * cos(phi[p]) + I * (mag * sin(phi[p]));
    }
    o_autocorrelate_c(otf, st->N, st->N, 1);
}

/* Channel::initPatch: phi_channel and otfNormalization (channel.cpp:908-952) */
static void init_channels(OState* st)
{
    const OProblem* pr = st->pr;
    for (int c = 0; c < pr->nCh; ++c) {
        int o = pr->chObj[c];
        double* phi = st->phiCh + (long)c * st->P2;
        memcpy(phi, pr->phiFixed + (long)c * st->P2, sizeof(double) * st->P2);
        const double* J = pr->jac + 8L * o;  /* J_T_xy[0..3], J_xy_T[4..7] */
        double rx = pr->resOffset[2 * c], ry = pr->resOffset[2 * c + 1];
        /* shiftToAlpha (modes.hpp:130-133): y = sx*Jxy[1][0]+sy*Jxy[1][1], x = sx*Jxy[0][0]+sy*Jxy[0][1] */
        double ax = rx * J[4] + ry * J[5];
        double ay = rx * J[6] + ry * J[7];
        int tx = pr->tilt[2 * o], ty = pr->tilt[2 * o + 1];
        if (tx >= 0 && fabs(rx) > 0) {
            const double* mode = pr->modes + ((long)o * st->K + tx) * st->P2;
            for (int p = 0; p < st->P2; ++p) phi[p] = phi[p] + (-ax) * mode[p];
        }
        if (ty >= 0 && fabs(ry) > 0) {
            const double* mode = pr->modes + ((long)o * st->K + ty) * st->P2;
            for (int p = 0; p < st->P2; ++p) phi[p] = phi[p] + (-ay) * mode[p];
        }
        /* measure otfNormalization on the first image of the channel with phi = 0 */
        int first = -1;
        for (int i = 0; i < pr->nImg; ++i) if (pr->imgCh[i] == c) { first = i; break; }
        st->otfNorm[c] = 1.0;
        if (first >= 0) {
            memset(st->phi + (long)first * st->P2, 0, sizeof(double) * st->P2);
            calc_pfotf(st, first);
            const ocplx* otf = st->OTF + (long)first * st->N2;
            double mx = 0;
            for (int k = 0; k < st->N2; ++k) { double a = cabs(otf[k]); if (a > mx) mx = a; }
            st->otfNorm[c] = 1.0 / sqrt(mx);
        }
    }
}

/* SubImage::getWindowedImg + initialize (subimage.cpp:106-209), no fitted plane */
static void initialize_img(OState* st, int i, int doReset)
{
    const OProblem* pr = st->pr;
    int o = img_obj(st, i), c = pr->imgCh[i];
    int s = st->s, N = st->N;
    double* img = st->td1;
    ocplx* oldFT = st->tc3;
    ocplx* ft = st->imgFT + (long)i * st->N2;
    if (doReset) memset(oldFT, 0, sizeof(ocplx) * st->N2);
    else memcpy(oldFT, ft, sizeof(ocplx) * st->N2);

    const float* blk = pr->data + (long)i * pr->dataY * pr->dataX;
    int y0 = pr->startY + st->curShift[2 * i + 1], x0 = pr->startX + st->curShift[2 * i];
    for (int y = 0; y < s; ++y)
        for (int x = 0; x < s; ++x) img[y * s + x] = blk[(long)(y0 + y) * pr->dataX + x0 + x];
    double avg = stats_mean(img, st->s2);
    for (int k = 0; k < st->s2; ++k) img[k] = (img[k] - avg) * st->window[k] + avg;
    double mean = stats_mean(img, st->s2);
    double sd = stats_stddev(img, st->s2, mean);
    st->mean[i] = mean; st->stddev[i] = sd;

    if (doReset) {
        ocplx* d2 = st->tc1;
        for (int k = 0; k < st->s2; ++k) d2[k] = (img[k] - mean) * st->noiseWindow[k];
        o_fft2(d2, s, s, -1);
        double nz = o_noise(d2, s, s, -1, -1) * pr->noiseFudge[c];
        st->noise[i] = nz;
        double rg = nz / sd;
        st->regGamma[o] += 0.10 * rg / st->nImgObj[o];   /* Object::addRegGamma (object.cpp:465-468) */
    }
    if (s != N) {
        int off = (N - s) / 2;
        for (int k = 0; k < st->N2; ++k) ft[k] = mean;
        for (int y = 0; y < s; ++y)
            for (int x = 0; x < s; ++x) ft[(long)(off + y) * N + off + x] = img[y * s + x];
    } else {
        for (int k = 0; k < st->N2; ++k) ft[k] = img[k];
    }
    o_fft2(ft, N, N, -1);
    o_reorder_c(ft, N, N);
    double* fs = st->ftSum + (long)o * st->N2;
    if (doReset) for (int k = 0; k < st->N2; ++k) fs[k] = fs[k] + cnorm(ft[k]);
    else for (int k = 0; k < st->N2; ++k) fs[k] += (cnorm(ft[k]) - cnorm(oldFT[k]));
}

/* Object::reInitialize (object.cpp:885-904) */
static void reinit_object(OState* st, int o, int doReset)
{
    if (doReset) { st->regGamma[o] = 0; memset(st->ftSum + (long)o * st->N2, 0, sizeof(double) * st->N2); }
    for (int i = 0; i < st->pr->nImg; ++i) if (img_obj(st, i) == o) initialize_img(st, i, doReset);
}

/* SubImage::adjustShifts (subimage.cpp:444-480) and Array::shift (array.hpp:1119-1140) */
static int array_shift(int* first, int size, int len, int n)
{
    if (n == 0) return 0;
    int last = *first + len - 1;
    if (n + *first < 0) n = -*first;
    if (abs(n + last) >= size) n = size - last - 1;
    *first += n;
    return n;
}

static int adjust_shifts(OState* st, int i, const double* a)
{
    const OProblem* pr = st->pr;
    int o = img_obj(st, i);
    const double* J = pr->jac + 8L * o;
    int tx = pr->tilt[2 * o], ty = pr->tilt[2 * o + 1];
    double ax = a[tx] + st->tiltOff[2 * i], ay = a[ty] + st->tiltOff[2 * i + 1];
    /* alphaToShift (modes.hpp:126-129) */
    double sx = ax * J[0] + ay * J[1];
    double sy = ax * J[2] + ay * J[3];
    int thx = -(int)lround(sx), thy = -(int)lround(sy);
    int firstX = pr->startX + st->curShift[2 * i], firstY = pr->startY + st->curShift[2 * i + 1];
    int actX = 0, actY = 0;
    if (thx) actX = array_shift(&firstX, pr->dataX, st->s, thx);
    if (thy) actY = array_shift(&firstY, pr->dataY, st->s, thy);
    if (actX || actY) {
        st->curShift[2 * i] += actX; st->curShift[2 * i + 1] += actY;
        st->tiltOff[2 * i] += actX * J[4] + actY * J[5];
        st->tiltOff[2 * i + 1] += actX * J[6] + actY * J[7];
        return 1;
    }
    return 0;
}

/* Solver::shiftAndInit (solver.cpp:583-605) */
static void shift_and_init(OState* st, const double* a, int doReset)
{
    const OProblem* pr = st->pr;
    int* shifted = calloc(pr->nObj, sizeof(int));
    for (int o = 0; o < pr->nObj; ++o) shifted[o] = doReset;
    for (int i = 0; i < pr->nImg; ++i) shifted[img_obj(st, i)] |= adjust_shifts(st, i, a + (long)i * st->K);
    for (int o = 0; o < pr->nObj; ++o) if (shifted[o]) reinit_object(st, o, doReset);
    free(shifted);
}

/* ---------- metric, P/Q, gradient ---------- */

static void apply_alpha(OState* st, const double* a)
{
    for (int i = 0; i < st->pr->nImg; ++i) {
        calc_phi(st, i, a + (long)i * st->K, st->phi + (long)i * st->P2);
        calc_pfotf(st, i);
    }
}

/* Object::initPQ + SubImage::addPQ + Object::addToPQ in tasks of <= 30 images per
 * channel (object.cpp:459-462, subimage.cpp:226-234, object.cpp:508-518, solver.cpp:756-797) */
static void accumulate_pq(OState* st)
{
    const OProblem* pr = st->pr;
    int N2 = st->N2;
    for (int o = 0; o < pr->nObj; ++o) {
        ocplx* P = st->Pc + (long)o * N2; double* Q = st->Qc + (long)o * N2;
        for (int k = 0; k < N2; ++k) { P[k] = 0; Q[k] = st->regGamma[o]; }
        for (int c = 0; c < pr->nCh; ++c) {
            if (pr->chObj[c] != o) continue;
            int imgs[4096]; int n = 0;
            for (int i = 0; i < pr->nImg; ++i) if (pr->imgCh[i] == c) imgs[n++] = i;
            for (int beg = 0; beg < n; beg += 30) {
                int end = beg + 30 < n ? beg + 30 : n;
                ocplx* tC = st->tc1; double* tD = st->td2;
                memset(tC, 0, sizeof(ocplx) * N2); memset(tD, 0, sizeof(double) * N2);
                for (int j = beg; j < end; ++j) {
                    const ocplx* otf = st->OTF + (long)imgs[j] * N2;
                    const ocplx* ft = st->imgFT + (long)imgs[j] * N2;
                    for (int k = 0; k < st->nOS[o]; ++k) {
                        int ind = st->os[o][k];
                        tD[ind] += cnorm(otf[ind]);
                        tC[ind] += conj(ft[ind]) * otf[ind];
                    }
                }
                for (int k = 0; k < st->nOS[o]; ++k) {
                    int ind = st->os[o][k];
                    Q[ind] += tD[ind];
                    P[ind] += tC[ind];
                }
            }
        }
    }
}

/* Solver::metric (solver.cpp:763-817), Object::calcMetric (object.cpp:597-615) */
static double metric_now(OState* st)
{
    const OProblem* pr = st->pr;
    accumulate_pq(st);
    double sum = 0;
    for (int o = 0; o < pr->nObj; ++o) {
        if (pr->weight[o] != 0.0) {
            const double* fs = st->ftSum + (long)o * st->N2;
            const ocplx* P = st->Pc + (long)o * st->N2; const double* Q = st->Qc + (long)o * st->N2;
            double m = 0;
            for (int k = 0; k < st->nOS[o]; ++k) {
                int ind = st->os[o][k];
                m += (fs[ind] - cnorm(P[ind]) / Q[ind]);
            }
            st->curMetric[o] = m / ((double)st->s * st->s);
        }
        sum += pr->weight[o] * st->curMetric[o];
    }
    if (pr->regAlpha > 0)
        for (int i = 0; i < pr->nPar; ++i) sum += 0.5 * st->alpha[i] * st->alpha[i] * st->regW[i];
    st->nF++;
    return sum;
}

/* Object::calcHelpers (object.cpp:531-550) */
static void calc_helpers(OState* st)
{
    for (int o = 0; o < st->pr->nObj; ++o) {
        long off = (long)o * st->N2;
        memset(st->PQ + off, 0, sizeof(ocplx) * st->N2);
        memset(st->PS + off, 0, sizeof(double) * st->N2);
        memset(st->QS + off, 0, sizeof(double) * st->N2);
        for (int k = 0; k < st->nOS[o]; ++k) {
            int ind = st->os[o][k];
            st->PQ[off + ind] = st->Pc[off + ind] * st->Qc[off + ind];
            st->PS[off + ind] = cnorm(st->Pc[off + ind]);
            st->QS[off + ind] = st->Qc[off + ind] * st->Qc[off + ind];
        }
    }
}

/* SubImage::calcVogelWeight (subimage.cpp:348-390) */
static void calc_vogel_weight(OState* st, int i)
{
    const OProblem* pr = st->pr;
    int o = img_obj(st, i), N = st->N, P = st->P, N2 = st->N2;
    long off = (long)o * N2;
    const ocplx* otf = st->OTF + (long)i * N2;
    const ocplx* ft = st->imgFT + (long)i * N2;
    const ocplx* pf = st->PF + (long)i * st->P2;
    ocplx* tmpOtf = st->tc1; ocplx* gl = st->tc2; ocplx* hj = st->tc3;
    memset(tmpOtf, 0, sizeof(ocplx) * N2);
    o_reorder_into_c(pf, P, P, tmpOtf, N, N, 0);
    memcpy(hj, tmpOtf, sizeof(ocplx) * N2);                 /* ift(hj): backward / N^2 */
    o_fft2(hj, N, N, +1);
    for (int k = 0; k < N2; ++k) hj[k] *= 1.0 / N2;
    memset(tmpOtf, 0, sizeof(ocplx) * N2);
    for (int k = 0; k < st->nOS[o]; ++k) {
        int ind = st->os[o][k];
        tmpOtf[ind] = (st->PQ[off + ind] * ft[ind] - st->PS[off + ind] * otf[ind]) / st->QS[off + ind];
    }
    o_reorder_c(tmpOtf, N, N);
    memcpy(gl, tmpOtf, sizeof(ocplx) * N2);                 /* ift(gl) */
    o_fft2(gl, N, N, +1);
    for (int k = 0; k < N2; ++k) gl[k] *= 1.0 / N2;
    for (int k = 0; k < N2; ++k) gl[k] = hj[k] * creal(gl[k]);
    memcpy(tmpOtf, gl, sizeof(ocplx) * N2);                 /* ft(gl), unnormalised */
    o_fft2(tmpOtf, N, N, -1);
    o_reorder_c(tmpOtf, N, N);
    double* vog = st->vogel + (long)i * st->P2;
    const double* pup = pr->pupil + (long)o * st->P2;
    for (int k = 0; k < st->nPS[o]; ++k) {
        int p = st->ps[o][k];
        vog[p] = cimag(conj(pf[p]) * tmpOtf[st->pio[o][k]]) * pup[p];
    }
}

/* SubImage::metricChange (subimage.cpp:261-278) */
static double metric_change(OState* st, int i, const ocplx* newOTF)
{
    int o = img_obj(st, i); long off = (long)o * st->N2;
    const ocplx* oldOTF = st->OTF + (long)i * st->N2;
    const ocplx* ft = st->imgFT + (long)i * st->N2;
    double dl = 0.0;
    for (int k = 0; k < st->nOS[o]; ++k) {
        int ind = st->os[o][k];
        ocplx dsj = newOTF[ind] - oldOTF[ind];
        ocplx dp = conj(ft[ind]) * dsj;
        ocplx p = st->Pc[off + ind]; double q = st->Qc[off + ind];
        double dq = 2.0 * (creal(oldOTF[ind]) * creal(dsj) + cimag(oldOTF[ind]) * cimag(dsj)) + cnorm(dsj);
        double dn = 2.0 * (creal(dp) * creal(p) + cimag(dp) * cimag(p)) + cnorm(dp);
        dl -= (q * dn - dq * cnorm(p)) / (q * (q + dq));
    }
    return dl / st->N2;
}

/* Solver::gradient (solver.cpp:860-894) with gradientVogel / gradientFiniteDifference */
static void gradient_now(OState* st)
{
    const OProblem* pr = st->pr;
    if (pr->regAlpha > 0) for (int i = 0; i < pr->nPar; ++i) st->gradAlpha[i] = st->alpha[i] * st->regW[i];
    else memset(st->gradAlpha, 0, sizeof(double) * pr->nPar);
    double* tphi = st->td2;
    for (int i = 0; i < pr->nImg; ++i) {
        int o = img_obj(st, i);
        double w = pr->weight[o];
        calc_vogel_weight(st, i);
        double* g = st->gradAlpha + (long)i * st->K;
        for (int m = 0; m < st->K; ++m) {
            if (!st->enabled[m]) continue;
            const double* mode = pr->modes + ((long)o * st->K + m) * st->P2;
            double val = 0;
            if (w != 0) {
                if (st->gradMethod == 2) {
                    double scale = -2.0 * st->area[o] / st->N2;
                    const double* vog = st->vogel + (long)i * st->P2;
                    double ret = 0;
                    for (int k = 0; k < st->nPS[o]; ++k) { int p = st->ps[o][k]; ret += scale * vog[p] * mode[p]; }
                    val = ret * w;
                } else {
                    double step = pr->graddiffStep * pr->wavelength[o];
                    memcpy(tphi, st->phi + (long)i * st->P2, sizeof(double) * st->P2);
                    for (int p = 0; p < st->P2; ++p) tphi[p] = tphi[p] + step * mode[p];
                    calc_otf(st, i, st->tc1, tphi);
                    val = metric_change(st, i, st->tc1) / step * w;
                }
            }
            g[m] = w * val;
        }
    }
    st->nDF++;
}

/* Constraints::apply / reverseAndAdd (constraints.hpp:128-146) */
static void ns_apply(const OProblem* pr, const double* in, double* out)
{
    memset(out, 0, sizeof(double) * pr->nFree);
    for (int e = 0; e < pr->nnz; ++e) out[pr->nsCol[e]] += pr->nsVal[e] * in[pr->nsRow[e]];
}

static void ns_reverse_add(const OProblem* pr, const double* in, const double* offset, double* out)
{
    memcpy(out, offset, sizeof(double) * pr->nPar);
    for (int e = 0; e < pr->nnz; ++e) out[pr->nsRow[e]] += pr->nsVal[e] * in[pr->nsCol[e]];
}

/* the metricNoise factor (sensitivity experiments; 1 when metricNoise is 0) */
static double noise_factor(OState* st)
{
    if (st->pr->metricNoise == 0.0) return 1.0;
    unsigned long long z = (st->noiseState += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    double u = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    return 1.0 + st->pr->metricNoise * u;
}

/* Solver::my_f / my_df (solver.cpp:219-234) */
static double my_f(OState* st, const double* beta)
{
    ns_reverse_add(st->pr, beta, st->alphaOffset, st->alpha);
    apply_alpha(st, st->alpha);
    return metric_now(st) * noise_factor(st);
}

static void my_df(OState* st, const double* beta, double* gb)
{
    ns_reverse_add(st->pr, beta, st->alphaOffset, st->alpha);
    apply_alpha(st, st->alpha);
    accumulate_pq(st);
    calc_helpers(st);
    gradient_now(st);
    ns_apply(st->pr, st->gradAlpha, gb);
}

/* Solver::my_fdf (solver.cpp:237-243): metric, then the gradient from the same P, Q */
static void my_fdf(OState* st, const double* beta, double* f, double* gb)
{
    ns_reverse_add(st->pr, beta, st->alphaOffset, st->alpha);
    apply_alpha(st, st->alpha);
    *f = metric_now(st) * noise_factor(st);
    calc_helpers(st);
    gradient_now(st);
    ns_apply(st->pr, st->gradAlpha, gb);
}

static double cb_f(void* ctx, const double* x) { return my_f((OState*)ctx, x); }
static void cb_df(void* ctx, const double* x, double* g) { my_df((OState*)ctx, x, g); }
static void cb_fdf(void* ctx, const double* x, double* f, double* g) { my_fdf((OState*)ctx, x, f, g); }

/* ---------- conjugate_rdx + bracket + brent ---------- */

typedef struct {
    const OFdf* fn; const double* x; const double* p; double lambda; double* x1; double* dx; int n;
} LineCtx;

/* take_step / value_at (conjugate_gradient.cpp:110-125) */
static void take_step(const double* x, const double* p, double step, double lambda, double* x1, double* dx, int n)
{
    for (int k = 0; k < n; ++k) dx[k] = 0.0 + (-step * lambda) * p[k];
    for (int k = 0; k < n; ++k) x1[k] = x[k] + 1.0 * dx[k];
}

static double value_at(LineCtx* c, double step)
{
    take_step(c->x, c->p, step, c->lambda, c->x1, c->dx, c->n);
    return c->fn->f(c->fn->ctx, c->x1);
}

#define SHFT(a, b, c, d) (a) = (b); (b) = (c); (c) = (d);
#define SIGN(a, b) ((b) >= 0.0 ? fabs(a) : -fabs(a))

/* math::bracket (helpers.cpp:58-129) */
static void bracket(LineCtx* F, double* a, double* b, double* mid, double* fa, double* fb, double* fmid, int limit)
{
    double u, fu, ulim, t;
    if (*fb > *fa) { t = *a; *a = *b; *b = t; t = *fa; *fa = *fb; *fb = t; }
    *mid = *b + GOLDEN * (*b - *a);
    *fmid = value_at(F, *mid);
    while (*fb > *fmid) {
        double r = (*b - *a) * (*fb - *fmid);
        double q = (*b - *mid) * (*fb - *fa);
        double m1 = fabs(q - r), m2 = TINY;
        u = (*b) - ((*b - *mid) * q - (*b - *a) * r) / (2.0 * SIGN((m1 > m2 ? m1 : m2), q - r));
        ulim = *b + limit * (*mid - *b);
        if ((*b - u) * (u - *mid) > 0.0) {
            fu = value_at(F, u);
            if (fu < *fmid) { *a = *b; *b = u; *fa = *fb; *fb = fu; break; }
            else if (fu > *fb) { *mid = u; *fmid = fu; break; }
            u = *mid + GOLDEN * (*mid - *b);
            fu = value_at(F, u);
        } else if ((*mid - u) * (u - ulim) > 0.0) {
            fu = value_at(F, u);
            if (fu < *fmid) {
                SHFT(*b, *mid, u, *mid + GOLDEN * (*mid - *b))
                SHFT(*fb, *fmid, fu, value_at(F, u))
            }
        } else if ((*mid - ulim) * (ulim - u) >= 0.0) {
            u = ulim;
            fu = value_at(F, u);
        } else {
            u = *mid + GOLDEN * (*mid - *b);
            fu = value_at(F, u);
        }
        SHFT(*a, *b, *mid, u)
        SHFT(*fa, *fb, *fmid, fu)
    }
    if (*a > *b) { t = *a; *a = *b; *b = t; t = *fa; *fa = *fb; *fb = t; }
    if (*mid > *b) { t = *mid; *mid = *b; *b = t; t = *fmid; *fmid = *fb; *fb = t; }
    if (*mid < *a) { t = *mid; *mid = *a; *a = t; t = *fmid; *fmid = *fa; *fa = t; }
}

/* math::brent (helpers.cpp:133-216) */
static void brent(LineCtx* F, double ax, double bx, double* xio, double* fxo, double tol, int limit)
{
    const double CGOLD = 2.0 - GOLDEN;
    double a = ax < bx ? ax : bx, b = ax > bx ? ax : bx;
    double x = *xio, u, v, w, fu, fv, fw, fx;
    u = v = w = x;
    fu = fv = fw = fx = value_at(F, x);
    double d = 0.0, e = 0.0;
    for (int iter = 1; iter <= limit; ++iter) {
        double xm = 0.5 * (a + b);
        double tol1 = tol * fabs(x) + TINY;
        double tol2 = 2.0 * tol1;
        if (fabs(x - xm) <= (tol2 - 0.5 * (b - a))) { *xio = x; *fxo = fx; return; }
        if (fabs(e) > tol1) {
            double r = (x - w) * (fx - fv);
            double q = (x - v) * (fx - fw);
            double p = (x - v) * q - (x - w) * r;
            q = 2.0 * (q - r);
            if (q > 0.0) p = -p;
            q = fabs(q);
            double etemp = e;
            e = d;
            if (fabs(p) >= fabs(0.5 * q * etemp) || p <= q * (a - x) || p >= q * (b - x)) {
                e = (x >= xm ? a - x : b - x);
                d = CGOLD * e;
            } else {
                d = p / q;
                u = x + d;
                if (u - a < tol2 || b - u < tol2) d = SIGN(tol1, xm - x);
            }
        } else {
            e = (x >= xm ? a - x : b - x);
            d = CGOLD * e;
        }
        if (fabs(d) < tol1) d *= (tol1 / fabs(d));
        u = ((fabs(d) >= tol1) ? x + d : x + SIGN(tol1, d));
        fu = value_at(F, u);
        if (fu <= fx) {
            if (u >= x) a = x; else b = x;
            SHFT(v, w, x, u)
            SHFT(fv, fw, fx, fu)
        } else {
            if (u < x) a = u; else b = u;
            if (fu <= fw || w == x) { v = w; w = u; fv = fw; fw = fu; }
            else if (fu <= fv || v == x || v == w) { v = u; fv = fu; }
        }
    }
    *xio = x; *fxo = fx;
}

static double dnrm2(const double* v, int n)
{
    double s = 0; for (int k = 0; k < n; ++k) s += v[k] * v[k];
    return sqrt(s);
}

typedef struct {
    int iter; double step, tol, pnorm, g0norm;
    double* x1; double* p; double* prevG;
} CGState;

/* conjugate_rdx_iterate (conjugate_gradient.cpp:128-256) */
static int cg_iterate(const OFdf* fn, CGState* cs, double* x, double* f, double* g, double* dx, int n)
{
    fn->df(fn->ctx, x, g);
    double pnorm = cs->pnorm, g0norm = cs->g0norm;
    double g1norm = dnrm2(g, n);
    double pg;
    if (cs->iter++ == 0) {
        memcpy(cs->p, g, sizeof(double) * n);
        pnorm = cs->pnorm = g1norm;
        pg = g1norm * g1norm;
    } else {
        if (g0norm == 0.0) { memset(dx, 0, sizeof(double) * n); return GSL_ENOPROG; }
        double g0g1 = 0;
        for (int k = 0; k < n; ++k) { double t = cs->prevG[k] - g[k]; g0g1 += t * g[k]; cs->prevG[k] = t; }
        double beta = g0g1 / (g0norm * g0norm);
        for (int k = 0; k < n; ++k) cs->p[k] = g[k] - beta * cs->p[k];
        pnorm = cs->pnorm = dnrm2(cs->p, n);
        pg = 0; for (int k = 0; k < n; ++k) pg += cs->p[k] * g[k];
    }
    double lambda = (pg >= 0.0) ? +1.0 / pnorm : -1.0 / pnorm;
    if (pnorm == 0.0) { memset(dx, 0, sizeof(double) * n); return GSL_ENOPROG; }
    double fa = *f, fb, fc;
    double stepa = 0.0, stepb = 0, stepc = 1E-21 / lambda;
    LineCtx lc = { fn, x, cs->p, lambda, cs->x1, dx, n };
    fc = value_at(&lc, stepc);
    bracket(&lc, &stepa, &stepc, &stepb, &fa, &fc, &fb, 100);
    brent(&lc, stepa, stepc, &stepb, &fb, 0.1, 100);
    take_step(x, cs->p, stepb, lambda, cs->x1, dx, n);
    memcpy(x, cs->x1, sizeof(double) * n);
    *f = fb;
    cs->step = 0.5 * (fabs(stepb) > 2.0E-9 ? fabs(stepb) : 2.0E-9);
    cs->g0norm = g1norm;
    memcpy(cs->prevG, g, sizeof(double) * n);
    cs->iter = cs->iter % n;
    return GSL_SUCCESS;
}

/* conjugate_rdx, bracket and brent on caller-supplied functions, so the tests can check
   them on functions whose minimisers are known in closed form.  The solve above runs the
   same cg_iterate, bracket and brent. */
struct OCg {
    const OFdf* fn;
    CGState cs;
    int n;
};

OCg* o_cg_new(int n)
{
    OCg* w = calloc(1, sizeof(OCg));
    w->n = n;
    w->cs.x1 = calloc(n ? n : 1, sizeof(double));
    w->cs.p = calloc(n ? n : 1, sizeof(double));
    w->cs.prevG = calloc(n ? n : 1, sizeof(double));
    return w;
}

void o_cg_free(OCg* w)
{
    if (!w) return;
    free(w->cs.x1); free(w->cs.p); free(w->cs.prevG); free(w);
}

/* conjugate_rdx_set (conjugate_gradient.cpp:72-86), f only (see the file header) */
void o_cg_set(OCg* w, const OFdf* fn, const double* x, double* f, double stepSize, double tol)
{
    w->fn = fn;
    w->cs.iter = 0;
    w->cs.step = stepSize;
    w->cs.tol = tol;
    *f = fn->f(fn->ctx, x);
}

int o_cg_iterate(OCg* w, double* x, double* f, double* g, double* dx)
{
    return cg_iterate(w->fn, &w->cs, x, f, g, dx, w->n);
}

/* the search direction of the last iteration (PR update of conjugate_gradient.cpp:165-182) */
void o_cg_direction(const OCg* w, double* p)
{
    memcpy(p, w->cs.p, sizeof(double) * w->n);
}

/* bracket / brent on a function of one variable: fn->f receives a one-element x.  The line
   runs from x = 0 along p = 1 with lambda = -1, so every probe is exactly at the step. */
void o_bracket(const OFdf* fn, double* a, double* b, double* mid, double* fa, double* fb, double* fmid, int limit)
{
    double x0 = 0.0, p = 1.0, x1 = 0.0, dx = 0.0;
    LineCtx lc = { fn, &x0, &p, -1.0, &x1, &dx, 1 };
    bracket(&lc, a, b, mid, fa, fb, fmid, limit);
}

void o_brent(const OFdf* fn, double ax, double bx, double* x, double* fx, double tol, int limit)
{
    double x0 = 0.0, p = 1.0, x1 = 0.0, dx = 0.0;
    LineCtx lc = { fn, &x0, &p, -1.0, &x1, &dx, 1 };
    brent(&lc, ax, bx, x, fx, tol, limit);
}

/* ---------- GETSTEP = BFGS_inv: GSL 1.16 vector_bfgs2 ----------
   The reference's default stepper (config.cpp:632; solver.cpp:362-368 selects
   gsl_multimin_fdfminimizer_vector_bfgs2 with init_step = 1e-4 lambda_max, tol = 1e-6).
   GSL is not in this tree; its published algorithm is restated here:
   - direction: BFGS update of the inverse Hessian applied implicitly,
     p' = g1 - A dx - B dg, B = dx.g1/dx.dg, A = -(1 + dg.dg/dx.dg) B + dg.g1/dx.dg,
     then p <- -sign(p.g1) p/|p| (multimin/vector_bfgs2.c);
   - line search: Fletcher's bracketing and sectioning (Practical Methods of
     Optimization, algorithms 2.6.2 and 2.6.4) with rho = 0.01, sigma = tol,
     tau1 = 9, tau2 = 0.05, tau3 = 0.5 and cubic interpolation where both ends have a
     slope (quadratic otherwise), at most 100 trial points (multimin/linear_minimize.c);
   - a line wrapper that caches f, the gradient and the slope at the last trial point
     and at alpha = 0, so no point is evaluated twice (multimin/linear_wrapper.c).
   Parity with GSL's exact arithmetic is unpinned (no GSL build here); the tests pin the
   mathematics: finite termination on quadratics and the Wolfe conditions at every
   accepted step. */

#define O_DBL_EPSILON 2.2204460492503131e-16

struct OBfgs2 {
    /* caching line wrapper over x + alpha p */
    const OFdf* fn;
    double *x, *g, *p, *xa, *ga;
    double fa, dfa, xKey, fKey, gKey, dfKey;
    /* stepper state */
    double step, g0norm, pnorm, deltaF, fp0;
    double *dx0, *dg0;
    double rho, sigma, tau1, tau2, tau3;
    int n;
};

static double o_dot(const double* a, const double* b, int n) { double s = 0; for (int k = 0; k < n; ++k) s += a[k] * b[k]; return s; }

static void lw_moveto(OBfgs2* w, double alpha)
{
    if (alpha == w->xKey) return;
    for (int k = 0; k < w->n; ++k) w->xa[k] = w->x[k] + alpha * w->p[k];
    w->xKey = alpha;
}

static double lw_f(OBfgs2* w, double alpha)
{
    if (alpha == w->fKey) return w->fa;
    lw_moveto(w, alpha);
    w->fa = w->fn->f(w->fn->ctx, w->xa);
    w->fKey = alpha;
    return w->fa;
}

static double lw_df(OBfgs2* w, double alpha)
{
    if (alpha == w->dfKey) return w->dfa;
    lw_moveto(w, alpha);
    if (alpha != w->gKey) { w->fn->df(w->fn->ctx, w->xa, w->ga); w->gKey = alpha; }
    w->dfa = o_dot(w->ga, w->p, w->n);
    w->dfKey = alpha;
    return w->dfa;
}

static void lw_fdf(OBfgs2* w, double alpha, double* f, double* df)
{
    if (alpha == w->fKey && alpha == w->dfKey) { *f = w->fa; *df = w->dfa; return; }
    if (alpha == w->fKey || alpha == w->dfKey) { *f = lw_f(w, alpha); *df = lw_df(w, alpha); return; }
    lw_moveto(w, alpha);
    w->fn->fdf(w->fn->ctx, w->xa, &w->fa, w->ga);
    w->fKey = alpha; w->gKey = alpha;
    w->dfa = o_dot(w->ga, w->p, w->n);
    w->dfKey = alpha;
    *f = w->fa; *df = w->dfa;
}

/* new origin alpha = 0 at the end point of the last line search, along the new p */
static void lw_change_direction(OBfgs2* w)
{
    memcpy(w->xa, w->x, sizeof(double) * w->n);
    w->xKey = 0.0;
    w->fKey = 0.0;
    memcpy(w->ga, w->g, sizeof(double) * w->n);
    w->gKey = 0.0;
    w->dfa = o_dot(w->ga, w->p, w->n);
    w->dfKey = 0.0;
}

/* real roots of a z^2 + b z + c, ascending (gsl_poly_solve_quadratic) */
static int o_solve_quadratic(double a, double b, double c, double* z0, double* z1)
{
    if (a == 0) {
        if (b == 0) return 0;
        *z0 = -c / b;
        return 1;
    }
    double disc = b * b - 4 * a * c;
    if (disc > 0) {
        if (b == 0) { double r = sqrt(-c / a); *z0 = -r; *z1 = r; }
        else {
            double sgnb = (b > 0 ? 1 : -1), temp = -0.5 * (b + sgnb * sqrt(disc));
            double r1 = temp / a, r2 = c / temp;
            if (r1 < r2) { *z0 = r1; *z1 = r2; } else { *z0 = r2; *z1 = r1; }
        }
        return 2;
    }
    if (disc == 0) { *z0 = -0.5 * b / a; *z1 = -0.5 * b / a; return 2; }
    return 0;
}

/* minimum on [zl, zh] of the quadratic through (0, f0) with slope fp0 and (1, f1) */
static double o_interp_quad(double f0, double fp0, double f1, double zl, double zh)
{
    double fl = f0 + zl * (fp0 + zl * (f1 - f0 - fp0));
    double fh = f0 + zh * (fp0 + zh * (f1 - f0 - fp0));
    double c = 2 * (f1 - f0 - fp0);
    double zmin = zl, fmin = fl;
    if (fh < fmin) { zmin = zh; fmin = fh; }
    if (c > 0) {
        double z = -fp0 / c;
        if (z > zl && z < zh) {
            double f = f0 + z * (fp0 + z * (f1 - f0 - fp0));
            if (f < fmin) { zmin = z; fmin = f; }
        }
    }
    return zmin;
}

static double o_cubic(double c0, double c1, double c2, double c3, double z) { return c0 + z * (c1 + z * (c2 + z * c3)); }

/* minimum on [zl, zh] of the cubic through (0, f0), (1, f1) with slopes fp0, fp1 */
static double o_interp_cubic(double f0, double fp0, double f1, double fp1, double zl, double zh)
{
    double eta = 3 * (f1 - f0) - 2 * fp0 - fp1, xi = fp0 + fp1 - 2 * (f1 - f0);
    double c0 = f0, c1 = fp0, c2 = eta, c3 = xi, z0, z1;
    double zmin = zl, fmin = o_cubic(c0, c1, c2, c3, zl), y;
    y = o_cubic(c0, c1, c2, c3, zh);
    if (y < fmin) { zmin = zh; fmin = y; }
    int nr = o_solve_quadratic(3 * c3, 2 * c2, c1, &z0, &z1);
    if (nr >= 1 && z0 > zl && z0 < zh) { y = o_cubic(c0, c1, c2, c3, z0); if (y < fmin) { zmin = z0; fmin = y; } }
    if (nr == 2 && z1 > zl && z1 < zh) { y = o_cubic(c0, c1, c2, c3, z1); if (y < fmin) { zmin = z1; fmin = y; } }
    return zmin;
}

/* trial point in [xmin, xmax] from the end points a, b (slope fpb may be NaN: quadratic) */
static double o_interpolate(double a, double fa, double fpa, double b, double fb, double fpb, double xmin,
                            double xmax)
{
    double zmin = (xmin - a) / (b - a), zmax = (xmax - a) / (b - a), z;
    if (zmin > zmax) { double t = zmin; zmin = zmax; zmax = t; }
    if (isfinite(fpb)) z = o_interp_cubic(fa, fpa * (b - a), fb, fpb * (b - a), zmin, zmax);
    else z = o_interp_quad(fa, fpa * (b - a), fb, zmin, zmax);
    return a + z * (b - a);
}

/* Fletcher's line search along the wrapper's p from alpha = 0 (cached); *alphaNew stays
   unchanged when the trial budget runs out */
static int o_line_minimize(OBfgs2* w, double alpha1, double* alphaNew)
{
    const double rho = w->rho, sigma = w->sigma, tau1 = w->tau1, tau2 = w->tau2, tau3 = w->tau3;
    double f0, fp0, falpha, falphaPrev, fpalpha, fpalphaPrev, delta, alphaNext;
    double alpha = alpha1, alphaPrev = 0.0, a, b, fa, fb, fpa, fpb;
    int i = 0;
    lw_fdf(w, 0.0, &f0, &fp0);
    falphaPrev = f0;
    fpalphaPrev = fp0;
    a = 0.0; b = alpha; fa = f0; fb = 0.0; fpa = fp0; fpb = 0.0;
    while (i++ < 100) {                                     /* bracketing */
        falpha = lw_f(w, alpha);
        if (falpha > f0 + alpha * rho * fp0 || falpha >= falphaPrev) {
            a = alphaPrev; fa = falphaPrev; fpa = fpalphaPrev;
            b = alpha; fb = falpha; fpb = NAN;
            break;
        }
        fpalpha = lw_df(w, alpha);
        if (fabs(fpalpha) <= -sigma * fp0) { *alphaNew = alpha; return GSL_SUCCESS; }
        if (fpalpha >= 0) {
            a = alpha; fa = falpha; fpa = fpalpha;
            b = alphaPrev; fb = falphaPrev; fpb = fpalphaPrev;
            break;
        }
        delta = alpha - alphaPrev;
        alphaNext = o_interpolate(alphaPrev, falphaPrev, fpalphaPrev, alpha, falpha, fpalpha, alpha + delta,
                                  alpha + tau1 * delta);
        alphaPrev = alpha; falphaPrev = falpha; fpalphaPrev = fpalpha;
        alpha = alphaNext;
    }
    while (i++ < 100) {                                     /* sectioning of [a, b] */
        delta = b - a;
        alpha = o_interpolate(a, fa, fpa, b, fb, fpb, a + tau2 * delta, b - tau3 * delta);
        falpha = lw_f(w, alpha);
        if ((a - alpha) * fpa <= O_DBL_EPSILON) return GSL_ENOPROG;     /* roundoff prevents progress */
        if (falpha > f0 + rho * alpha * fp0 || falpha >= fa) {
            b = alpha; fb = falpha; fpb = NAN;
        } else {
            fpalpha = lw_df(w, alpha);
            if (fabs(fpalpha) <= -sigma * fp0) { *alphaNew = alpha; return GSL_SUCCESS; }
            if (((b - a) >= 0 && fpalpha >= 0) || ((b - a) <= 0 && fpalpha <= 0)) {
                b = a; fb = fa; fpb = fpa;
            }
            a = alpha; fa = falpha; fpa = fpalpha;
        }
    }
    return GSL_SUCCESS;
}

OBfgs2* o_bfgs2_new(int n)
{
    OBfgs2* w = calloc(1, sizeof(OBfgs2));
    w->n = n;
    w->x = calloc(n, sizeof(double)); w->g = calloc(n, sizeof(double)); w->p = calloc(n, sizeof(double));
    w->xa = calloc(n, sizeof(double)); w->ga = calloc(n, sizeof(double));
    w->dx0 = calloc(n, sizeof(double)); w->dg0 = calloc(n, sizeof(double));
    return w;
}

void o_bfgs2_free(OBfgs2* w)
{
    if (!w) return;
    free(w->x); free(w->g); free(w->p); free(w->xa); free(w->ga); free(w->dx0); free(w->dg0); free(w);
}

/* vector_bfgs2_set: f and g at x, first direction -g/|g| */
void o_bfgs2_set(OBfgs2* w, const OFdf* fn, const double* x, double* f, double* g, double stepSize, double tol)
{
    const int n = w->n;
    w->fn = fn;
    w->step = stepSize;
    w->deltaF = 0;
    fn->fdf(fn->ctx, x, f, g);
    memcpy(w->x, x, sizeof(double) * n);
    memcpy(w->g, g, sizeof(double) * n);
    w->g0norm = dnrm2(w->g, n);
    for (int k = 0; k < n; ++k) w->p[k] = g[k] * (-1 / w->g0norm);
    w->pnorm = dnrm2(w->p, n);
    w->fp0 = -w->g0norm;
    memcpy(w->xa, w->x, sizeof(double) * n);              /* the wrapper at alpha = 0 */
    w->xKey = 0.0;
    w->fa = *f;
    w->fKey = 0.0;
    memcpy(w->ga, w->g, sizeof(double) * n);
    w->gKey = 0.0;
    w->dfa = o_dot(w->ga, w->p, n);
    w->dfKey = 0.0;
    w->rho = 0.01; w->sigma = tol; w->tau1 = 9; w->tau2 = 0.05; w->tau3 = 0.5;
}

/* vector_bfgs2_iterate: x, f, g move to the line minimum; dx = the step taken */
int o_bfgs2_iterate(OBfgs2* w, double* x, double* f, double* g, double* dx)
{
    const int n = w->n;
    double alpha = 0.0, alpha1;
    const double f0 = *f;
    if (w->pnorm == 0.0 || w->g0norm == 0.0 || w->fp0 == 0) {
        memset(dx, 0, sizeof(double) * n);
        return GSL_ENOPROG;
    }
    if (w->deltaF < 0) {
        double del = fmax(-w->deltaF, 10 * O_DBL_EPSILON * fabs(f0));
        alpha1 = fmin(1.0, 2.0 * del / (-w->fp0));
    } else {
        alpha1 = fabs(w->step);
    }
    int status = o_line_minimize(w, alpha1, &alpha);
    if (status != GSL_SUCCESS) return status;
    {   /* update_position: f and g at alpha, cached or evaluated together */
        double fa, dfa;
        lw_fdf(w, alpha, &fa, &dfa);
        *f = w->fa;
        memcpy(x, w->xa, sizeof(double) * n);
        memcpy(g, w->ga, sizeof(double) * n);
    }
    w->deltaF = *f - f0;
    for (int k = 0; k < n; ++k) { w->dx0[k] = x[k] - w->x[k]; dx[k] = w->dx0[k]; w->dg0[k] = g[k] - w->g[k]; }
    double dxg = o_dot(w->dx0, g, n), dgg = o_dot(w->dg0, g, n), dxdg = o_dot(w->dx0, w->dg0, n);
    double dgnorm = dnrm2(w->dg0, n), A, B;
    if (dxdg != 0) {
        B = dxg / dxdg;
        A = -(1.0 + dgnorm * dgnorm / dxdg) * B + dgg / dxdg;
    } else {
        B = 0; A = 0;
    }
    for (int k = 0; k < n; ++k) w->p[k] = g[k] + (-A) * w->dx0[k] + (-B) * w->dg0[k];
    memcpy(w->g, g, sizeof(double) * n);
    memcpy(w->x, x, sizeof(double) * n);
    w->g0norm = dnrm2(w->g, n);
    w->pnorm = dnrm2(w->p, n);
    double pg = o_dot(w->p, g, n), dir = (pg >= 0.0) ? -1.0 : +1.0;
    for (int k = 0; k < n; ++k) w->p[k] *= dir / w->pnorm;
    w->pnorm = dnrm2(w->p, n);
    w->fp0 = o_dot(w->p, w->g, n);
    lw_change_direction(w);
    return GSL_SUCCESS;
}

/* ---------- restoration (object.cpp:278-443, utils.cpp:304-392) ---------- */

static void smooth2(double* d, int ny, int nx, int n, double* tmp)
{
    for (int y = 0; y < ny; ++y) {
        int yl = y - n > 0 ? y - n : 0, yh = y + n < ny ? y + n : ny;
        for (int x = 0; x < nx; ++x) {
            int xl = x - n > 0 ? x - n : 0, xh = x + n < nx ? x + n : nx;
            double acc = 0; int cnt = 0;
            for (int yy = yl; yy < yh; ++yy)
                for (int xx = xl; xx < xh; ++xx) { acc += d[yy * nx + xx]; cnt++; }
            tmp[y * nx + x] = acc / cnt;
        }
    }
    memcpy(d, tmp, sizeof(double) * ny * nx);
}

static void connected_region(double* d, int ny, int nx, int y0, int x0)
{
    unsigned char* mask = calloc((long)ny * nx, 1);
    int* stack = malloc(sizeof(int) * 4 * ny * nx + 16);
    int sp = 0;
    stack[sp++] = y0 * nx + x0;
    while (sp) {
        int idx = stack[--sp];
        int y = idx / nx, x = idx % nx;
        if (mask[idx]) continue;
        if (!((y == y0 && x == x0) || d[idx] > 0.0)) continue;
        mask[idx] = 1;
        if (y + 1 < ny) stack[sp++] = idx + nx;
        if (y - 1 >= 0) stack[sp++] = idx - nx;
        if (x + 1 < nx) stack[sp++] = idx + 1;
        if (x - 1 >= 0) stack[sp++] = idx - 1;
    }
    for (long k = 0; k < (long)ny * nx; ++k) d[k] *= mask[k];
    free(mask); free(stack);
}

/* ScharmerFilter (utils.cpp:334-392) on a centred ny x nx grid */
static void scharmer_filter(ocplx* data, const double* q2, int ny, int nx, double noise_power, double cutoff)
{
    double noiseFactor = noise_power * ny * nx;
    long n = (long)ny * nx;
    double* pt = malloc(sizeof(double) * n); double* rr = malloc(sizeof(double) * n);
    double* tmp = malloc(sizeof(double) * n);
    for (long k = 0; k < n; ++k) pt[k] = cnorm(data[k]) * q2[k];
    smooth2(pt, ny, nx, 1, tmp);
    for (long k = 0; k < n; ++k) pt[k] = noiseFactor / pt[k];
    for (int y = 0; y < ny; ++y) {
        rr[y * nx] = pt[y * nx];
        for (int x = 1; x < nx; ++x) rr[y * nx + x] = pt[y * nx + nx - x];
    }
    double* col = malloc(sizeof(double) * ny);
    for (int x = 0; x < nx; ++x) {
        double v = 1.0 - 0.5 * (rr[x] + pt[x]);
        rr[x] = (v < 0.0) ? 0.0 : v;
        for (int y = 1; y < ny; ++y) col[y] = rr[y * nx + x];
        for (int y = 1; y < ny; ++y) {
            double w = 1.0 - 0.5 * (col[ny - y] + pt[y * nx + x]);
            rr[y * nx + x] = (w < 0.0) ? 0.0 : w;
        }
    }
    free(col);
    for (long k = 0; k < n; ++k) { if (rr[k] < 0.2) rr[k] = 0.0; if (rr[k] > 1.0) rr[k] = 1.0; }
    int yH = ny / 2, xH = nx / 2;
    rr[yH * nx + xH] = 1.0;
    connected_region(rr, ny, nx, yH, xH);
    smooth2(rr, ny, nx, 4, tmp);
    double c2 = cutoff * cutoff;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            long k = (long)y * nx + x;
            if ((double)((y - yH) * (y - yH) + (x - xH) * (x - xH)) > c2) data[k] = 0.0;
            else data[k] *= rr[k];
        }
    free(pt); free(rr); free(tmp);
}

/* The restored object o on the full N x N grid, before the centred crop (object.cpp:298-331):
   sums of SubImage::restore, conj, safeDivide, Scharmer filter, normalised inverse
   transform (avgObjFT's flags: de-centre, no image reorder) */
static void restore_full(OState* st, int o, double* full, int normalise)
{
    const OProblem* pr = st->pr;
    int N = st->N, N2 = st->N2;
    ocplx* ao = calloc(N2, sizeof(ocplx)); double* dn = calloc(N2, sizeof(double));
    ocplx* tmp = malloc(sizeof(ocplx) * N2);
    double avgNoiseVar = 0; int cnt = 0;
    for (int i = 0; i < pr->nImg; ++i) {
        if (img_obj(st, i) != o) continue;
        const ocplx* otf = st->OTF + (long)i * N2; const ocplx* ft = st->imgFT + (long)i * N2;
        for (int k = 0; k < st->nOS[o]; ++k) {       /* SubImage::restore (subimage.cpp:242-258) */
            int ind = st->os[o][k];
            dn[ind] += cnorm(otf[ind]);
            ao[ind] += conj(ft[ind]) * otf[ind];
        }
        avgNoiseVar += sqr(st->noise[i]);
        cnt++;
    }
    for (int k = 0; k < N2; ++k) ao[k] = conj(ao[k]);
    avgNoiseVar /= cnt;
    for (int k = 0; k < N2; ++k) ao[k] = (fabs(dn[k]) > 0) ? ao[k] / dn[k] : 0;   /* safeDivide, eps 0 */
    if (!pr->noFilter) scharmer_filter(ao, dn, N, N, avgNoiseVar, 0.90 * pr->freqCutoff[o]);
    o_reorder_into_c(ao, N, N, tmp, 0, 0, 0);
    o_fft2(tmp, N, N, +1);
    for (int k = 0; k < N2; ++k) full[k] = normalise ? creal(tmp[k]) / N2 : creal(tmp[k]);
    free(ao); free(dn); free(tmp);
}

void o_scharmer_filter(ocplx* data, const double* q2, int ny, int nx, double noisePower, double cutoff)
{
    scharmer_filter(data, q2, ny, nx, noisePower, cutoff);
}

static void restore_objects(OState* st, double* imgOut)
{
    const OProblem* pr = st->pr;
    int N = st->N, s = st->s, off = (N - s) / 2;
    double* full = malloc(sizeof(double) * st->N2);
    for (int o = 0; o < pr->nObj; ++o) {
        restore_full(st, o, full, 1);
        double* out = imgOut + (long)o * s * s;
        for (int y = 0; y < s; ++y)
            for (int x = 0; x < s; ++x) out[y * s + x] = full[(long)(y + off) * N + x + off];
    }
    free(full);
}

/* Object::restorePatch's optional outputs (object.cpp:332-407), each cropped to the centred
   s x s (object.cpp:393-408) and rounded to float as ObjectData stores them (data.hpp:109-112):
   - PSF of image i (SubImage::getPSF, subimage.cpp:615-634): the OTF's own inverse transform
     (its flags, subimage.cpp:89-90: de-centre, backward, normalise by 1/N^2, re-centre the
     image), real part;
   - average PSF of object o: the per-image PSFs summed in double, rounded to float, then
     scaled by 1/imgCount;
   - convolved object (SubImage::convolveImage, subimage.hpp:110-115, and
     FourierTransform::convolve, fouriertransform.hpp:309-324): the float restored object
     (full grid, before the crop and the fitted plane) forward transformed, centred and
     normalised, times the centred OTF, de-centred, backward transformed, real part;
   - residual (subimage.hpp:118-130): the float convolved object minus the windowed, padded
     image (the inverse of imgFT), rounded to float; the same arithmetic with and without the
     convolved objects saved.
   psf: [nImg][s][s] (mask & 32) or [nObj][s][s] (mask & 64, which wins when both are set,
   config.cpp:497-500); cobj (mask & 2) and res (mask & 128): [nImg][s][s]. */
void o_state_restore_outputs(OState* st, const double* alpha, int mask, double* psf, double* cobj, double* res)
{
    const OProblem* pr = st->pr;
    const int N = st->N, N2 = st->N2, s = st->s, off = (N - s) / 2;
    const double nrm = 1.0 / N2;
    if ((mask & 32) && (mask & 64)) mask &= ~32;
    o_state_metric(st, alpha);
    double* full = malloc(sizeof(double) * N2);
    double* acc = malloc(sizeof(double) * N2);
    double* conv = malloc(sizeof(double) * N2);
    ocplx* t = malloc(sizeof(ocplx) * N2); ocplx* u = malloc(sizeof(ocplx) * N2);
    ocplx* fo = malloc(sizeof(ocplx) * N2);
    for (int o = 0; o < pr->nObj; ++o) {
        restore_full(st, o, full, 0);
        for (int k = 0; k < N2; ++k) full[k] = (float)((double)(float)full[k] * nrm);   /* od.img: float, ift<float> */
        for (int k = 0; k < N2; ++k) fo[k] = full[k];
        o_fft2(fo, N, N, -1);
        o_reorder_c(fo, N, N);
        for (int k = 0; k < N2; ++k) fo[k] *= nrm;
        memset(acc, 0, sizeof(double) * N2);
        int cnt = 0;
        for (int i = 0; i < pr->nImg; ++i) {
            if (img_obj(st, i) != o) continue;
            cnt++;
            const ocplx* otf = st->OTF + (long)i * N2;
            if (mask & (32 | 64)) {
                o_reorder_into_c(otf, N, N, t, 0, 0, 0);
                o_fft2(t, N, N, +1);
                for (int k = 0; k < N2; ++k) conv[k] = creal(t[k] * nrm);
                o_reorder_d(conv, N, N);
                if (mask & 64) for (int k = 0; k < N2; ++k) acc[k] += conv[k];
                else
                    for (int y = 0; y < s; ++y)
                        for (int x = 0; x < s; ++x)
                            psf[((long)i * s + y) * s + x] = (float)conv[(long)(y + off) * N + x + off];
            }
            if (mask & (2 | 128)) {
                for (int k = 0; k < N2; ++k) u[k] = fo[k] * otf[k];
                o_reorder_into_c(u, N, N, t, 0, 0, 0);
                o_fft2(t, N, N, +1);
                for (int k = 0; k < N2; ++k) conv[k] = (float)creal(t[k]);
                if (mask & 2)
                    for (int y = 0; y < s; ++y)
                        for (int x = 0; x < s; ++x)
                            cobj[((long)i * s + y) * s + x] = conv[(long)(y + off) * N + x + off];
                if (mask & 128) {
                    o_reorder_into_c(st->imgFT + (long)i * N2, N, N, t, 0, 0, 0);
                    o_fft2(t, N, N, +1);
                    for (int y = 0; y < s; ++y)
                        for (int x = 0; x < s; ++x) {
                            long k = (long)(y + off) * N + x + off;
                            res[((long)i * s + y) * s + x] = (float)(conv[k] - creal(t[k]) * nrm);
                        }
                }
            }
        }
        if (mask & 64)
            for (int y = 0; y < s; ++y)
                for (int x = 0; x < s; ++x)
                    psf[((long)o * s + y) * s + x] = (float)((double)(float)acc[(long)(y + off) * N + x + off] * (1.0 / cnt));
    }
    free(full); free(acc); free(conv); free(t); free(u); free(fo);
}

/* ---------- GETSTEP = steepest_descent and BFGS: GSL 1.16 steepest_descent and
   vector_bfgs (solver.cpp:355-361, 369-372; GSL is not in the reference tree, its
   published algorithms restated: multimin/steepest_descent.c, vector_bfgs.c and the
   directional minimiser they share, directional_minimize.c) ----------
   Steepest descent: from x, trial x - step g/|g|; a trial that raises f shrinks the step
   by tol and is retried; on success the step doubles (or, after a shrink, shrinks once
   more).  vector_bfgs: along -+p (downhill), a trial at the current step; if it lowers f
   the step doubles, else a parabolic guess from (0, fa, slope) and (stepc, fc) finds a
   point below fa (halving towards 0), then a safeguarded parabolic / golden-section
   search (at most 10 trials) stops when |p.g lambda/|g|| < tol; the new direction is the
   BFGS update p' = g1 - A dx - B dg, restarted to g every n iterations.
   Parity with GSL's arithmetic is unpinned; the tests pin the mathematics. */

struct OStep {
    const OFdf* fn;
    int n, kind, iter;                 /* kind 1: steepest descent, 3: vector_bfgs */
    double step, maxStep, tol, pnorm, g0norm;
    double *x1, *g1, *dx1, *x2, *p, *x0, *g0, *dx0, *dg0;
};

OStep* o_step_new(int n, int kind)
{
    OStep* w = calloc(1, sizeof(OStep));
    w->n = n;
    w->kind = kind;
    double** v[] = {&w->x1, &w->g1, &w->dx1, &w->x2, &w->p, &w->x0, &w->g0, &w->dx0, &w->dg0};
    for (size_t i = 0; i < sizeof v / sizeof v[0]; ++i) *v[i] = calloc(n ? n : 1, sizeof(double));
    return w;
}

void o_step_free(OStep* w)
{
    if (!w) return;
    free(w->x1); free(w->g1); free(w->dx1); free(w->x2); free(w->p); free(w->x0); free(w->g0);
    free(w->dx0); free(w->dg0); free(w);
}

static int o_vec_equal(const double* a, const double* b, int n)
{
    for (int k = 0; k < n; ++k) if (a[k] != b[k]) return 0;
    return 1;
}

/* x1 = x + dx, dx = -step lambda p */
static void o_take_step(const double* x, const double* p, double step, double lambda, double* x1, double* dx, int n)
{
    for (int k = 0; k < n; ++k) dx[k] = 0.0 + (-step * lambda) * p[k];
    for (int k = 0; k < n; ++k) x1[k] = x[k] + 1.0 * dx[k];
}

void o_step_set(OStep* w, const OFdf* fn, const double* x, double* f, double* g, double stepSize, double tol)
{
    const int n = w->n;
    w->fn = fn;
    w->iter = 0;
    w->step = stepSize;
    w->maxStep = stepSize;
    w->tol = tol;
    fn->fdf(fn->ctx, x, f, g);
    if (w->kind == 3) {                 /* the first direction is the gradient */
        memcpy(w->x0, x, sizeof(double) * n);
        memcpy(w->p, g, sizeof(double) * n);
        memcpy(w->g0, g, sizeof(double) * n);
        w->pnorm = w->g0norm = dnrm2(g, n);
    }
}

static int o_sd_iterate(OStep* w, double* x, double* f, double* g, double* dx)
{
    const int n = w->n;
    const double f0 = *f, gnorm = dnrm2(g, n);
    double step = w->step, f1;
    int shrunk = 0;
    if (gnorm == 0.0) { memset(dx, 0, sizeof(double) * n); return GSL_ENOPROG; }
    for (;;) {
        for (int k = 0; k < n; ++k) dx[k] = 0.0 + (-step / gnorm) * g[k];
        for (int k = 0; k < n; ++k) w->x1[k] = x[k] + 1.0 * dx[k];
        if (o_vec_equal(x, w->x1, n)) return GSL_ENOPROG;
        w->fn->fdf(w->fn->ctx, w->x1, &f1, w->g1);
        if (f1 > f0) { shrunk = 1; step *= w->tol; continue; }
        break;
    }
    step = shrunk ? step * w->tol : step * 2.0;
    w->step = step;
    memcpy(x, w->x1, sizeof(double) * n);
    memcpy(g, w->g1, sizeof(double) * n);
    *f = f1;
    return GSL_SUCCESS;
}

/* a point between 0 and stepc with f below fa: parabolic guess from the slope pg at 0,
   moving towards 0 while f does not fall (directional_minimize.c: intermediate_point) */
static void o_intermediate(OStep* w, const double* x, double lambda, double pg, double stepc, double fa,
                           double fc, double* g, double* stepOut, double* fOut)
{
    const int n = w->n;
    double stepb, fb;
    for (;;) {
        const double u = fabs(pg * lambda * stepc);
        stepb = 0.5 * stepc * u / ((fc - fa) + u);
        o_take_step(x, w->p, stepb, lambda, w->x1, w->dx1, n);
        if (o_vec_equal(x, w->x1, n)) {            /* the trial does not move: give up */
            *stepOut = 0;
            *fOut = fa;
            w->fn->df(w->fn->ctx, w->x1, g);
            return;
        }
        fb = w->fn->f(w->fn->ctx, w->x1);
        if (fb >= fa && stepb > 0.0) { fc = fb; stepc = stepb; continue; }
        break;
    }
    *stepOut = stepb;
    *fOut = fb;
    w->fn->df(w->fn->ctx, w->x1, g);
}

/* safeguarded parabolic / golden-section search in (stepa, stepc) around stepb; the
   best point goes to x2 (directional_minimize.c: minimize) */
static void o_dirmin(OStep* w, const double* x, double lambda, double stepa, double stepb, double stepc,
                     double fa, double fb, double fc, double* dx2, double* g, double* stepOut, double* fOut,
                     double* gnormOut)
{
    const int n = w->n;
    double u = stepb, v = stepa, wv = stepc, fu = fb, fv = fa, fw = fc;
    double old2 = fabs(wv - v), old1 = fabs(v - u);
    memcpy(w->x2, w->x1, sizeof(double) * n);
    memcpy(dx2, w->dx1, sizeof(double) * n);
    *fOut = fb;
    *stepOut = stepb;
    *gnormOut = dnrm2(g, n);
    for (int iter = 1; iter <= 10; ++iter) {
        const double dw = wv - u, dv = v - u;
        const double e1 = (fv - fu) * dw * dw + (fu - fw) * dv * dv;
        const double e2 = 2.0 * ((fv - fu) * dw + (fu - fw) * dv);
        const double du = e2 != 0.0 ? e1 / e2 : 0.0;
        double stepm;
        if (du > 0.0 && du < stepc - stepb && fabs(du) < 0.5 * old2) stepm = u + du;
        else if (du < 0.0 && du > stepa - stepb && fabs(du) < 0.5 * old2) stepm = u + du;
        else if (stepc - stepb > stepb - stepa) stepm = 0.38 * (stepc - stepb) + stepb;
        else stepm = stepb - 0.38 * (stepb - stepa);
        o_take_step(x, w->p, stepm, lambda, w->x1, w->dx1, n);
        const double fm = w->fn->f(w->fn->ctx, w->x1);
        if (fm > fb) {
            if (fm < fv) { wv = v; v = stepm; fw = fv; fv = fm; }
            else if (fm < fw) { wv = stepm; fw = fm; }
            if (stepm < stepb) { stepa = stepm; fa = fm; }
            else { stepc = stepm; fc = fm; }
            continue;
        }
        old2 = old1;
        old1 = fabs(u - stepm);
        wv = v; v = u; u = stepm;
        fw = fv; fv = fu; fu = fm;
        memcpy(w->x2, w->x1, sizeof(double) * n);
        memcpy(dx2, w->dx1, sizeof(double) * n);
        w->fn->df(w->fn->ctx, w->x1, g);
        const double pg = o_dot(w->p, g, n), gnorm1 = dnrm2(g, n);
        *fOut = fm;
        *stepOut = stepm;
        *gnormOut = gnorm1;
        if (fabs(pg * lambda / gnorm1) < w->tol) return;
        if (stepm < stepb) { stepc = stepb; fc = fb; stepb = stepm; fb = fm; }
        else { stepa = stepb; fa = fb; stepb = stepm; fb = fm; }
    }
}

static int o_bfgs1_iterate(OStep* w, double* x, double* f, double* g, double* dx)
{
    const int n = w->n;
    const double fa = *f, stepc = w->step;
    double stepb, fb, g1norm;
    if (w->pnorm == 0.0 || w->g0norm == 0.0) { memset(dx, 0, sizeof(double) * n); return GSL_ENOPROG; }
    const double pg = o_dot(w->p, g, n), dir = pg >= 0.0 ? +1.0 : -1.0;
    o_take_step(x, w->p, stepc, dir / w->pnorm, w->x1, dx, n);
    const double fc = w->fn->f(w->fn->ctx, w->x1);
    if (fc < fa) {                                     /* the trial step is downhill: take it */
        w->step = stepc * 2.0;
        *f = fc;
        memcpy(x, w->x1, sizeof(double) * n);
        w->fn->df(w->fn->ctx, w->x1, g);
        return GSL_SUCCESS;
    }
    o_intermediate(w, x, dir / w->pnorm, pg, stepc, fa, fc, g, &stepb, &fb);
    if (stepb == 0.0) return GSL_ENOPROG;
    o_dirmin(w, x, dir / w->pnorm, 0.0, stepb, stepc, fa, fb, fc, dx, g, &w->step, f, &g1norm);
    memcpy(x, w->x2, sizeof(double) * n);
    w->iter = (w->iter + 1) % n;
    if (w->iter == 0) {                                 /* restart along the gradient */
        memcpy(w->p, g, sizeof(double) * n);
        w->pnorm = g1norm;
    } else {
        for (int k = 0; k < n; ++k) { w->dx0[k] = x[k] + (-1.0) * w->x0[k]; w->dg0[k] = g[k] + (-1.0) * w->g0[k]; }
        const double dxg = o_dot(w->dx0, g, n), dgg = o_dot(w->dg0, g, n), dxdg = o_dot(w->dx0, w->dg0, n);
        const double dgnorm = dnrm2(w->dg0, n);
        double A = 0, B = 0;
        if (dxdg != 0) {
            B = dxg / dxdg;
            A = -(1.0 + dgnorm * dgnorm / dxdg) * B + dgg / dxdg;
        }
        for (int k = 0; k < n; ++k) w->p[k] = g[k] + (-A) * w->dx0[k] + (-B) * w->dg0[k];
        w->pnorm = dnrm2(w->p, n);
    }
    memcpy(w->g0, g, sizeof(double) * n);
    memcpy(w->x0, x, sizeof(double) * n);
    w->g0norm = dnrm2(w->g0, n);
    return GSL_SUCCESS;
}

int o_step_iterate(OStep* w, double* x, double* f, double* g, double* dx)
{
    return w->kind == 1 ? o_sd_iterate(w, x, f, g, dx) : o_bfgs1_iterate(w, x, f, g, dx);
}

/* ---------- the driver: Solver::run (solver.cpp:330-580) ---------- */

int o_solve(const OProblem* pr, OResult* res)
{
    OState* st = o_state_new(pr);
    int nFree = pr->nFree, nPar = pr->nPar, K = pr->K;
    double maxWl = 0;
    for (int o = 0; o < pr->nObj; ++o) if (pr->wavelength[o] > maxWl) maxWl = pr->wavelength[o];
    init_channels(st);

    double* x = calloc(nFree, sizeof(double)); double* g = calloc(nFree, sizeof(double));
    double* dx = calloc(nFree, sizeof(double)); double* beta0 = calloc(nFree, sizeof(double));
    CGState cs; memset(&cs, 0, sizeof(cs));
    cs.x1 = calloc(nFree, sizeof(double)); cs.p = calloc(nFree, sizeof(double)); cs.prevG = calloc(nFree, sizeof(double));
    double f = 0;

    double init_step = 1E-4 * maxWl, init_tol = 1E-6;
    OBfgs2* bfgs = pr->getstep == 4 ? o_bfgs2_new(nFree) : NULL;
    OStep* other = (pr->getstep == 1 || pr->getstep == 3) ? o_step_new(nFree, pr->getstep) : NULL;
    if (pr->getstep == 3) { init_step = 1E8; init_tol = 0.01; }   /* solver.cpp:359-360 */
    OFdf fn = { st, cb_f, cb_df, cb_fdf };
    st->gradMethod = 1;
    double previousMetric = 0, thisMetric = 0, gradNorm = 0, tol = 1.0E-1;
    int failCount = 0, totalIterations = 0, maxIterations = 10, maxFails = 3, status = 0;
    memset(st->alphaOffset, 0, sizeof(double) * nPar);
    if (pr->alphaInit) memcpy(st->alphaOffset, pr->alphaInit, sizeof(double) * nPar);
    shift_and_init(st, st->alphaOffset, 1);
    res->nMetrics = 0;
    double initialMetric = my_f(st, beta0);
    if (res->metrics && res->nMetrics < pr->maxMetrics) res->metrics[res->nMetrics++] = initialMetric;
    double nTotalPixels = (double)pr->nImg * pr->s * pr->s;
    double gradScale = 1.0 / (nTotalPixels * nTotalPixels);
    int exitLoop = 0;
    for (int modeCount = pr->modeStart; !exitLoop; modeCount += pr->modeStep) {
        if (modeCount > K) modeCount = K;
        for (int m = 0; m < modeCount; ++m) st->enabled[m] = 1;
        if (modeCount == K || pr->maxIter == 0) {
            exitLoop = 1; tol = pr->FTOL; maxIterations = pr->maxIter;
        }
        /* gsl_multimin_fdfminimizer_set + the stepper's set: conjugate_rdx_set
           (conjugate_gradient.cpp:72-86) or vector_bfgs2_set */
        memcpy(x, beta0, sizeof(double) * nFree);
        memset(dx, 0, sizeof(double) * nFree);
        if (bfgs) {
            o_bfgs2_set(bfgs, &fn, x, &f, g, init_step, init_tol);
        } else if (other) {
            o_step_set(other, &fn, x, &f, g, init_step, init_tol);
        } else {
            cs.iter = 0; cs.step = init_step; cs.tol = init_tol;
            f = my_f(st, x);
        }
        int iter = 0, successCount = 0, done = 0;
        while ((!done || iter < pr->minIter) && iter < maxIterations) {
            iter++;
            status = bfgs ? o_bfgs2_iterate(bfgs, x, &f, g, dx)
                   : other ? o_step_iterate(other, x, &f, g, dx) : cg_iterate(&fn, &cs, x, &f, g, dx, nFree);
            if (failCount > maxFails) { status = GSL_FAILURE; exitLoop = 1; done = 1; break; }
            st->gradMethod = 2;
            if (status == GSL_ENOPROG) { failCount++; break; }
            thisMetric = f;
            if (!isfinite(thisMetric)) { failCount++; break; }
            if (res->metrics && res->nMetrics < pr->maxMetrics) res->metrics[res->nMetrics++] = thisMetric;
            gradNorm = dnrm2(g, nFree) * gradScale;
            if (!isfinite(gradNorm)) { failCount++; break; }
            if (gradNorm < pr->FTOL) { exitLoop = 1; done = 1; break; }
            if (iter > 1) {
                double rel = 2.0 * fabs(thisMetric - previousMetric) / (fabs(thisMetric) + fabs(previousMetric) + pr->EPS);
                if (rel < tol) {
                    if (++successCount >= pr->nDoneIter) { successCount = 0; done = 1; }
                } else successCount = 0;
            }
            previousMetric = thisMetric;
        }
        thisMetric = my_f(st, x);
        totalIterations += iter;
        if (!exitLoop) {
            shift_and_init(st, st->alpha, 0);
            memcpy(st->alphaOffset, st->alpha, sizeof(double) * nPar);
        }
    }
    res->initialMetric = initialMetric;
    res->finalMetric = thisMetric;
    res->totalIterations = totalIterations;
    res->status = status;
    res->failCount = failCount;
    res->nF = st->nF; res->nDF = st->nDF;
    memcpy(res->alpha, st->alpha, sizeof(double) * nPar);
    if (res->shifts) memcpy(res->shifts, st->curShift, sizeof(int) * 2 * pr->nImg);
    if (res->img) restore_objects(st, res->img);
    free(x); free(g); free(dx); free(beta0); free(cs.x1); free(cs.p); free(cs.prevG);
    o_bfgs2_free(bfgs);
    o_step_free(other);
    o_state_free(st);
    return 0;
}

/* ---------- stage-wise access for parity tests ---------- */

void o_state_init(OState* st)
{
    init_channels(st);
    shift_and_init(st, st->alphaOffset, 1);
}

double o_state_metric(OState* st, const double* alpha)
{
    memcpy(st->alpha, alpha, sizeof(double) * st->pr->nPar);
    apply_alpha(st, st->alpha);
    return metric_now(st);
}

static void grad_common(OState* st, const double* alpha, int nEnabled, double* g, int method)
{
    memcpy(st->alpha, alpha, sizeof(double) * st->pr->nPar);
    for (int m = 0; m < st->K; ++m) st->enabled[m] = m < nEnabled;
    apply_alpha(st, st->alpha);
    accumulate_pq(st);
    calc_helpers(st);
    st->gradMethod = method;
    gradient_now(st);
    memcpy(g, st->gradAlpha, sizeof(double) * st->pr->nPar);
}

void o_state_grad_vogel(OState* st, const double* alpha, int nEnabled, double* g) { grad_common(st, alpha, nEnabled, g, 2); }
void o_state_grad_fd(OState* st, const double* alpha, int nEnabled, double* g) { grad_common(st, alpha, nEnabled, g, 1); }

void o_state_restore_full(OState* st, const double* alpha, int o, double* full)
{
    o_state_metric(st, alpha);
    restore_full(st, o, full, 1);
}

void o_state_restore(OState* st, const double* alpha, double* img)
{
    o_state_metric(st, alpha);
    restore_objects(st, img);
}

void o_state_get(OState* st, const char* what, double* out)
{
    const OProblem* pr = st->pr;
    if (!strcmp(what, "otf")) memcpy(out, st->OTF, sizeof(ocplx) * pr->nImg * st->N2);
    else if (!strcmp(what, "imgft")) memcpy(out, st->imgFT, sizeof(ocplx) * pr->nImg * st->N2);
    else if (!strcmp(what, "ftsum")) memcpy(out, st->ftSum, sizeof(double) * pr->nObj * st->N2);
    else if (!strcmp(what, "P")) memcpy(out, st->Pc, sizeof(ocplx) * pr->nObj * st->N2);
    else if (!strcmp(what, "Q")) memcpy(out, st->Qc, sizeof(double) * pr->nObj * st->N2);
    else if (!strcmp(what, "phi")) memcpy(out, st->phi, sizeof(double) * pr->nImg * st->P2);
    else if (!strcmp(what, "vogel")) memcpy(out, st->vogel, sizeof(double) * pr->nImg * st->P2);
    else if (!strcmp(what, "phich")) memcpy(out, st->phiCh, sizeof(double) * pr->nCh * st->P2);
    else if (!strcmp(what, "window")) memcpy(out, st->window, sizeof(double) * st->s2);
    else if (!strcmp(what, "noisewindow")) memcpy(out, st->noiseWindow, sizeof(double) * st->s2);
}

double o_state_scalar(OState* st, const char* what, int idx)
{
    if (!strcmp(what, "gamma")) return st->regGamma[idx];
    if (!strcmp(what, "otfnorm")) return st->otfNorm[idx];
    if (!strcmp(what, "area")) return st->area[idx];
    if (!strcmp(what, "noise")) return st->noise[idx];
    if (!strcmp(what, "stddev")) return st->stddev[idx];
    if (!strcmp(what, "mean")) return st->mean[idx];
    if (!strcmp(what, "npupil")) return st->nPS[idx];
    if (!strcmp(what, "notf")) return st->nOS[idx];
    if (!strcmp(what, "nf")) return st->nF;
    if (!strcmp(what, "ndf")) return st->nDF;
    return NAN;
}
