/*
 * CPU FP64 oracle for the MOMFBD patch solver: TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithm (ISP-SST/redux,
 * src/lib/momfbd/{solver,subimage,object,channel}.cpp, src/lib/util/gsl/
 * conjugate_gradient.cpp, src/lib/math/helpers.cpp, src/lib/image/
 * {fouriertransform,utils}.cpp).  It is used by tests/ as the checker, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg -- never by the
 * product path.  Every function cites the reference file:line it follows.
 *
 * Conventions (FFTW): forward transform e^{-i}, backward e^{+i}, both
 * unnormalised.  Arrays are row-major, y slow / x fast.
 */
#ifndef MFBD_ORACLE_H
#define MFBD_ORACLE_H

#include <complex.h>
#include <stdint.h>

typedef double complex ocplx;

/* ---- Fourier helpers (fft.c) ---- */
void o_fft2(ocplx* a, int ny, int nx, int sign);
void o_reorder_into_c(const ocplx* in, int iny, int inx, ocplx* out, int ony, int onx, int center);
void o_reorder_into_d(const double* in, int iny, int inx, double* out, int ony, int onx, int center);
void o_reorder_c(ocplx* a, int ny, int nx);
void o_reorder_d(double* a, int ny, int nx);
void o_autocorrelate_c(ocplx* a, int ny, int nx, int center);
void o_autocorrelate_d(double* a, int ny, int nx, int center);
double o_noise(const ocplx* ft, int ny, int nx, int mask, double limit);

/* ---- solver problem description (solver.c) ---- */
typedef struct {
    int s, P, N, K;               /* patch side, pupil grid, OTF grid (=2P), modes */
    int nObj, nCh, nImg;
    /* objects */
    const double* weight;         /* [nObj] */
    const double* wavelength;     /* [nObj] */
    const double* pupil;          /* [nObj][P*P]                                   */
    const double* modes;          /* [nObj][K][P*P]  scaled to RMS 1/lambda           */
    const int*    tilt;           /* [nObj][2]  tiltMode.x, tiltMode.y (mode index)  */
    const double* jac;            /* [nObj][8]  J_T_xy[2][2], J_xy_T[2][2]            */
    const double* freqCutoff;     /* [nObj] */
    /* channels */
    const int*    chObj;          /* [nCh] */
    const double* phiFixed;       /* [nCh][P*P] diversity phase */
    const double* resOffset;      /* [nCh][2]  residual offset (x, y) in pixels */
    const double* noiseFudge;     /* [nCh] */
    /* images, ordered object -> channel -> frame */
    const int*    imgCh;          /* [nImg] */
    const float*  data;           /* [nImg][dataY][dataX] */
    int dataY, dataX;
    int startY, startX;           /* patch start inside the data block */
    /* nullspace, sparse, entries sorted by (row, col) */
    int nPar, nFree, nnz;
    const int* nsRow; const int* nsCol; const double* nsVal;
    /* settings (config.cpp:629-639 defaults unless changed) */
    double FTOL, EPS, graddiffStep, regAlpha;
    int minIter, maxIter, nDoneIter, modeStart, modeStep;
    int apodSize;                 /* <0: patchSize/8 */
    int noFilter;
    const double* alphaInit;      /* [nImg][K] or NULL */
    int maxMetrics;
    int getstep;                  /* 1: steepest_descent, 0/2: conjugate_rdx, 3: vector_bfgs, 4: vector_bfgs2 (BFGS_inv) */
    /* Sensitivity experiments only (0 in every parity use): each metric value the solver's
       minimiser sees is multiplied by 1 + metricNoise*u, u uniform in [-1, 1) from a
       splitmix64 stream seeded by noiseSeed.  It models an implementation whose metric
       differs from this one at that relative level (another summation order, FP32). */
    double metricNoise;
    unsigned long long noiseSeed;
} OProblem;

typedef struct {
    double* alpha;                /* [nImg][K] */
    double* img;                  /* [nObj][s][s] restored objects (NULL: skip) */
    double* metrics;              /* [maxMetrics] metric history (NULL: skip) */
    int nMetrics;
    double initialMetric, finalMetric;
    int totalIterations, nF, nDF, status, failCount;
    int* shifts;                  /* [nImg][2] final integer shift (x, y), may be NULL */
} OResult;

/* Full Solver::run restatement on one patch.  Returns 0 on success. */
int o_solve(const OProblem* pr, OResult* res);

/* Stateful evaluation handle for parity checks of individual stages. */
typedef struct OState OState;
OState* o_state_new(const OProblem* pr);
void o_state_free(OState* st);
/* SubImage::initialize for every image (doReset=1 as on the first call). */
void o_state_init(OState* st);
/* metric L(alpha) as Solver::metric after applyAlpha (solver.cpp:639,763) */
double o_state_metric(OState* st, const double* alpha);
/* Vogel gradient d/dalpha (solver.cpp:860-894) at alpha, enabled = first nEnabled modes */
void o_state_grad_vogel(OState* st, const double* alpha, int nEnabled, double* g);
/* finite-difference gradient (subimage.cpp:281-292) at alpha */
void o_state_grad_fd(OState* st, const double* alpha, int nEnabled, double* g);
/* restored objects at alpha (object.cpp:278-443), img [nObj][s][s] */
void o_state_restore(OState* st, const double* alpha, double* img);
/* optional restoration outputs at alpha (object.cpp:332-407): mask bits 2 convolved objects,
   32 PSF per image, 64 PSF averaged per object, 128 residuals; each [.][s][s] */
void o_state_restore_outputs(OState* st, const double* alpha, int mask, double* psf, double* cobj, double* res);
/* export per-image/per-object arrays after the last evaluation */
void o_state_get(OState* st, const char* what, double* out);
double o_state_scalar(OState* st, const char* what, int idx);



/* ---- GSL 1.16 vector_bfgs2 restatement (solver.c), generic over f / df / fdf ---- */
typedef struct {
    void* ctx;
    double (*f)(void* ctx, const double* x);
    void (*df)(void* ctx, const double* x, double* g);
    void (*fdf)(void* ctx, const double* x, double* f, double* g);
} OFdf;
typedef struct OBfgs2 OBfgs2;
OBfgs2* o_bfgs2_new(int n);
void o_bfgs2_free(OBfgs2* w);
void o_bfgs2_set(OBfgs2* w, const OFdf* fn, const double* x, double* f, double* g, double stepSize, double tol);
int o_bfgs2_iterate(OBfgs2* w, double* x, double* f, double* g, double* dx);
/* GSL 1.16 steepest_descent (kind 1) and vector_bfgs (kind 3) restatements (solver.c) */
typedef struct OStep OStep;
OStep* o_step_new(int n, int kind);
void o_step_free(OStep* w);
void o_step_set(OStep* w, const OFdf* fn, const double* x, double* f, double* g, double stepSize, double tol);
int o_step_iterate(OStep* w, double* x, double* f, double* g, double* dx);

/* ---- conjugate_rdx / bracket / brent restatements (solver.c) on caller functions ---- */
typedef struct OCg OCg;
OCg* o_cg_new(int n);
void o_cg_free(OCg* w);
void o_cg_set(OCg* w, const OFdf* fn, const double* x, double* f, double stepSize, double tol);
int o_cg_iterate(OCg* w, double* x, double* f, double* g, double* dx);
void o_cg_direction(const OCg* w, double* p);
void o_bracket(const OFdf* fn, double* a, double* b, double* mid, double* fa, double* fb, double* fmid, int limit);
void o_brent(const OFdf* fn, double ax, double bx, double* x, double* fx, double tol, int limit);
/* Object::restorePatch pieces (solver.c): the ScharmerFilter restatement on a centred grid,
   and object o's restored object on the full N x N grid before the crop (normalised
   inverse transform) at alpha */
void o_scharmer_filter(ocplx* data, const double* q2, int ny, int nx, double noisePower, double cutoff);
void o_state_restore_full(OState* st, const double* alpha, int o, double* full);

/* ---- mosaic of restored patches (mosaic.c; object.cpp:1160-1207) ---- */
void o_blend_window(double* w, int nRows, int nCols, int blend, int margin);
int o_mosaic(int n, int s, const float* patches, const int* posY, const int* posX, int clip,
             float* img, int* rows, int* cols, int* r0, int* c0);

#endif
