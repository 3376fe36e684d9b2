/*
 * momfbd_hip.h -- C-ABI of the MI355X MOMFBD patch solver (libmomfbd_hip.so).
 *
 * Drop-in boundary: the reference's per-patch solve, Solver::run(PatchData::Ptr)
 * (ISP-SST/redux src/lib/momfbd/solver.cpp:330-580), called by MomfbdJob::run
 * (src/lib/momfbd/momfbdjob.cpp:576-590) once per patch after data->initPatch().
 * A context plays the role of Solver + thread::TmpStorage (include/redux/momfbd/
 * solver.hpp:33-179): it owns every device buffer and one HIP stream.  The caller
 * owns all host arrays; they are read during the call and never retained.
 *
 * Plain C: pointers, sizes and int status codes.  No exceptions cross the ABI.
 * A context is single-threaded-use; separate contexts may run on separate threads.
 */
#ifndef MOMFBD_HIP_H
#define MOMFBD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference throws std::logic_error or returns GSL codes) */
#define MFBD_OK 0
#define MFBD_EINVAL 1        /* bad argument / inconsistent sizes          */
#define MFBD_ENOMEM 2        /* device or host allocation failed           */
#define MFBD_EDEVICE 3       /* HIP runtime error                          */
#define MFBD_ENONFINITE 4    /* non-finite metric or gradient              */
#define MFBD_ESTATE 5        /* call out of order (e.g. solve before load) */
#define MFBD_ESCRATCH 6      /* larger than the fixed scratch of an allocation-free call      */

/* mode bases (GlobalCfg::modeBasis, config.cpp:629) */
#define MFBD_BASIS_ZERNIKE 0
#define MFBD_BASIS_KL 1

/* arithmetic of the hot path */
#define MFBD_FP64 0          /* all stages in FP64 (the reference's arithmetic)              */
#define MFBD_FP32 1          /* FP32 phase/FFT/OTF, FP64 P/Q/metric reductions and CG scalars */
#define MFBD_MIXED 2         /* as MFBD_FP32 with the mode expansion phi = B alpha on BF16 matrix
                                cores (BF16 basis, alpha as two BF16 parts, FP32 accumulation) */

/* gradient forms for mfbd_eval_grad */
#define MFBD_GRAD_FD 1       /* SubImage::gradientFiniteDifference (subimage.cpp:281-292)   */
#define MFBD_GRAD_VOGEL 2    /* SubImage::gradientVogel (subimage.cpp:312-326, 348-390)      */
#define MFBD_GRAD_ADJOINT 3  /* exact dL/dalpha (adjoint of the OTF chain); no reference counterpart */

/* per-patch exit status, mirrors solver.cpp:470-516 */
#define MFBD_PATCH_CONVERGED 0   /* relative-change or gradient-norm criterion met        */
#define MFBD_PATCH_MAXITER 1     /* last stage stopped at MAX_ITER                        */
#define MFBD_PATCH_FAILED 2      /* more than 3 failures (GSL_FAILURE path)               */
#define MFBD_PATCH_NO_PROGRESS 3 /* last stage ended by GSL_ENOPROG (solver.cpp:480-483)  */
#define MFBD_PATCH_NONFINITE 4   /* last stage ended by a non-finite metric or gradient
                                    (solver.cpp:488-499)                                  */

typedef struct {
    int s;              /* patch side, NUM_POINTS (Object::patchSize)                   */
    int P;              /* pupil grid side, pupilPixels; the OTF grid is N = 2P         */
    int K;              /* number of modes (job.nModes)                                 */
    int nObj;           /* objects per patch                                            */
    int nCh;            /* channels per patch (all objects)                             */
    int nImg;           /* images per patch, ordered object -> channel -> frame         */
    int dataY, dataX;   /* raw data block per image (patch plus shift margin)           */
    int startY, startX; /* patch start inside the block (ChannelData::patchStart)       */
} mfbd_geometry;

typedef struct {        /* GlobalCfg solver knobs, defaults in config.cpp:629-639 */
    double FTOL, EPS, graddiffStep, regAlpha;
    int minIter, maxIter, nDoneIter, modeStart, modeStep;
    int apodSize;       /* APODIZATION, < 0: patchSize/8 (solver.cpp:117-126)          */
    int noFilter;       /* NO_FILTER: skip the Scharmer filter in restoration           */
    int precision;      /* MFBD_FP64, MFBD_FP32 or MFBD_MIXED                          */
    int maxBatch;       /* patches solved concurrently (1 = the reference's one-at-a-time) */
    int gradient;       /* MFBD_GRAD_SCHEDULE_REFERENCE: finite differences for the first
                           gradient, Vogel after (solver.cpp:385,478); or MFBD_GRAD_ADJOINT
                           for the exact dL/dalpha throughout                              */
    int getstep;        /* GETSTEP: MFBD_GETSTEP_STEEPEST, _CONJUGATE, _BFGS or _BFGS_INV
                           (GlobalCfg::getstepMethod); 0 is the conjugate gradient */
    int linePhases;     /* 1: the conjugate-gradient line search's probes take their phases as
                           base + step * direction, both computed once per line
                           (Solver::my_precalc / metricAt, solver.cpp:246-327, which the
                           reference leaves disabled), instead of the mode GEMM per probe;
                           equal in exact arithmetic, the phases differ by rounding          */
    int lineRounds;     /* > 0 (with linePhases and the conjugate gradient): each patch's line
                           search (math::bracket + math::brent as conjugate_rdx_iterate calls
                           them, helpers.cpp:58-216) runs on the device, one captured HIP
                           graph per round of probes, and the host reads the lines' states
                           back every lineRounds rounds; the same probes and result as the
                           host's line search.  0: the host steps the line search.  Forms
                           without a fused transform (N = 256) keep the host's.            */
    int lineGraph;      /* with lineRounds: 1 replays one captured HIP graph per round of
                           probes (launched one by one anyway while kernel events time each
                           launch, MFBD_KERNEL_TIMING), 0 launches the round's kernels one by
                           one (the same kernels and results)                               */
} mfbd_params;

#define MFBD_GRAD_SCHEDULE_REFERENCE 0
/* steppers, numbered as the reference's getstepMethod (config.cpp:37, 102-114) */
#define MFBD_GETSTEP_STEEPEST 1    /* GSM_SDSC: GSL steepest_descent */
#define MFBD_GETSTEP_CONJUGATE 2   /* GSM_CNJG: conjugate_rdx (conjugate_gradient.cpp:128-256)   */
#define MFBD_GETSTEP_BFGS 3        /* GSM_BFGS: GSL vector_bfgs (init_step 1e8, tol 0.01) */
#define MFBD_GETSTEP_BFGS_INV 4    /* GSM_BFGS_inv, the reference default: GSL vector_bfgs2      */
/* saveMask bits of mfbd_restore_outputs (SaveFlags, config.hpp:31-33) */
#define MFBD_SAVE_COBJ 2
#define MFBD_SAVE_PSF 32
#define MFBD_SAVE_PSF_AVG 64
#define MFBD_SAVE_RESIDUAL 128

typedef struct {
    int status;             /* MFBD_PATCH_*                                  */
    int iterations;         /* total CG iterations over all mode stages      */
    int nF, nDF;            /* metric / gradient evaluations                 */
    int failCount;
    double initialMetric, finalMetric;
} mfbd_result;

typedef struct mfbd_ctx mfbd_ctx;

/* ---- context (Solver::Solver + Solver::init, solver.cpp:64-183) ---- */
int mfbd_ctx_create(const mfbd_geometry* geo, const mfbd_params* prm, int device, mfbd_ctx** out);
int mfbd_ctx_destroy(mfbd_ctx* ctx);
const char* mfbd_last_error(const mfbd_ctx* ctx);

/* Patch structure: chObj[nCh] object of each channel, imgCh[nImg] channel of each
 * image; objWeight/objWavelength/freqCutoff per object (Object::weight,
 * wavelength, frequencyCutoff; object.cpp:206-268). */
int mfbd_set_structure(mfbd_ctx* ctx, const int* chObj, const int* imgCh, const double* objWeight,
                       const double* objWavelength, const double* freqCutoff);

/* Pupil [P*P] (Pupil, pupil.cpp:221-277), scaled modes [nObj][K][P*P]
 * (ModeSet via GlobalData::get(info, 1/lambda), data.cpp:658-686), tilt mode
 * indices [nObj][2] and Jacobians [nObj][8] = J_T_xy (2x2), J_xy_T (2x2)
 * (ModeSet::measureJacobian, modes.cpp:464-512).  Replaces Object::initProcessing's
 * pupil/mode fetch. */
int mfbd_set_basis(mfbd_ctx* ctx, const double* pupil, const double* modes, const int* tiltXY, const double* jac);

/* Sparse nullspace N (nPar x nFree), entries sorted by (row, col): alpha = alpha0 + N beta,
 * g_beta = N^T g_alpha (Constraints::apply/reverse, constraints.hpp:128-159). */
int mfbd_set_constraints(mfbd_ctx* ctx, int nPar, int nFree, int nnz, const int* rows, const int* cols,
                         const double* vals);

/* Load n patches (replaces PatchData::initPatch + Channel::initPatch,
 * channel.cpp:846-953): data [n][nImg][dataY][dataX] float32, phiFixed [nCh][P*P]
 * diversity phases, resOffset [n][nCh][2] residual (x,y) offsets, noiseFudge [nCh].
 * Frames are taken as already normalised (ChannelData::images after initPatch). */
int mfbd_load_patches(mfbd_ctx* ctx, int n, const float* data, const double* phiFixed, const double* resOffset,
                      const double* noiseFudge);

/* Ragged batches (uneven mode and frame counts in one context): patch `patch` of the last
 * load solves for its first nModes modes only, has only the images with imgOn[j] != 0
 * ([nImg]; at least one per object; the frames of the others are ignored) and its own null
 * space N ((row, col, value) triplets as in mfbd_set_constraints, rows only on its images
 * and modes).  The reference runs such patches as separate jobs of their own geometry
 * (object.cpp:789-790); within a context of the largest K and frame counts this gives the
 * same results, since a missing image has a zero OTF and a missing mode a zero alpha.
 * Call after mfbd_load_patches and before the solve; a new load resets every patch. */
int mfbd_set_patch_shape(mfbd_ctx* ctx, int patch, int nModes, const unsigned char* imgOn, int nFree, int nnz,
                         const int* rows, const int* cols, const double* vals);

/* Solver::run for every loaded patch (batched maxBatch at a time).  alpha0
 * [n][nImg][K] warm start (Solver::loadInit) or NULL; results[n]; alphaOut
 * [n][nImg][K]; imgOut [n][nObj][s][s] restored objects (Object::restorePatch,
 * object.cpp:278-443), either output may be NULL. */
int mfbd_solve(mfbd_ctx* ctx, const double* alpha0, mfbd_result* results, double* alphaOut, float* imgOut);

/* Solve the loaded patches of n contexts at once, one host thread per context, so their
 * launches overlap on the device: the grid classes of an uneven batch (patches of another
 * size need a context of their own geometry).  Arguments per context as for mfbd_solve
 * (alpha0 or imgOut may be NULL as a whole or per context); each context's results are
 * those mfbd_solve gives it alone.  Returns the first failing context's code. */
int mfbd_solve_group(int n, mfbd_ctx* const* ctxs, const double* const* alpha0, mfbd_result* const* results,
                     double* const* alphaOut, float* const* imgOut);

/* ---- stage-wise entry points (parity checks, autograd harness) ----
 * They evaluate on the frames' current state: the first initialisation at zero shifts
 * (made on first use, or again by mfbd_init_patches), or, after mfbd_solve, the state the
 * solve left (each image's final integer shift and tilt offset), so that the restoration
 * of a patch's alphaOut reproduces its imgOut. */
/* First SubImage::initialize (doReset) of every loaded patch at zero alpha
 * (Solver::shiftAndInit(alpha, true), solver.cpp:583-605). */
int mfbd_init_patches(mfbd_ctx* ctx);
/* L(alpha) for n loaded patches: Solver::applyAlpha + Solver::metric
 * (solver.cpp:639-657, 763-817).  alpha [n][nImg][K], L [n]. */
int mfbd_eval_metric(mfbd_ctx* ctx, int n, const int* patchIdx, const double* alpha, double* L);
/* Gradient w.r.t. alpha of the first nEnabled modes (others 0): calcPQ + gradient
 * (solver.cpp:820-894) with method MFBD_GRAD_*.  gAlpha [n][nImg][K]. */
int mfbd_eval_grad(mfbd_ctx* ctx, int n, const int* patchIdx, const double* alpha, int nEnabled, int method,
                   double* gAlpha);
/* Restored objects at alpha: Object::restorePatch.  img [n][nObj][s][s]. */
int mfbd_restore(mfbd_ctx* ctx, int n, const int* patchIdx, const double* alpha, float* img);
/* Optional restoration outputs of n patches at their alphas ([n][nImg][K], as for
 * mfbd_restore): the PSF, convolved-object and residual blocks of Object::restorePatch
 * (object.cpp:332-407), selected by saveMask bits (the reference's SaveFlags,
 * config.hpp:31-33; other bits are ignored, and PSF_AVG wins over PSF as in
 * config.cpp:497-500).  Each image is cropped to the centred s x s; per patch:
 *   MFBD_SAVE_PSF       psf  [n][nImg][s][s]  PSF of each image (SubImage::getPSF)
 *   MFBD_SAVE_PSF_AVG   psf  [n][nObj][s][s]  mean PSF of each object's images
 *   MFBD_SAVE_COBJ      cobj [n][nImg][s][s]  restored object convolved with the image's OTF
 *   MFBD_SAVE_RESIDUAL  res  [n][nImg][s][s]  convolved object minus the windowed frame
 * Images a patch lacks (mfbd_set_patch_shape) are zero.  Called after mfbd_solve with the
 * patches' alphaOut rows, it gives what restorePatch saves with the solution. */
int mfbd_restore_outputs(mfbd_ctx* ctx, int n, const int* patchIdx, const double* alpha, int saveMask, float* psf,
                         float* cobj, float* res);
/* Diagnostics: copy a device array for patch p ("S" OTF on support, "D", "P", "Q" (as
 * the last gradient evaluation or restoration left them), "ftsum", "phi") or host state ("gamma_host", "noise_host", "otfnorm", "window",
 * "shifts" [nImg][2] integer cut-out shifts (x, y), "tiltoff" [nImg][2], "sizes" = pupil
 * support, OTF support and stored (Hermitian half) OTF support point counts) in double;
 * returns the element count via *count. */
int mfbd_debug_get(mfbd_ctx* ctx, const char* what, int p, double* out, int* count);
/* Per-kernel timing (environment MFBD_KERNEL_TIMING set when the patches were
 * loaded): name[i], summed milliseconds, launches and images processed, accumulated
 * over all calls; returns the number of entries (<= cap). */
int mfbd_kernel_stats(mfbd_ctx* ctx, int cap, char (*names)[48], double* ms, int* launches, long long* units);

/* ---- after the solve: mosaic of restored patches (Object::doMozaic, object.cpp:1160-1207,
 *      which the ANA/FITS writers call, object.cpp:1213-1310) ---- */
/* Untrimmed extent of the mosaic of patches at (posY, posX): maxPosY + patchSize + 1 rows,
 * maxPosX + patchSize + 1 columns (object.cpp:1187-1188). */
int mfbd_mosaic_size(int nPatches, int patchSize, const int* posY, const int* posX, int* imgRows, int* imgCols);
/* redux::image::mozaic (utils.hpp:440-502) on the device: patches [n][patchSize][patchSize]
 * weighted by apodizeInPlace(ones, blend, blend, margin, margin) with margin = patchSize/8,
 * blend = (patchSize - 2 margin)/3, FP64 sums divided where the weight sum exceeds 1e-6;
 * with clip (RF_NO_CLIP unset) img_trim(1e-15) drops empty border rows and columns
 * (utils.hpp:521-553).  img must hold the untrimmed extent; on return it holds
 * imgRows x imgCols, the window starting at (firstRow, firstCol) of the untrimmed mosaic. */
int mfbd_mosaic(int nPatches, int patchSize, const float* patches, const int* posY, const int* posX, int clip,
                int device, float* img, int* imgRows, int* imgCols, int* firstRow, int* firstCol);

/* mfbd_mosaic without any device or host allocation, for callers that cannot allocate (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  No counterpart
 * in the reference.  Same arguments and results as mfbd_mosaic (bit for bit); the patches,
 * window, positions and the untrimmed mosaic are staged in a fixed 32 MiB device array that
 * each GPU holds until the process ends.  Limits (MFBD_ESCRATCH beyond them): patchSize at
 * most 2048, and 8 (2 patchSize + patchSize^2) + 8 nPatches + 4 nPatches patchSize^2
 * + 4 rows cols bytes (each of the six regions rounded up to 256 bytes; rows x cols the
 * untrimmed extent of mfbd_mosaic_size) at most 32 MiB: e.g. 256 patches of 88^2 into a
 * 1024^2 mosaic take 12.2 MiB.  All calls on one GPU share the array, so calls that
 * overlap on one GPU (from other host threads) race. */
int mfbd_mosaic_noalloc(int nPatches, int patchSize, const float* patches, const int* posY, const int* posX, int clip,
                        int device, float* img, int* imgRows, int* imgCols, int* firstRow, int* firstCol);

/* ---- output files (Object::writeAna / writeFits, object.cpp:1212-1334) ---- */
/* ANA (.f0) file, uncompressed (Ana::write, fileana.cpp:150-225, 498-560): datyp 0..4 =
 * int8, int16, int32, float32, float64; dims fastest first (columns, rows, ...); text is
 * the header text (NULL for none; more than 255 characters take further 512-byte blocks,
 * up to 7935).  mfbd_write_ana_compressed writes the compressed form. */
int mfbd_write_ana(const char* path, int datyp, int ndim, const int* dims, const void* data, const char* text);
/* ANA file of 2-D int16 (datyp 1) or int32 (datyp 2) data [ny][nx] with the crunch
 * compression at bit slice `slice` (Ana::write with compression, fileana.cpp:498-560;
 * anacrunch / anacrunch32, anacompress.cpp), stored raw when that is not smaller. */
int mfbd_write_ana_compressed(const char* path, int datyp, int nx, int ny, const void* data, const char* text,
                              int slice);
/* FITS primary image (Fits::write, filefits.cpp:1756-1818): BITPIX 8, 16, 32, 64, -32 or
 * -64; dims fastest first (NAXIS1 = columns); data in host order, written big-endian. */
int mfbd_write_fits(const char* path, int bitpix, int ndim, const int* dims, const void* data);
/* The .momfbd result container (FileMomfbd::write, filemomfbd.cpp:391-564 with
 * PatchInfo::write, filled as Object::writeMomfbd does, object.cpp:1336-1565): every
 * number in host byte order behind a one-byte endian flag.  Layout written, in order:
 *   int8 big-endian flag; int32 length + version string (with its NUL); int32 length +
 *   modified-time string (ISO "YYYY-MM-DDTHH:MM:SS", UTC); int32 length + time string;
 *   int32 length + date string; int32 region[4] (1-based first x, last x, first y, last y,
 *   the union of the patches'); int8 hasModes [float pix2cf, cf2pix; int32 nPH, nModes;
 *   float pupil[nPH][nPH]; per mode: uint8 type 1 and float mode[nPH][nPH]]; int32 nCh;
 *   int16 clipStartX[nCh], clipEndX[nCh], clipStartY[nCh], clipEndY[nCh]; int32 nPatchesY,
 *   nPatchesX, nPoints; per patch (y slow): int32 region[4], offx, offy (0), nCh, nim[nCh],
 *   dx[nCh], dy[nCh]; int8 1 and float img[ny][nx]; int32 npsf + psf; int32 nobj + convolved
 *   objects; int32 nres + residuals; int32 nalpha [+ int32 nm + float alpha[nalpha][nm]];
 *   int32 ndiv [+ int32 nphx, nphy + per diversity: uint8 type 0 and float div[nphy][nphx]];
 *   then, with names, int32 nFileNames + per name int32 length + the name with its NUL.
 * The version must be 20190408 or later (the layout above, which the reference writes and
 * reads from that version on).  Arrays absent (NULL or zero count) are written as counts 0. */
typedef struct {
    const char* version;        /* e.g. "20240612.1" (reduxCommitTime's "%4d%02d%02d.%01d")        */
    const char* modified;       /* ISO time, or NULL for the current UTC time                        */
    const char* time;           /* observation time of day ("N/A" when unknown)                      */
    const char* date;           /* observation date                                                  */
    int nCh;                    /* channels of the object                                            */
    const int16_t* clip;        /* [4][nCh]: clipStartX, clipEndX, clipStartY, clipEndY (1-based)    */
    int nPH, nModes;            /* pupil grid side and modes; nPH = 0 or pupil NULL: no modes block   */
    float pix2cf, cf2pix;       /* ModeSet::shiftToAlpha / alphaToShift of (1, 1), averaged          */
    const float* pupil;         /* [nPH][nPH]                                                        */
    const float* modes;         /* [nModes][nPH][nPH]                                                */
    int nPatchesY, nPatchesX, nPoints;   /* patch grid (y slow) and patch side                        */
    const int32_t* region;      /* [nPatchesY * nPatchesX][4] 1-based first x, last x, first y, last y */
    const int32_t* nim;         /* [nCh] images per channel                                          */
    const int32_t* dx;          /* [patches][nCh] channel offsets (ChannelData::channelOffset)       */
    const int32_t* dy;          /* [patches][nCh]                                                    */
    const float* img;           /* [patches][nPoints][nPoints] restored objects                      */
    int npsf; const float* psf; /* [patches][npsf][nPoints][nPoints]                                 */
    int nobj; const float* cobj;/* [patches][nobj][nPoints][nPoints] convolved objects               */
    int nres; const float* res; /* [patches][nres][nPoints][nPoints] residuals                       */
    int nalpha, nm; const float* alpha;  /* [patches][nalpha][nm]                                   */
    int ndiv; const float* div; /* [ndiv][nPH][nPH] diversity phases (the same in every patch)       */
    int nFileNames; const char* const* fileNames;
} mfbd_momfbd;
int mfbd_write_momfbd(const char* path, const mfbd_momfbd* m);
const char* mfbd_file_last_error(void);

/* ---- per-job setup (host; GlobalData/Object::initObject equivalents) ---- */
const char* mfbd_setup_last_error(void);
/* Pupil::calculatePupilSize (pupil.cpp:78-100) */
int mfbd_pupil_size(double wavelength, int nPixels, double telescopeD, double arcSecsPerPixel,
                    double* freqCutoff, double* radius, int* pupilPixels);
/* makePupil + Pupil::normalize (utils.cpp:68-121, pupil.cpp:280-299) */
int mfbd_make_pupil(int P, double radius, double coRadius, double* pupil, double* area);
/* Pupil::generateSupport (pupil.cpp:245-277): pupil support, its OTF-grid index, OTF support */
int mfbd_pupil_support(int P, const double* pupil, int* nPS, int* ps, int* pio, int* nOS, int* os, double* area);
/* Zernike::getCovariance (zernike.cpp:102-148) */
int mfbd_noll_covariance(int i, int j, double* cov);
/* ModeSet::generate (modes.cpp:348-439): unit-RMS Zernike or KL modes [K][P*P] */
int mfbd_make_modes(int P, double radius, double angle, int nModes, const int* modeIds, int basis, int klMin,
                    int klMax, double klCutoff, double* modes, double* atmRms, int* tiltXY);
/* mfbd_make_modes with the Zernike polynomials, their normalisation and the KL sums on the
 * device (FP64; the host version evaluates the radial polynomials in long double): the
 * same plan of modes (ModeSet / PupilMode, modes.cpp:158-205; karhunenLoeveExpansion,
 * zernike.cpp:367-479), for jobs whose KL_MAX_MODE (2000 by default) makes the host
 * evaluation of every Zernike the startup cost.  Errors: mfbd_basis_last_error(). */
int mfbd_make_modes_device(int P, double radius, double angle, int nModes, const int* modeIds, int basis, int klMin,
                           int klMax, double klCutoff, int device, double* modes, double* atmRms, int* tiltXY);
const char* mfbd_basis_last_error(void);
/* ModeSet::getNorms + normalize(scale) (modes.cpp:443-535) */
int mfbd_normalize_modes(int P, int K, double* modes, const double* pupil, double area, double scale);
/* ModeSet::measureJacobian (modes.cpp:464-512) */
int mfbd_measure_jacobian(int P, int K, const double* modes, const double* pupil, double scale, int* tiltXY,
                          double* J, double* Ji);
/* Constraints::init + nullspace (constraints.cpp:85-198, 429-469, 536-622) */
int mfbd_make_constraints(int nObj, const int* nChPerObj, const int* nFramesPerCh, const int* wavefronts, int K,
                          int* nPar, int* nFree, int* nnz, int* rows, int* cols, double* vals, int capacity);

#ifdef __cplusplus
}
#endif

#endif /* MOMFBD_HIP_H */
