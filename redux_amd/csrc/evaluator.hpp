// Interface between the host-side solver (CG/line search state machines, one per
// patch) and the device pipeline that evaluates the MOMFBD metric, gradients and
// restoration sums for a batch of patches in one set of kernel launches.
#pragma once

#include <memory>
#include <string>
#include <vector>

namespace mfbd {

// Lanes of the device evaluator: up to kMaxEvalLanes asynchronous evaluation lanes (each a
// contiguous share of the batch, with its own scratch buffers and stream) and the
// synchronous lane kSyncLane (set-up, stage-wise calls, restoration sums).
constexpr int kMaxEvalLanes = 4;
constexpr int kSyncLane = kMaxEvalLanes;

struct Geometry {
    int s, P, N, K, nObj, nCh, nImg, dataY, dataX, startY, startX;
};

struct Structure {
    std::vector<int> chObj, imgCh, imgObj;   // per channel / per image
    std::vector<double> weight, wavelength, freqCutoff;
    std::vector<int> nImgObj;                 // images per object
    double graddiffStep = 1e-2;               // GRADDIFF_STEP (times the wavelength)
};

// One evaluation request: metric at alpha for patch `slot`, optionally a gradient.
struct EvalReq {
    int slot = 0;
    bool wantGrad = false;
    int gradMethod = 2;     // 1 finite difference, 2 Vogel, 3 adjoint
    int nEnabled = 0;       // gradient for modes [0, nEnabled)
    const double* alpha = nullptr;   // [nImg][K], tilt offsets already added
    // line-search probes without the mode GEMM (Solver::my_precalc / metricAt,
    // solver.cpp:246-327): line 1 starts a line at this slot (base phases from alpha, the
    // direction's phases B dAlpha) and probes it; line 2 probes the slot's current line.
    // The probe's phases are base + lineS * direction.  Never with wantGrad.
    int line = 0;
    double lineS = 0;
    bool saveBase = false;           // line 0: its phases also become the slot's line base (a
                                     // gradient that starts a conjugate-gradient iteration)
    const double* dAlpha = nullptr;  // line 1: [nImg][K], no tilt offsets
    double f = 0;                    // out: metric
    double* g = nullptr;             // out: gradient [nImg][K] (when wantGrad)
};

// (Re)initialisation of one image: cutout at the current integer shift.
struct ImgInit {
    int slot, img, shiftX, shiftY;
};

struct KernelStat {
    std::string name;
    double ms = 0;
    int launches = 0;
    long long units = 0;   // images (tiles) processed, summed over launches
};

// A device line search (conjugate gradient with line phases): a new line of patch `idx` (its
// index in the lane's list) from metric fa along the direction dAlpha ([nImg][K], no tilt
// offsets) with conjugate_rdx's lambda; its base phases are those of the slot's last
// gradient evaluation with saveBase.
struct LsStart {
    int idx;
    double fa, lambda;
    const double* dAlpha;
};

struct LsResult {
    bool done;
    double step, f;         // the line's end (brent's x and f(x))
    int probes;             // metric evaluations of the line
};

class DeviceEvaluator {
public:
    virtual ~DeviceEvaluator() = default;
    // Device line search of a lane's patches (slots, fixed for the solve): ls_begin sets up
    // the lane's state and, with `graph`, captures one round (advance every line to its next
    // probe, evaluate the probes, their metrics) as a hipGraph; ls_start enqueues new lines;
    // ls_run runs the round `rounds` times (graph replays or direct launches); ls_end waits
    // and reports each patch's line.  Returns false from ls_begin when this form has no
    // device line search.
    virtual bool ls_begin(int lane, const std::vector<int>& slots, bool graph) = 0;
    virtual void ls_start(int lane, const std::vector<LsStart>& starts) = 0;
    virtual void ls_run(int lane, int rounds) = 0;
    virtual void ls_end(int lane, std::vector<LsResult>& out) = 0;
    virtual void ls_clear() = 0;
    virtual void set_basis(const double* pupil, const double* modes) = 0;
    virtual void set_structure(const Structure& st, const std::vector<double>& otfNorm) = 0;
    // raw frames of `n` patches into slots [first, first+n); phiCh [n][nCh][P*P]
    virtual void load(int first, int n, const float* data, const double* phiCh) = 0;
    virtual void set_phi_channel(int slot, const double* phiCh) = 0;
    // windowing, padding and FFT of the listed images, then ftSum of the touched objects;
    // the synchronous lane completes before returning, an evaluation lane only enqueues
    virtual void init_images(const std::vector<ImgInit>& imgs, const std::vector<double>& window, int lane) = 0;
    // windowed-frame standard deviation and noise level (before NOISE_FUDGE) of the
    // listed images at their shifts; patch side <= 128
    virtual void frame_stats(const std::vector<ImgInit>& imgs, const std::vector<double>& window,
                             const std::vector<double>& noiseWindow, double* noise, double* stddev) = 0;
    // regularisation gamma of n slots, gamma[k * nObj + o] for slot slots[k]
    virtual void set_gamma(int n, const int* slots, const double* gamma) = 0;
    // which of the nImg images the patch in `slot` has (ragged frame counts): a missing
    // image gets a zero pupil amplitude, so it adds nothing to P, Q, the metric or gradient
    virtual void set_images_enabled(int slot, const unsigned char* on) = 0;
    virtual void evaluate(std::vector<EvalReq>& reqs) = 0;
    // asynchronous form on an evaluation lane: begin enqueues and returns, end waits and fills
    // f / g; the two lanes let the host work on one half of a batch while the device
    // evaluates the other
    virtual void evaluate_begin(int lane, std::vector<EvalReq>& reqs) = 0;
    virtual void evaluate_end(int lane, std::vector<EvalReq>& reqs) = 0;
    // The restored objects (Object::restorePatch, object.cpp:298-331) of n patches at their
    // alphas (alpha[k] points at [nImg][K] for slots[k]), on the synchronous lane while the
    // evaluation lanes run: begin enqueues the sums, the Scharmer filter (when filter; nf
    // [n][nObj]: the noise power times N^2, c2 [nObj]: the squared cut-off radius), the
    // inverse transform and the crop, and returns; ready says whether they are done; end
    // waits and copies out [n][nObj][s][s] (full: the N x N grid, [n][nObj][N][N]).  One
    // restoration at a time.
    virtual void restore_objects_begin(int n, const int* slots, const double* const* alpha, const double* nf,
                                       const double* c2, bool filter, bool full) = 0;
    virtual bool restore_ready() = 0;
    virtual void restore_objects_end(float* out) = 0;
    // Optional restoration outputs of n slots at their alphas (object.cpp:332-407), for the
    // images imgs[k] of slot k (their current shifts), each cropped to the centred s x s of
    // the N x N grid and written as floats, zero for images a patch lacks: kOutPsf -> psf
    // [n][nImg][s][s], the PSF (normalised inverse transform of the OTF, re-centred), or
    // kOutPsfAvg -> psf [n][nObj][s][s], those PSFs averaged per object; kOutCobj -> cobj
    // [n][nImg][s][s], the object of the image (objFull: [n][nObj][N][N], natural layout)
    // convolved with its OTF; kOutRes -> res, that convolved object minus the windowed,
    // padded frame.  On the synchronous lane.
    static constexpr int kOutPsf = 1, kOutCobj = 2, kOutRes = 4, kOutPsfAvg = 8;
    virtual void restore_extra(int n, const int* slots, const double* const* alpha,
                               const std::vector<std::vector<ImgInit>>& imgs, const std::vector<float>& objFull,
                               int want, float* psf, float* cobj, float* res) = 0;
    virtual int nOS() const = 0;            // OTF support points (full)
    virtual int nOSStored() const = 0;      // points stored on the device (Hermitian half)
    // full support point u -> stored point h (value as stored), or -1-h (its conjugate)
    virtual const std::vector<int>& support_map() const = 0;
    virtual int nPS() const = 0;            // pupil support points
    virtual const std::vector<int>& otf_support() const = 0;
    virtual void debug_get(const std::string& what, int slot, std::vector<double>& out) = 0;
    virtual std::vector<KernelStat> kernel_stats() = 0;
    virtual void reset_stats() = 0;
    virtual int capacity() const = 0;
    virtual void bind_thread() = 0;          // make the context's device current on this thread
    // pinned [slots][nImg][K] buffer: a request whose alpha points at its slot's row is
    // uploaded from here without a host-side copy
    virtual double* alpha_staging() = 0;
};

std::unique_ptr<DeviceEvaluator> make_device_evaluator(const Geometry& g, int precision, int maxSlots,
                                                       int maxBatch, int device, std::string& err);

}  // namespace mfbd
