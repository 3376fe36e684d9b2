// Host side of the MI355X MOMFBD patch solver: the C-ABI context, the per-patch
// Solver::run state machine and the batch driver.
//
// Each patch runs the reference's control flow (src/lib/momfbd/solver.cpp:330-580:
// mode stepping, conjugate_rdx iterations with bracket + Brent line search,
// convergence tests, shift re-initialisation between stages) as a C++20 coroutine.
// Every metric or gradient evaluation is a co_await: the driver resumes all
// patches until each waits for an evaluation, then evaluates all pending requests
// in one batched device pass (many patches x all images per launch) and resumes
// them.  With maxBatch = 1 this is the reference's one-patch-at-a-time drop-in.
#include "../../include/momfbd_hip.h"
#include "evaluator.hpp"
#include "hostfft.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <coroutine>
#include <cstring>
#include <exception>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

using namespace mfbd;

namespace {

constexpr double GOLDEN = 1.6180339887498949025;   // constants.hpp:26
constexpr double TINY = 1.0E-32;                   // helpers.cpp:56
constexpr int ST_SUCCESS = 0, ST_FAILURE = -1, ST_ENOPROG = 27;

// ------------------------------------------------------------ coroutines ----
template <typename T>
struct Task {
    struct promise_type {
        T value{};
        std::coroutine_handle<> cont;
        std::exception_ptr exc;
        Task get_return_object() { return Task{std::coroutine_handle<promise_type>::from_promise(*this)}; }
        std::suspend_always initial_suspend() noexcept { return {}; }
        struct Final {
            bool await_ready() noexcept { return false; }
            std::coroutine_handle<> await_suspend(std::coroutine_handle<promise_type> h) noexcept {
                auto c = h.promise().cont;
                return c ? c : std::noop_coroutine();
            }
            void await_resume() noexcept {}
        };
        Final final_suspend() noexcept { return {}; }
        void return_value(T v) { value = v; }
        void unhandled_exception() { exc = std::current_exception(); }
    };
    std::coroutine_handle<promise_type> h;
    explicit Task(std::coroutine_handle<promise_type> hh) : h(hh) {}
    Task(Task&& o) noexcept : h(o.h) { o.h = {}; }
    Task(const Task&) = delete;
    ~Task() { if (h) h.destroy(); }
    bool await_ready() { return false; }
    std::coroutine_handle<> await_suspend(std::coroutine_handle<> c) {
        h.promise().cont = c;
        return h;
    }
    T await_resume() {
        if (h.promise().exc) std::rethrow_exception(h.promise().exc);
        return h.promise().value;
    }
};

// ------------------------------------------------------------ parallel for ----
void parallel_for(int n, const std::function<void(int)>& f, int maxThreads = 16) {
    int nt = (int)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), (unsigned)std::max(1, maxThreads));
    nt = std::min(nt, n);
    if (nt <= 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    std::exception_ptr exc;
    std::atomic<bool> bad{false};
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (int i; (i = next.fetch_add(1)) < n;) {
                try { f(i); } catch (...) { if (!bad.exchange(true)) exc = std::current_exception(); }
            }
        });
    for (auto& t : th) t.join();
    if (exc) std::rethrow_exception(exc);
}

// Persistent worker pool: run(n, f) calls f(0..n-1) on the workers and the caller.
class Pool {
public:
    explicit Pool(int nt) {
        for (int t = 0; t < nt; ++t) th_.emplace_back([this] { loop(); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> l(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    void run(int count, std::function<void(int)> fn) {
        if (count <= 0) return;
        if (th_.empty() || count == 1) {
            for (int i = 0; i < count; ++i) fn(i);
            return;
        }
        {
            std::lock_guard<std::mutex> l(m_);
            f_ = std::move(fn);
            n_ = count;
            next_ = 0;
            pending_ = (int)th_.size();
            exc_ = nullptr;
            ++gen_;
        }
        cv_.notify_all();
        work();
        {
            std::unique_lock<std::mutex> l(m_);
            done_.wait(l, [&] { return pending_ == 0; });
        }
        if (exc_) std::rethrow_exception(exc_);
    }

private:
    void loop() {
        long seen = 0;
        while (true) {
            {
                std::unique_lock<std::mutex> l(m_);
                cv_.wait(l, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            work();
            {
                std::lock_guard<std::mutex> l(m_);
                if (--pending_ == 0) done_.notify_one();
            }
        }
    }
    void work() {
        for (int i; (i = next_.fetch_add(1)) < n_;) {
            try {
                f_(i);
            } catch (...) {
                std::lock_guard<std::mutex> l(m_);
                if (!exc_) exc_ = std::current_exception();
            }
        }
    }
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    std::function<void(int)> f_;
    int n_ = 0, pending_ = 0;
    std::atomic<int> next_{0};
    long gen_ = 0;
    bool stop_ = false;
    std::exception_ptr exc_;
};

// ----------------------------------------------- host statistics & filters ----
double mean_rev(const double* d, long n) {       // ArrayStats::getMinMaxMean (arraystats.cpp:17-45)
    double sum = 0;
    long cnt = 0;
    for (const double* p = d + n; p-- > d;)
        if (std::isfinite(*p)) { sum += *p; cnt++; }
    return cnt ? sum / cnt : sum;
}

double stddev_rev(const double* d, long n, double mean) {   // getRmsStddev (arraystats.cpp:59-81)
    double rms = 0;
    long cnt = 0;
    for (const double* p = d + n; p-- > d;)
        if (*p == *p) { rms += *p * *p; cnt++; }
    if (!cnt) return 0;
    rms /= cnt;
    return std::sqrt(rms - mean * mean);
}

void apodize1(double* data, int n, double target) {         // math::apodize (functions.cpp:39-50)
    double mean = 0.5 * (*data + target), amp = 0.5 * (*data - target), step = M_PI / (n - 1);
    for (int i = 1; i < n; ++i) data[i] = mean + amp * std::cos(i * step);
}

void apodize_in_place(double* d, int nR, int nC, int blend) {   // utils.cpp:583-613
    int rb = std::min(blend, nC / 2), cb = std::min(blend, nR / 2);
    std::vector<double> t(std::max(rb, cb) + 2, 0.0);
    apodize1(t.data(), rb + 2, 1.0);
    for (int c = 0; c < nC; ++c)
        for (int r = 0; r < rb; ++r) { d[r * nC + c] *= t[r + 1]; d[(nR - r - 1) * nC + c] *= t[r + 1]; }
    std::fill(t.begin(), t.end(), 0.0);
    apodize1(t.data(), cb + 2, 1.0);
    for (int r = 0; r < nR; ++r)
        for (int c = 0; c < cb; ++c) { d[r * nC + c] *= t[c + 1]; d[r * nC + nC - c - 1] *= t[c + 1]; }
}

// FourierTransform::noise (fouriertransform.cpp:448-495), FULLCOMPLEX, mask = -1, limit = -1
double noise_level(const cd* ft, int ny, int nx) {
    int mask = std::max(ny, nx) / 6 + 1;
    double limits = (double)ny * nx / 4.0;
    double np = 0;
    long cnt = 0;
    for (int x = mask; x < nx - mask + 1; ++x) {
        double xx = std::min(x, nx - x);
        xx *= xx;
        for (int y = mask; y < ny - mask + 1; ++y) {
            double yy = std::min(y, ny - y);
            if (xx + yy * yy < limits) continue;
            np += std::norm(ft[(long)y * nx + x]);
            cnt++;
        }
    }
    return cnt ? std::sqrt(np / ((double)cnt * ny * nx)) : np;
}

}  // namespace

// ------------------------------------------------------------------ context ----


// alpha = alpha_offset + N beta of one patch (Constraints::reverseAndAdd / apply,
// constraints.hpp:128-159): N as (row, col, value) triplets
struct NullSpace {
    int nFree = 0;
    std::vector<int> row, col;
    std::vector<double> val;
};

struct mfbd_ctx {
    Geometry g{};
    mfbd_params prm{};
    int device = 0;
    std::string err;
    std::unique_ptr<DeviceEvaluator> dev;
    Structure st;
    bool haveStructure = false, haveBasis = false, haveCons = false;
    // basis
    std::vector<double> pupil, modes, jac, otfNorm;
    std::vector<int> tilt;
    // constraints
    int nPar = 0, nFree = 0;
    NullSpace ns;                                // the job's
    // ragged batches: per patch, the modes it solves for, the images it has and its own
    // null space (nFree 0: the job's), all within the context's K and nImg
    std::vector<int> nModesP;
    std::vector<std::vector<unsigned char>> imgOn;
    std::vector<NullSpace> nsP;
    // patches
    int nPatch = 0;
    std::vector<float> data;          // host copy of raw frames [n][nImg][dataY][dataX]
    std::vector<double> phiFixed, resOffset, noiseFudge;
    std::vector<double> window, noiseWindow;
    // per-patch host state (stage-wise API)
    std::vector<std::vector<int>> shifts;       // [patch][nImg*2]
    std::vector<std::vector<double>> tiltOff;   // [patch][nImg*2]
    std::vector<std::vector<double>> noise;     // [patch][nImg]
    std::vector<std::vector<double>> gamma;     // [patch][nObj]
    bool initDone = false;
    // the patches hold the state the last mfbd_solve left (each image's final shift and tilt
    // offset); stage-wise calls evaluate on it until mfbd_init_patches or a new load
    bool solved = false;
    std::unique_ptr<Pool> pool;
    bool deviceLS = false;            // this solve's line searches run on the device (lineRounds)
    // host-side timing of the batch driver: evaluate / resume / re-init phases
    std::vector<KernelStat> hostStats = std::vector<KernelStat>(7);
};

namespace {

int set_err(mfbd_ctx* c, int code, const std::string& m) {
    if (c) c->err = m;
    return code;
}

const NullSpace& ns_of(const mfbd_ctx* c, int p) {
    return c->nsP.empty() || c->nsP[p].nFree == 0 ? c->ns : c->nsP[p];
}

bool img_on(const mfbd_ctx* c, int p, int j) { return c->imgOn.empty() || c->imgOn[p][j]; }

// phi_channel per patch: phi_fixed - residual-tilt (Channel::initPatch, channel.cpp:908-939)
void phi_channels(const mfbd_ctx* c, int p, std::vector<double>& out) {
    const int P2 = c->g.P * c->g.P, K = c->g.K;
    out.assign((size_t)c->g.nCh * P2, 0.0);
    for (int ch = 0; ch < c->g.nCh; ++ch) {
        const int o = c->st.chObj[ch];
        double* phi = &out[(size_t)ch * P2];
        std::memcpy(phi, &c->phiFixed[(size_t)ch * P2], sizeof(double) * P2);
        const double* J = &c->jac[8 * o];
        const double rx = c->resOffset[((size_t)p * c->g.nCh + ch) * 2], ry = c->resOffset[((size_t)p * c->g.nCh + ch) * 2 + 1];
        const double ax = rx * J[4] + ry * J[5], ay = rx * J[6] + ry * J[7];
        const int tx = c->tilt[2 * o], ty = c->tilt[2 * o + 1];
        if (tx >= 0 && std::fabs(rx) > 0) {
            const double* m = &c->modes[((size_t)o * K + tx) * P2];
            for (int k = 0; k < P2; ++k) phi[k] = phi[k] + (-ax) * m[k];
        }
        if (ty >= 0 && std::fabs(ry) > 0) {
            const double* m = &c->modes[((size_t)o * K + ty) * P2];
            for (int k = 0; k < P2; ++k) phi[k] = phi[k] + (-ay) * m[k];
        }
    }
}

// otfNormalization = 1/sqrt(max |S|) of the aberration-free OTF (channel.cpp:941-952);
// max |S| = S(0) = sum pupil^2 for the autocorrelation of the pupil field
double otf_normalization(const mfbd_ctx* c) {
    const int P = c->g.P, N = 2 * P;
    std::vector<cd> a((size_t)N * N, 0.0);
    for (int i = 0; i < P * P; ++i)
        if (c->pupil[i] > 1e-9) a[(size_t)(i / P + P / 2) * N + i % P + P / 2] = c->pupil[i];
    fft2d(a.data(), N, N, -1);
    for (auto& v : a) v = std::norm(v);
    fft2d(a.data(), N, N, +1);
    double mx = 0;
    for (auto& v : a) mx = std::max(mx, std::abs(v) / ((double)N * N));
    return 1.0 / std::sqrt(mx);
}

// windowed image of one frame at a shift (SubImage::getWindowedImg), its stats and
// noise level (SubImage::initialize with doReset, subimage.cpp:155-187)
void frame_noise(const mfbd_ctx* c, int p, int j, int sx, int sy, double& noise, double& sd) {
    const int s = c->g.s;
    std::vector<double> img((size_t)s * s);
    const float* blk = &c->data[((size_t)p * c->g.nImg + j) * c->g.dataY * c->g.dataX];
    const int y0 = c->g.startY + sy, x0 = c->g.startX + sx;
    for (int y = 0; y < s; ++y)
        for (int x = 0; x < s; ++x) img[(size_t)y * s + x] = blk[(size_t)(y0 + y) * c->g.dataX + x0 + x];
    const double avg = mean_rev(img.data(), (long)s * s);
    for (size_t k = 0; k < img.size(); ++k) img[k] = (img[k] - avg) * c->window[k] + avg;
    const double mean = mean_rev(img.data(), (long)s * s);
    sd = stddev_rev(img.data(), (long)s * s, mean);
    std::vector<cd> d2((size_t)s * s);
    for (size_t k = 0; k < img.size(); ++k) d2[k] = (img[k] - mean) * c->noiseWindow[k];
    fft2d(d2.data(), s, s, -1);
    noise = noise_level(d2.data(), s, s) * c->noiseFudge[c->st.imgCh[j]];
}

// reset initialisation of patch p at the given shifts: noise, gamma, device D/ftSum
void reset_init(mfbd_ctx* c, const std::vector<int>& patches) {
    const int nImg = c->g.nImg, nObj = c->g.nObj;
    std::vector<double> noiseAll((size_t)patches.size() * nImg), sdAll((size_t)patches.size() * nImg);
    std::vector<ImgInit> imgs;
    for (int p : patches)
        for (int j = 0; j < nImg; ++j) imgs.push_back({p, j, c->shifts[p][2 * j], c->shifts[p][2 * j + 1]});
    if (c->g.s <= 128) {
        c->dev->frame_stats(imgs, c->window, c->noiseWindow, noiseAll.data(), sdAll.data());
        for (size_t k = 0; k < noiseAll.size(); ++k) noiseAll[k] *= c->noiseFudge[c->st.imgCh[k % nImg]];
    } else {
        parallel_for((int)(patches.size() * nImg), [&](int k) {
            const int p = patches[k / nImg], j = k % nImg;
            frame_noise(c, p, j, c->shifts[p][2 * j], c->shifts[p][2 * j + 1], noiseAll[k], sdAll[k]);
        });
    }
    std::vector<double> gamAll(patches.size() * nObj, 0.0);
    std::vector<int> nIO(nObj);
    for (size_t a = 0; a < patches.size(); ++a) {
        const int p = patches[a];
        c->noise[p].assign(noiseAll.begin() + a * nImg, noiseAll.begin() + (a + 1) * nImg);
        double* gam = &gamAll[a * nObj];
        std::fill(nIO.begin(), nIO.end(), 0);
        for (int j = 0; j < nImg; ++j)
            if (img_on(c, p, j)) nIO[c->st.imgObj[j]]++;
        for (int j = 0; j < nImg; ++j) {                      // Object::addRegGamma
            if (!img_on(c, p, j)) continue;
            const int o = c->st.imgObj[j];
            gam[o] += 0.10 * (noiseAll[a * nImg + j] / sdAll[a * nImg + j]) / nIO[o];
        }
        c->gamma[p].assign(gam, gam + nObj);
    }
    c->dev->set_gamma((int)patches.size(), patches.data(), gamAll.data());
    c->dev->init_images(imgs, c->window, kSyncLane);
}

// SubImage::adjustShifts (subimage.cpp:444-480) for one image; returns true if shifted
bool adjust_shift(mfbd_ctx* c, int p, int j, const double* a) {
    if (!img_on(c, p, j)) return false;
    const int o = c->st.imgObj[j];
    const double* J = &c->jac[8 * o];
    const int tx = c->tilt[2 * o], ty = c->tilt[2 * o + 1];
    if (tx < 0 || ty < 0) return false;
    double* off = &c->tiltOff[p][2 * j];
    int* sh = &c->shifts[p][2 * j];
    const double ax = a[tx] + off[0], ay = a[ty] + off[1];
    const double sx = ax * J[0] + ay * J[1], sy = ax * J[2] + ay * J[3];
    if (!std::isfinite(sx) || !std::isfinite(sy)) return false;     // a failed stage's alphas: no re-cut
    const int thx = -(int)std::lround(sx), thy = -(int)std::lround(sy);
    auto shift1 = [&](int first, int size, int n) {           // Array::shift (array.hpp:1119-1140)
        if (n == 0) return 0;
        const int last = first + c->g.s - 1;
        if (n + first < 0) n = -first;
        if (std::abs(n + last) >= size) n = size - last - 1;
        return n;
    };
    const int actX = thx ? shift1(c->g.startX + sh[0], c->g.dataX, thx) : 0;
    const int actY = thy ? shift1(c->g.startY + sh[1], c->g.dataY, thy) : 0;
    if (actX || actY) {
        sh[0] += actX;
        sh[1] += actY;
        off[0] += actX * J[4] + actY * J[5];
        off[1] += actX * J[6] + actY * J[7];
        return true;
    }
    return false;
}

// alpha with the tilt offsets folded in (SubImage::addToPhi, subimage.cpp:509-532)
void effective_alpha(const mfbd_ctx* c, int p, const double* alpha, double* out) {
    const int K = c->g.K;
    std::memcpy(out, alpha, sizeof(double) * c->g.nImg * K);
    for (int j = 0; j < c->g.nImg; ++j) {
        const int o = c->st.imgObj[j];
        const int tx = c->tilt[2 * o], ty = c->tilt[2 * o + 1];
        if (tx >= 0) out[(size_t)j * K + tx] += c->tiltOff[p][2 * j];
        if (ty >= 0 && ty != tx) out[(size_t)j * K + ty] += c->tiltOff[p][2 * j + 1];
    }
}

// The restoration's per-(patch, object) noise term (Object::restorePatch, object.cpp:306-312):
// the mean of noise^2 over the object's images the patch has, times N^2 as ScharmerFilter
// scales it (utils.cpp:336); [n][nObj]
std::vector<double> restore_noise(const mfbd_ctx* c, const int* patches, int n) {
    const int N = 2 * c->g.P, nO = c->g.nObj;
    std::vector<double> nf((size_t)n * nO);
    for (int k = 0; k < n; ++k) {
        const int p = patches[k];
        for (int o = 0; o < nO; ++o) {
            double nv = 0;
            int cnt = 0;
            for (int j = 0; j < c->g.nImg; ++j)
                if (c->st.imgObj[j] == o && img_on(c, p, j)) { nv += c->noise[p][j] * c->noise[p][j]; cnt++; }
            nv /= cnt;
            nf[(size_t)k * nO + o] = nv * N * N;
        }
    }
    return nf;
}

// the filter's squared cut-off radius per object, 0.9 of the frequency cut-off
// (object.cpp:313-315)
std::vector<double> restore_cutoffs(const mfbd_ctx* c) {
    std::vector<double> c2(c->g.nObj);
    for (int o = 0; o < c->g.nObj; ++o) {
        const double cutoff = 0.90 * c->st.freqCutoff[o];
        c2[o] = cutoff * cutoff;
    }
    return c2;
}

// --------------------------------------------------------- patch solver ----
struct EvalSlot {
    bool pending = false;
    bool wantGrad = false;
    int method = 2, nEnabled = 0;
    double* alphaEff = nullptr;        // this patch's row of the device's pinned alpha staging
    int line = 0;                      // EvalReq::line: 1 start a line, 2 probe it
    double lineS = 0;
    bool saveBase = false;             // EvalReq::saveBase
    std::vector<double> dAlpha;        // the line's direction in alpha ([nImg][K])
    double f = 0;
    std::vector<double> gAlpha;
    std::coroutine_handle<> h;
};

struct PatchRun {
    mfbd_ctx* c = nullptr;
    int p = 0;
    const NullSpace* ns = nullptr;  // this patch's null space
    int K = 0;                      // modes this patch solves for
    EvalSlot ev;
    bool wantShiftInit = false;     // pending re-initialisation after a stage
    std::vector<int> shiftedObjs;
    std::coroutine_handle<> initH;
    // solver state (solver.cpp:134-144, conjugate_gradient.cpp:29-38)
    std::vector<double> alpha, alphaOffset, x, g, dx, x1, p_, prevG, beta0, ga;
    bool lineStart = false;         // the next line-search probe starts a line (linePhases)
    struct DeviceLine {             // a line search handed to the device (lineRounds)
        bool pending = false, active = false;
        double fa = 0, lambda = 0;  // the line's start: metric at x and 1 / |p| signed downhill
        double step = 0, f = 0;     // its end: brent's x and f(x)
        int probes = 0;
        std::coroutine_handle<> h;
    } dl;
    int iterCG = 0;
    double pnorm = 0, g0norm = 0, step = 0;
    int nEnabled = 0, gradMethod = 1;
    int nF = 0, nDF = 0;
    mfbd_result res{};
    std::vector<double> metrics;
    // vector_bfgs2 (GETSTEP = BFGS_inv): the line wrapper over x0 + alpha p caches f, the
    // gradient and the slope at its last trial point (keys: the alpha they belong to)
    struct Bfgs2 {
        std::vector<double> x0, g0, p, xa, ga, dx0, dg0;
        double fa = 0, dfa = 0, xKey = 0, fKey = 0, gKey = 0, dfKey = 0;
        double step = 0, g0norm = 0, pnorm = 0, deltaF = 0, fp0 = 0, sigma = 0;
    } bf;
    // steepest_descent / vector_bfgs (GETSTEP = steepest_descent / BFGS)
    struct Step1 {
        int iter = 0;
        double step = 0, tol = 0, pnorm = 0, g0norm = 0;
        std::vector<double> x1, g1, dx1, x2, p, x0, g0, dx0, dg0;
    } sd;
};

struct EvalAwait {
    PatchRun* pr;
    bool await_ready() { return false; }
    void await_suspend(std::coroutine_handle<> h) { pr->ev.h = h; pr->ev.pending = true; }
    void await_resume() {}
};

struct LineAwait {
    PatchRun* pr;
    bool await_ready() { return false; }
    void await_suspend(std::coroutine_handle<> h) { pr->dl.h = h; pr->dl.pending = true; }
    void await_resume() {}
};

struct InitAwait {
    PatchRun* pr;
    bool await_ready() { return false; }
    void await_suspend(std::coroutine_handle<> h) { pr->initH = h; pr->wantShiftInit = true; }
    void await_resume() {}
};

void ns_reverse_add(const mfbd_ctx* c, const NullSpace& ns, const double* in, const double* offset, double* out) {
    std::memcpy(out, offset, sizeof(double) * c->nPar);
    for (size_t e = 0; e < ns.row.size(); ++e) out[ns.row[e]] += ns.val[e] * in[ns.col[e]];
}

void ns_apply(const NullSpace& ns, const double* in, double* out) {
    std::fill(out, out + ns.nFree, 0.0);
    for (size_t e = 0; e < ns.row.size(); ++e) out[ns.col[e]] += ns.val[e] * in[ns.row[e]];
}

Task<double> my_f(PatchRun& r, const double* beta) {   // Solver::my_f (solver.cpp:219-225)
    ns_reverse_add(r.c, *r.ns, beta, r.alphaOffset.data(), r.alpha.data());
    r.ev.wantGrad = false;
    r.ev.line = 0;
    r.ev.saveBase = false;
    effective_alpha(r.c, r.p, r.alpha.data(), r.ev.alphaEff);
    co_await EvalAwait{&r};
    r.nF++;
    co_return r.ev.f;
}

Task<int> my_df(PatchRun& r, const double* beta, double* gb, bool saveBase = false) {   // Solver::my_df (solver.cpp:228-234)
    ns_reverse_add(r.c, *r.ns, beta, r.alphaOffset.data(), r.alpha.data());
    r.ev.wantGrad = true;
    r.ev.line = 0;
    r.ev.saveBase = saveBase;
    r.ev.method = r.gradMethod;
    r.ev.nEnabled = r.nEnabled;
    effective_alpha(r.c, r.p, r.alpha.data(), r.ev.alphaEff);
    co_await EvalAwait{&r};
    r.nDF++;
    ns_apply(*r.ns, r.ev.gAlpha.data(), gb);
    co_return 0;
}

// value_at / take_step (conjugate_gradient.cpp:110-125).  With linePhases the probe at
// x - stp lambda p is evaluated as the line's phases base + s direction, s = -stp lambda
// (the reference's metricAt, solver.cpp:294-327): the first probe of a line uploads the base
// alpha (the gradient's point x) and the direction in alpha, N p; later probes send s only.
Task<double> value_at(PatchRun& r, double stp, double lambda) {
    const int n = r.ns->nFree;
    if (r.c->prm.linePhases) {
        r.ev.wantGrad = false;
        r.ev.saveBase = false;
        r.ev.line = r.lineStart ? 1 : 2;
        r.ev.lineS = -stp * lambda;
        r.lineStart = false;
        co_await EvalAwait{&r};
        r.nF++;
        co_return r.ev.f;
    }
    for (int k = 0; k < n; ++k) r.dx[k] = 0.0 + (-stp * lambda) * r.p_[k];
    for (int k = 0; k < n; ++k) r.x1[k] = r.x[k] + 1.0 * r.dx[k];
    double f = co_await my_f(r, r.x1.data());
    co_return f;
}

#define SHFT(a, b, c, d) (a) = (b); (b) = (c); (c) = (d);
#define SIGN(a, b) ((b) >= 0.0 ? std::fabs(a) : -std::fabs(a))

struct Bracket { double a, b, mid, fa, fb, fmid; };

// math::bracket (helpers.cpp:58-129)
Task<Bracket> bracket(PatchRun& r, double lambda, Bracket B, int limit) {
    double u, fu, ulim;
    if (B.fb > B.fa) { std::swap(B.a, B.b); std::swap(B.fa, B.fb); }
    B.mid = B.b + GOLDEN * (B.b - B.a);
    B.fmid = co_await value_at(r, B.mid, lambda);
    while (B.fb > B.fmid) {
        double rr = (B.b - B.a) * (B.fb - B.fmid);
        double q = (B.b - B.mid) * (B.fb - B.fa);
        u = B.b - ((B.b - B.mid) * q - (B.b - B.a) * rr) / (2.0 * SIGN(std::max(std::fabs(q - rr), TINY), q - rr));
        ulim = B.b + limit * (B.mid - B.b);
        if ((B.b - u) * (u - B.mid) > 0.0) {
            fu = co_await value_at(r, u, lambda);
            if (fu < B.fmid) { B.a = B.b; B.b = u; B.fa = B.fb; B.fb = fu; break; }
            else if (fu > B.fb) { B.mid = u; B.fmid = fu; break; }
            u = B.mid + GOLDEN * (B.mid - B.b);
            fu = co_await value_at(r, u, lambda);
        } else if ((B.mid - u) * (u - ulim) > 0.0) {
            fu = co_await value_at(r, u, lambda);
            if (fu < B.fmid) {
                SHFT(B.b, B.mid, u, B.mid + GOLDEN * (B.mid - B.b))
                double fnew = co_await value_at(r, u, lambda);
                SHFT(B.fb, B.fmid, fu, fnew)
            }
        } else if ((B.mid - ulim) * (ulim - u) >= 0.0) {
            u = ulim;
            fu = co_await value_at(r, u, lambda);
        } else {
            u = B.mid + GOLDEN * (B.mid - B.b);
            fu = co_await value_at(r, u, lambda);
        }
        SHFT(B.a, B.b, B.mid, u)
        SHFT(B.fa, B.fb, B.fmid, fu)
    }
    if (B.a > B.b) { std::swap(B.a, B.b); std::swap(B.fa, B.fb); }
    if (B.mid > B.b) { std::swap(B.mid, B.b); std::swap(B.fmid, B.fb); }
    if (B.mid < B.a) { std::swap(B.mid, B.a); std::swap(B.fmid, B.fa); }
    co_return B;
}

struct BrentOut { double x, fx; };

// math::brent (helpers.cpp:133-216)
Task<BrentOut> brent(PatchRun& r, double lambda, double ax, double bx, double x, double tol, int limit) {
    const double CGOLD = 2.0 - GOLDEN;
    double a = std::min(ax, bx), b = std::max(ax, bx);
    double u, v, w, fu, fv, fw, fx;
    u = v = w = x;
    fx = co_await value_at(r, x, lambda);
    fu = fv = fw = fx;
    double d = 0.0, e = 0.0;
    for (int iter = 1; iter <= limit; ++iter) {
        double xm = 0.5 * (a + b);
        double tol1 = tol * std::fabs(x) + TINY, tol2 = 2.0 * tol1;
        if (std::fabs(x - xm) <= (tol2 - 0.5 * (b - a))) co_return BrentOut{x, fx};
        if (std::fabs(e) > tol1) {
            double rr = (x - w) * (fx - fv);
            double q = (x - v) * (fx - fw);
            double pp = (x - v) * q - (x - w) * rr;
            q = 2.0 * (q - rr);
            if (q > 0.0) pp = -pp;
            q = std::fabs(q);
            double etemp = e;
            e = d;
            if (std::fabs(pp) >= std::fabs(0.5 * q * etemp) || pp <= q * (a - x) || pp >= q * (b - x)) {
                e = (x >= xm ? a - x : b - x);
                d = CGOLD * e;
            } else {
                d = pp / q;
                u = x + d;
                if (u - a < tol2 || b - u < tol2) d = SIGN(tol1, xm - x);
            }
        } else {
            e = (x >= xm ? a - x : b - x);
            d = CGOLD * e;
        }
        if (std::fabs(d) < tol1) d *= (tol1 / std::fabs(d));
        u = ((std::fabs(d) >= tol1) ? x + d : x + SIGN(tol1, d));
        fu = co_await value_at(r, u, lambda);
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
    co_return BrentOut{x, fx};
}

double dnrm2(const std::vector<double>& v) {
    double s = 0;
    for (double a : v) s += a * a;
    return std::sqrt(s);
}

// conjugate_rdx_iterate (conjugate_gradient.cpp:128-256)
Task<int> cg_iterate(PatchRun& r, double& f) {
    const int n = r.ns->nFree;
    co_await my_df(r, r.x.data(), r.g.data(), r.c->prm.linePhases != 0);   // its phases: the next line's base
    double pnorm = r.pnorm, g0norm = r.g0norm;
    const double g1norm = dnrm2(r.g);
    double pg;
    if (r.iterCG++ == 0) {
        r.p_ = r.g;
        pnorm = r.pnorm = g1norm;
        pg = g1norm * g1norm;
    } else {
        if (g0norm == 0.0) { std::fill(r.dx.begin(), r.dx.end(), 0.0); co_return ST_ENOPROG; }
        double g0g1 = 0;
        for (int k = 0; k < n; ++k) {
            double t = r.prevG[k] - r.g[k];
            g0g1 += t * r.g[k];
            r.prevG[k] = t;
        }
        const double beta = g0g1 / (g0norm * g0norm);
        for (int k = 0; k < n; ++k) r.p_[k] = r.g[k] - beta * r.p_[k];
        pnorm = r.pnorm = dnrm2(r.p_);
        pg = 0;
        for (int k = 0; k < n; ++k) pg += r.p_[k] * r.g[k];
    }
    const double lambda = (pg >= 0.0) ? +1.0 / pnorm : -1.0 / pnorm;
    if (pnorm == 0.0) { std::fill(r.dx.begin(), r.dx.end(), 0.0); co_return ST_ENOPROG; }
    if (r.c->prm.linePhases) {
        // the line through x along p: alpha(x - stp lambda p) = alpha(x) + s N p; r.alpha
        // holds alpha(x) from the gradient evaluation above
        std::vector<double>& da = r.ev.dAlpha;
        da.assign(r.alpha.size(), 0.0);
        for (size_t e = 0; e < r.ns->row.size(); ++e) da[r.ns->row[e]] += r.ns->val[e] * r.p_[r.ns->col[e]];
        r.lineStart = true;
    }
    const double fa = f;
    BrentOut bo;
    if (r.c->deviceLS) {
        // the same bracket and brent, stepped on the device from the line's phases
        r.dl.fa = fa;
        r.dl.lambda = lambda;
        co_await LineAwait{&r};
        r.nF += r.dl.probes;
        bo = {r.dl.step, r.dl.f};
    } else {
        const double stepc = 1E-21 / lambda;
        const double fc = co_await value_at(r, stepc, lambda);
        Bracket B = co_await bracket(r, lambda, Bracket{0.0, stepc, 0.0, fa, fc, 0.0}, 100);
        bo = co_await brent(r, lambda, B.a, B.b, B.mid, 0.1, 100);
    }
    for (int k = 0; k < n; ++k) r.dx[k] = 0.0 + (-bo.x * lambda) * r.p_[k];
    for (int k = 0; k < n; ++k) r.x1[k] = r.x[k] + 1.0 * r.dx[k];
    r.x = r.x1;
    f = bo.fx;
    r.step = 0.5 * std::max(std::fabs(bo.x), 2.0E-9);
    r.g0norm = g1norm;
    r.prevG = r.g;
    r.iterCG = r.iterCG % n;
    co_return ST_SUCCESS;
}

// Solver::my_fdf (solver.cpp:237-243): the metric and the gradient of one point, one request
Task<double> my_fdf(PatchRun& r, const double* beta, double* gb) {
    ns_reverse_add(r.c, *r.ns, beta, r.alphaOffset.data(), r.alpha.data());
    r.ev.wantGrad = true;
    r.ev.line = 0;
    r.ev.saveBase = false;
    r.ev.method = r.gradMethod;
    r.ev.nEnabled = r.nEnabled;
    effective_alpha(r.c, r.p, r.alpha.data(), r.ev.alphaEff);
    co_await EvalAwait{&r};
    r.nF++;
    r.nDF++;
    ns_apply(*r.ns, r.ev.gAlpha.data(), gb);
    co_return r.ev.f;
}

// ---- GETSTEP = BFGS_inv: GSL 1.16 vector_bfgs2, the reference's default stepper
// (config.cpp:632, solver.cpp:362-368).  GSL is not part of the reference tree; its
// published algorithm is restated (multimin/vector_bfgs2.c, linear_minimize.c,
// linear_wrapper.c): BFGS direction p' = g1 - A dx - B dg (B = dx.g1/dx.dg,
// A = -(1 + |dg|^2/dx.dg) B + dg.g1/dx.dg), normalised downhill; Fletcher's line search
// (bracketing, then sectioning; rho 0.01, sigma = tol, tau 9 / 0.05 / 0.5, cubic
// interpolation when both ends carry a slope, 100 trial points); a line wrapper caching
// f, g and the slope at the last trial point.
constexpr double DBL_EPS = 2.2204460492503131e-16;

double dotv(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0;
    for (size_t k = 0; k < a.size(); ++k) s += a[k] * b[k];
    return s;
}

void lw_moveto(PatchRun& r, double alpha) {
    auto& w = r.bf;
    if (alpha == w.xKey) return;
    for (size_t k = 0; k < w.xa.size(); ++k) w.xa[k] = w.x0[k] + alpha * w.p[k];
    w.xKey = alpha;
}

Task<double> lw_f(PatchRun& r, double alpha) {
    auto& w = r.bf;
    if (alpha == w.fKey) co_return w.fa;
    lw_moveto(r, alpha);
    w.fa = co_await my_f(r, w.xa.data());
    w.fKey = alpha;
    co_return w.fa;
}

Task<double> lw_df(PatchRun& r, double alpha) {
    auto& w = r.bf;
    if (alpha == w.dfKey) co_return w.dfa;
    lw_moveto(r, alpha);
    if (alpha != w.gKey) {
        co_await my_df(r, w.xa.data(), w.ga.data());
        w.gKey = alpha;
    }
    w.dfa = dotv(w.ga, w.p);
    w.dfKey = alpha;
    co_return w.dfa;
}

struct FDf { double f, df; };

Task<FDf> lw_fdf(PatchRun& r, double alpha) {
    auto& w = r.bf;
    if (alpha == w.fKey && alpha == w.dfKey) co_return FDf{w.fa, w.dfa};
    if (alpha == w.fKey || alpha == w.dfKey) {
        const double f = co_await lw_f(r, alpha);
        const double df = co_await lw_df(r, alpha);
        co_return FDf{f, df};
    }
    lw_moveto(r, alpha);
    w.fa = co_await my_fdf(r, w.xa.data(), w.ga.data());
    w.fKey = alpha;
    w.gKey = alpha;
    w.dfa = dotv(w.ga, w.p);
    w.dfKey = alpha;
    co_return FDf{w.fa, w.dfa};
}

// real roots of a z^2 + b z + c, ascending (gsl_poly_solve_quadratic)
int solve_quadratic(double a, double b, double c, double& z0, double& z1) {
    if (a == 0) {
        if (b == 0) return 0;
        z0 = -c / b;
        return 1;
    }
    const double disc = b * b - 4 * a * c;
    if (disc > 0) {
        if (b == 0) {
            const double rr = std::sqrt(-c / a);
            z0 = -rr;
            z1 = rr;
        } else {
            const double sgnb = b > 0 ? 1 : -1, temp = -0.5 * (b + sgnb * std::sqrt(disc));
            const double r1 = temp / a, r2 = c / temp;
            z0 = std::min(r1, r2);
            z1 = std::max(r1, r2);
        }
        return 2;
    }
    if (disc == 0) {
        z0 = z1 = -0.5 * b / a;
        return 2;
    }
    return 0;
}

double interp_quad(double f0, double fp0, double f1, double zl, double zh) {
    const double fl = f0 + zl * (fp0 + zl * (f1 - f0 - fp0)), fh = f0 + zh * (fp0 + zh * (f1 - f0 - fp0));
    const double c = 2 * (f1 - f0 - fp0);
    double zmin = zl, fmin = fl;
    if (fh < fmin) { zmin = zh; fmin = fh; }
    if (c > 0) {
        const double z = -fp0 / c;
        if (z > zl && z < zh) {
            const double f = f0 + z * (fp0 + z * (f1 - f0 - fp0));
            if (f < fmin) { zmin = z; fmin = f; }
        }
    }
    return zmin;
}

double interp_cubic(double f0, double fp0, double f1, double fp1, double zl, double zh) {
    const double c0 = f0, c1 = fp0, c2 = 3 * (f1 - f0) - 2 * fp0 - fp1, c3 = fp0 + fp1 - 2 * (f1 - f0);
    auto cub = [&](double z) { return c0 + z * (c1 + z * (c2 + z * c3)); };
    double zmin = zl, fmin = cub(zl), z0 = 0, z1 = 0;
    auto check = [&](double z) {
        const double y = cub(z);
        if (y < fmin) { zmin = z; fmin = y; }
    };
    check(zh);
    const int nr = solve_quadratic(3 * c3, 2 * c2, c1, z0, z1);
    if (nr >= 1 && z0 > zl && z0 < zh) check(z0);
    if (nr == 2 && z1 > zl && z1 < zh) check(z1);
    return zmin;
}

double interpolate(double a, double fa, double fpa, double b, double fb, double fpb, double xmin, double xmax) {
    double zmin = (xmin - a) / (b - a), zmax = (xmax - a) / (b - a);
    if (zmin > zmax) std::swap(zmin, zmax);
    const double z = std::isfinite(fpb) ? interp_cubic(fa, fpa * (b - a), fb, fpb * (b - a), zmin, zmax)
                                        : interp_quad(fa, fpa * (b - a), fb, zmin, zmax);
    return a + z * (b - a);
}

// linear_minimize.c: Fletcher's line search from alpha = 0 (cached); alphaNew stays as it
// was when the trial budget runs out
Task<int> line_minimize(PatchRun& r, double alpha1, double& alphaNew) {
    const double rho = 0.01, sigma = r.bf.sigma, tau1 = 9, tau2 = 0.05, tau3 = 0.5;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    const FDf z = co_await lw_fdf(r, 0.0);
    const double f0 = z.f, fp0 = z.df;
    double falphaPrev = f0, fpalphaPrev = fp0, alpha = alpha1, alphaPrev = 0.0;
    double a = 0.0, b = alpha, fa = f0, fb = 0.0, fpa = fp0, fpb = 0.0, falpha, fpalpha;
    int i = 0;
    while (i++ < 100) {                                     // bracketing
        falpha = co_await lw_f(r, alpha);
        if (falpha > f0 + alpha * rho * fp0 || falpha >= falphaPrev) {
            a = alphaPrev; fa = falphaPrev; fpa = fpalphaPrev;
            b = alpha; fb = falpha; fpb = nan;
            break;
        }
        fpalpha = co_await lw_df(r, alpha);
        if (std::fabs(fpalpha) <= -sigma * fp0) {
            alphaNew = alpha;
            co_return ST_SUCCESS;
        }
        if (fpalpha >= 0) {
            a = alpha; fa = falpha; fpa = fpalpha;
            b = alphaPrev; fb = falphaPrev; fpb = fpalphaPrev;
            break;
        }
        const double delta = alpha - alphaPrev;
        const double next = interpolate(alphaPrev, falphaPrev, fpalphaPrev, alpha, falpha, fpalpha, alpha + delta,
                                        alpha + tau1 * delta);
        alphaPrev = alpha; falphaPrev = falpha; fpalphaPrev = fpalpha;
        alpha = next;
    }
    while (i++ < 100) {                                     // sectioning of [a, b]
        const double delta = b - a;
        alpha = interpolate(a, fa, fpa, b, fb, fpb, a + tau2 * delta, b - tau3 * delta);
        falpha = co_await lw_f(r, alpha);
        if ((a - alpha) * fpa <= DBL_EPS) co_return ST_ENOPROG;    // roundoff prevents progress
        if (falpha > f0 + rho * alpha * fp0 || falpha >= fa) {
            b = alpha; fb = falpha; fpb = nan;
        } else {
            fpalpha = co_await lw_df(r, alpha);
            if (std::fabs(fpalpha) <= -sigma * fp0) {
                alphaNew = alpha;
                co_return ST_SUCCESS;
            }
            if ((delta >= 0 && fpalpha >= 0) || (delta <= 0 && fpalpha <= 0)) {
                b = a; fb = fa; fpb = fpa;
            }
            a = alpha; fa = falpha; fpa = fpalpha;
        }
    }
    co_return ST_SUCCESS;
}

// vector_bfgs2_set at x = r.x: f and g (both used), first direction -g/|g|
Task<int> bfgs2_set(PatchRun& r, double& f, double stepSize, double tol) {
    auto& w = r.bf;
    const size_t n = r.x.size();
    for (auto* v : {&w.x0, &w.g0, &w.p, &w.xa, &w.ga, &w.dx0, &w.dg0}) v->assign(n, 0.0);
    w.step = stepSize;
    w.deltaF = 0;
    w.sigma = tol;
    f = co_await my_fdf(r, r.x.data(), r.g.data());
    w.x0 = r.x;
    w.g0 = r.g;
    w.g0norm = dnrm2(w.g0);
    for (size_t k = 0; k < n; ++k) w.p[k] = r.g[k] * (-1 / w.g0norm);
    w.pnorm = dnrm2(w.p);
    w.fp0 = -w.g0norm;
    w.xa = w.x0;                                         // the wrapper at alpha = 0
    w.xKey = 0.0;
    w.fa = f;
    w.fKey = 0.0;
    w.ga = w.g0;
    w.gKey = 0.0;
    w.dfa = dotv(w.ga, w.p);
    w.dfKey = 0.0;
    co_return ST_SUCCESS;
}

// vector_bfgs2_iterate: r.x, f, r.g move to the line minimum, r.dx is the step
Task<int> bfgs2_iterate(PatchRun& r, double& f) {
    auto& w = r.bf;
    const size_t n = r.x.size();
    double alpha = 0.0, alpha1;
    const double f0 = f;
    if (w.pnorm == 0.0 || w.g0norm == 0.0 || w.fp0 == 0) {
        std::fill(r.dx.begin(), r.dx.end(), 0.0);
        co_return ST_ENOPROG;
    }
    if (w.deltaF < 0) {
        const double del = std::max(-w.deltaF, 10 * DBL_EPS * std::fabs(f0));
        alpha1 = std::min(1.0, 2.0 * del / (-w.fp0));
    } else {
        alpha1 = std::fabs(w.step);
    }
    const int status = co_await line_minimize(r, alpha1, alpha);
    if (status != ST_SUCCESS) co_return status;
    co_await lw_fdf(r, alpha);                           // update_position: f and g at alpha
    f = w.fa;
    r.x = w.xa;
    r.g = w.ga;
    w.deltaF = f - f0;
    for (size_t k = 0; k < n; ++k) {
        w.dx0[k] = r.x[k] - w.x0[k];
        r.dx[k] = w.dx0[k];
        w.dg0[k] = r.g[k] - w.g0[k];
    }
    const double dxg = dotv(w.dx0, r.g), dgg = dotv(w.dg0, r.g), dxdg = dotv(w.dx0, w.dg0), dgnorm = dnrm2(w.dg0);
    double A = 0, B = 0;
    if (dxdg != 0) {
        B = dxg / dxdg;
        A = -(1.0 + dgnorm * dgnorm / dxdg) * B + dgg / dxdg;
    }
    for (size_t k = 0; k < n; ++k) w.p[k] = r.g[k] + (-A) * w.dx0[k] + (-B) * w.dg0[k];
    w.g0 = r.g;
    w.x0 = r.x;
    w.g0norm = dnrm2(w.g0);
    w.pnorm = dnrm2(w.p);
    const double pg = dotv(w.p, r.g), dir = (pg >= 0.0) ? -1.0 : +1.0;
    for (size_t k = 0; k < n; ++k) w.p[k] *= dir / w.pnorm;
    w.pnorm = dnrm2(w.p);
    w.fp0 = dotv(w.p, w.g0);
    w.xa = w.x0;                                         // change_direction: new origin, same f
    w.xKey = 0.0;
    w.fKey = 0.0;
    w.ga = w.g0;
    w.gKey = 0.0;
    w.dfa = dotv(w.ga, w.p);
    w.dfKey = 0.0;
    co_return ST_SUCCESS;
}

// ---- GETSTEP = steepest_descent and BFGS: GSL 1.16 steepest_descent and vector_bfgs
// (solver.cpp:355-361, 369-372), their published algorithms restated (multimin/
// steepest_descent.c, vector_bfgs.c, directional_minimize.c).  Steepest descent tries
// x - step g/|g|, shrinking the step by tol while f rises, then doubles it (or, after a
// shrink, shrinks it once more).  vector_bfgs tries the current step along the downhill
// direction; if f does not fall it finds a lower point by a parabolic guess from the slope
// (halving towards 0), then refines it by a safeguarded parabolic / golden-section search
// of at most 10 trials that stops when |p.g lambda/|g|| < tol; the next direction is the
// BFGS update p' = g1 - A dx - B dg, restarted along g every n iterations.
Task<int> step1_set(PatchRun& r, double& f, double stepSize, double tol, bool bfgs) {
    auto& w = r.sd;
    const size_t n = r.x.size();
    for (auto* v : {&w.x1, &w.g1, &w.dx1, &w.x2, &w.p, &w.x0, &w.g0, &w.dx0, &w.dg0}) v->assign(n, 0.0);
    w.iter = 0;
    w.step = stepSize;
    w.tol = tol;
    f = co_await my_fdf(r, r.x.data(), r.g.data());
    if (bfgs) {
        w.x0 = r.x;
        w.p = r.g;
        w.g0 = r.g;
        w.pnorm = w.g0norm = dnrm2(r.g);
    }
    co_return ST_SUCCESS;
}

Task<int> sd_iterate(PatchRun& r, double& f) {
    auto& w = r.sd;
    const size_t n = r.x.size();
    const double f0 = f, gnorm = dnrm2(r.g);
    double step = w.step, f1 = 0;
    bool shrunk = false;
    if (gnorm == 0.0) {
        std::fill(r.dx.begin(), r.dx.end(), 0.0);
        co_return ST_ENOPROG;
    }
    for (;;) {
        for (size_t k = 0; k < n; ++k) r.dx[k] = 0.0 + (-step / gnorm) * r.g[k];
        for (size_t k = 0; k < n; ++k) w.x1[k] = r.x[k] + 1.0 * r.dx[k];
        if (w.x1 == r.x) co_return ST_ENOPROG;
        f1 = co_await my_fdf(r, w.x1.data(), w.g1.data());
        if (f1 > f0) {
            shrunk = true;
            step *= w.tol;
            continue;
        }
        break;
    }
    w.step = shrunk ? step * w.tol : step * 2.0;
    r.x = w.x1;
    r.g = w.g1;
    f = f1;
    co_return ST_SUCCESS;
}

void take_step(const std::vector<double>& x, const std::vector<double>& p, double step, double lambda,
               std::vector<double>& x1, std::vector<double>& dx) {
    for (size_t k = 0; k < x.size(); ++k) dx[k] = 0.0 + (-step * lambda) * p[k];
    for (size_t k = 0; k < x.size(); ++k) x1[k] = x[k] + 1.0 * dx[k];
}

// a point in (0, stepc) below fa (directional_minimize.c: intermediate_point); the
// gradient at it goes to r.g.  Returns (step, f), step 0 if the trial does not move x.
Task<std::pair<double, double>> intermediate_point(PatchRun& r, double lambda, double pg, double stepc, double fa,
                                                   double fc) {
    auto& w = r.sd;
    double stepb, fb;
    for (;;) {
        const double u = std::fabs(pg * lambda * stepc);
        stepb = 0.5 * stepc * u / ((fc - fa) + u);
        take_step(r.x, w.p, stepb, lambda, w.x1, w.dx1);
        if (w.x1 == r.x) {
            co_await my_df(r, w.x1.data(), r.g.data());
            co_return std::make_pair(0.0, fa);
        }
        fb = co_await my_f(r, w.x1.data());
        if (fb >= fa && stepb > 0.0) {
            fc = fb;
            stepc = stepb;
            continue;
        }
        break;
    }
    co_await my_df(r, w.x1.data(), r.g.data());
    co_return std::make_pair(stepb, fb);
}

// directional_minimize.c: minimize; the best point in w.x2 (its step in r.dx), f and the
// gradient norm there out, w.step set to its step
Task<std::pair<double, double>> dir_minimize(PatchRun& r, double lambda, double stepa, double stepb, double stepc,
                                             double fa, double fb, double fc) {
    auto& w = r.sd;
    const size_t n = r.x.size();
    double u = stepb, v = stepa, ww = stepc, fu = fb, fv = fa, fw = fc;
    double old2 = std::fabs(ww - v), old1 = std::fabs(v - u);
    w.x2 = w.x1;
    r.dx = w.dx1;
    double fOut = fb, gnorm = dnrm2(r.g);
    w.step = stepb;
    for (int iter = 1; iter <= 10; ++iter) {
        const double dw = ww - u, dv = v - u;
        const double e1 = (fv - fu) * dw * dw + (fu - fw) * dv * dv;
        const double e2 = 2.0 * ((fv - fu) * dw + (fu - fw) * dv);
        const double du = e2 != 0.0 ? e1 / e2 : 0.0;
        double stepm;
        if (du > 0.0 && du < stepc - stepb && std::fabs(du) < 0.5 * old2) stepm = u + du;
        else if (du < 0.0 && du > stepa - stepb && std::fabs(du) < 0.5 * old2) stepm = u + du;
        else if (stepc - stepb > stepb - stepa) stepm = 0.38 * (stepc - stepb) + stepb;
        else stepm = stepb - 0.38 * (stepb - stepa);
        take_step(r.x, w.p, stepm, lambda, w.x1, w.dx1);
        const double fm = co_await my_f(r, w.x1.data());
        if (fm > fb) {
            if (fm < fv) { ww = v; v = stepm; fw = fv; fv = fm; }
            else if (fm < fw) { ww = stepm; fw = fm; }
            if (stepm < stepb) { stepa = stepm; fa = fm; }
            else { stepc = stepm; fc = fm; }
            continue;
        }
        old2 = old1;
        old1 = std::fabs(u - stepm);
        ww = v; v = u; u = stepm;
        fw = fv; fv = fu; fu = fm;
        w.x2 = w.x1;
        r.dx = w.dx1;
        co_await my_df(r, w.x1.data(), r.g.data());
        double pg = 0;
        for (size_t k = 0; k < n; ++k) pg += w.p[k] * r.g[k];
        const double gnorm1 = dnrm2(r.g);
        fOut = fm;
        w.step = stepm;
        gnorm = gnorm1;
        if (std::fabs(pg * lambda / gnorm1) < w.tol) break;
        if (stepm < stepb) { stepc = stepb; fc = fb; stepb = stepm; fb = fm; }
        else { stepa = stepb; fa = fb; stepb = stepm; fb = fm; }
    }
    co_return std::make_pair(fOut, gnorm);
}

Task<int> bfgs1_iterate(PatchRun& r, double& f) {
    auto& w = r.sd;
    const size_t n = r.x.size();
    const double fa = f, stepc = w.step;
    if (w.pnorm == 0.0 || w.g0norm == 0.0) {
        std::fill(r.dx.begin(), r.dx.end(), 0.0);
        co_return ST_ENOPROG;
    }
    double pg = 0;
    for (size_t k = 0; k < n; ++k) pg += w.p[k] * r.g[k];
    const double dir = pg >= 0.0 ? +1.0 : -1.0, lambda = dir / w.pnorm;
    take_step(r.x, w.p, stepc, lambda, w.x1, r.dx);
    const double fc = co_await my_f(r, w.x1.data());
    if (fc < fa) {                                       // the trial step is downhill: take it
        w.step = stepc * 2.0;
        f = fc;
        r.x = w.x1;
        co_await my_df(r, w.x1.data(), r.g.data());
        co_return ST_SUCCESS;
    }
    const auto [stepb, fb] = co_await intermediate_point(r, lambda, pg, stepc, fa, fc);
    if (stepb == 0.0) co_return ST_ENOPROG;
    const auto [fm, g1norm] = co_await dir_minimize(r, lambda, 0.0, stepb, stepc, fa, fb, fc);
    f = fm;
    r.x = w.x2;
    w.iter = (w.iter + 1) % (int)n;
    if (w.iter == 0) {                                   // restart along the gradient
        w.p = r.g;
        w.pnorm = g1norm;
    } else {
        for (size_t k = 0; k < n; ++k) {
            w.dx0[k] = r.x[k] + (-1.0) * w.x0[k];
            w.dg0[k] = r.g[k] + (-1.0) * w.g0[k];
        }
        const double dxg = dotv(w.dx0, r.g), dgg = dotv(w.dg0, r.g), dxdg = dotv(w.dx0, w.dg0), dgnorm = dnrm2(w.dg0);
        double A = 0, B = 0;
        if (dxdg != 0) {
            B = dxg / dxdg;
            A = -(1.0 + dgnorm * dgnorm / dxdg) * B + dgg / dxdg;
        }
        for (size_t k = 0; k < n; ++k) w.p[k] = r.g[k] + (-A) * w.dx0[k] + (-B) * w.dg0[k];
        w.pnorm = dnrm2(w.p);
    }
    w.g0 = r.g;
    w.x0 = r.x;
    w.g0norm = dnrm2(w.g0);
    co_return ST_SUCCESS;
}

// Solver::run (solver.cpp:330-580), starting after the reset initialisation
Task<int> run_patch(PatchRun& r) {
    mfbd_ctx* c = r.c;
    const mfbd_params& pr = c->prm;
    const int K = c->g.K, nModes = r.K;
    const bool adjoint = pr.gradient == MFBD_GRAD_ADJOINT;
    const bool bfgs = pr.getstep == MFBD_GETSTEP_BFGS_INV;
    const bool sdesc = pr.getstep == MFBD_GETSTEP_STEEPEST, bfgs1 = pr.getstep == MFBD_GETSTEP_BFGS;
    double maxWl = 0;                                   // init_step = 1e-4 lambda_max (solver.cpp:351)
    for (int o = 0; o < c->g.nObj; ++o) maxWl = std::max(maxWl, c->st.wavelength[o]);
    const double initStep = 1E-4 * maxWl;
    r.gradMethod = adjoint ? MFBD_GRAD_ADJOINT : MFBD_GRAD_FD;   // GM_DIFF for the first gradient
    double prev = 0, thisMetric = 0, gradNorm = 0, tol = 1.0E-1;
    int failCount = 0, totalIter = 0, maxIterations = 10, status = 0;
    const int maxFails = 3;
    bool lastDone = false;
    int lastBreak = 0;          // how the last stage's loop ended early: 1 no progress, 2 non-finite
    double initial = co_await my_f(r, r.beta0.data());
    r.metrics.push_back(initial);
    int nImgP = 0;
    for (int j = 0; j < c->g.nImg; ++j) nImgP += img_on(c, r.p, j);
    const double nTotalPixels = (double)nImgP * c->g.s * c->g.s;
    const double gradScale = 1.0 / (nTotalPixels * nTotalPixels);
    bool exitLoop = false;
    for (int modeCount = pr.modeStart; !exitLoop; modeCount += pr.modeStep) {
        modeCount = std::min(modeCount, nModes);
        r.nEnabled = std::max(r.nEnabled, modeCount);
        if (modeCount == nModes || pr.maxIter == 0) {
            exitLoop = true;
            tol = pr.FTOL;
            maxIterations = pr.maxIter;
        }
        // gsl_multimin_fdfminimizer_set at x = 0 with the stepper's set: conjugate_rdx_set
        // (f(x); the gradient the reference also computes here is overwritten before use)
        // or vector_bfgs2_set (f and g, both used)
        std::fill(r.x.begin(), r.x.end(), 0.0);
        std::fill(r.dx.begin(), r.dx.end(), 0.0);
        r.iterCG = 0;
        double f = 0;
        if (bfgs) co_await bfgs2_set(r, f, initStep, 1E-6);
        else if (sdesc) co_await step1_set(r, f, initStep, 1E-6, false);
        else if (bfgs1) co_await step1_set(r, f, 1E8, 0.01, true);      // solver.cpp:359-360
        else f = co_await my_f(r, r.x.data());
        int iter = 0, successCount = 0;
        bool done = false;
        lastBreak = 0;
        while ((!done || iter < pr.minIter) && iter < maxIterations) {
            iter++;
            status = bfgs ? co_await bfgs2_iterate(r, f)
                     : sdesc ? co_await sd_iterate(r, f)
                     : bfgs1 ? co_await bfgs1_iterate(r, f) : co_await cg_iterate(r, f);
            if (failCount > maxFails) { status = ST_FAILURE; exitLoop = true; done = true; break; }
            if (!adjoint) r.gradMethod = MFBD_GRAD_VOGEL;   // Vogel from here on (solver.cpp:478)
            if (status == ST_ENOPROG) { failCount++; lastBreak = 1; break; }
            thisMetric = f;
            if (!std::isfinite(thisMetric)) { failCount++; lastBreak = 2; break; }
            r.metrics.push_back(thisMetric);
            gradNorm = dnrm2(r.g) * gradScale;
            if (!std::isfinite(gradNorm)) { failCount++; lastBreak = 2; break; }
            if (gradNorm < pr.FTOL) { exitLoop = true; done = true; break; }
            if (iter > 1) {
                double rel = 2.0 * std::fabs(thisMetric - prev) / (std::fabs(thisMetric) + std::fabs(prev) + pr.EPS);
                if (rel < tol) {
                    if (++successCount >= pr.nDoneIter) { successCount = 0; done = true; }
                } else successCount = 0;
            }
            prev = thisMetric;
        }
        thisMetric = co_await my_f(r, r.x.data());
        totalIter += iter;
        lastDone = done;
        if (!exitLoop) {
            // Solver::shiftAndInit() + alpha_offset <- alpha
            r.shiftedObjs.assign(c->g.nObj, 0);
            for (int j = 0; j < c->g.nImg; ++j)
                if (adjust_shift(c, r.p, j, &r.alpha[(size_t)j * K])) r.shiftedObjs[c->st.imgObj[j]] = 1;
            bool any = false;
            for (int v : r.shiftedObjs) any |= v != 0;
            if (any) co_await InitAwait{&r};
            r.alphaOffset = r.alpha;
        }
    }
    r.res.initialMetric = initial;
    r.res.finalMetric = thisMetric;
    r.res.iterations = totalIter;
    r.res.failCount = failCount;
    // the exits of solver.cpp:470-516, told apart: more than maxFails failures; the last
    // stage ended by GSL_ENOPROG or by a non-finite metric or gradient; converged by the
    // relative-change or gradient-norm test; or stopped at its iteration limit
    r.res.status = (status == ST_FAILURE) ? MFBD_PATCH_FAILED
                   : lastBreak == 1 ? MFBD_PATCH_NO_PROGRESS
                   : lastBreak == 2 ? MFBD_PATCH_NONFINITE
                   : (lastDone || maxIterations == 0) ? MFBD_PATCH_CONVERGED : MFBD_PATCH_MAXITER;
    co_return 0;
}

// re-initialisation of the shifted objects' images (Object::reInitialize, doReset = false)
void shift_inits(mfbd_ctx* c, std::vector<PatchRun*>& rs, int lane) {
    std::vector<ImgInit> imgs;
    for (auto* r : rs)
        for (int j = 0; j < c->g.nImg; ++j)
            if (r->shiftedObjs[c->st.imgObj[j]] && img_on(c, r->p, j))
                imgs.push_back({r->p, j, c->shifts[r->p][2 * j], c->shifts[r->p][2 * j + 1]});
    c->dev->init_images(imgs, c->window, lane);
}

// Evaluation lanes: each a contiguous share of the batch on a stream of its own; while the
// device evaluates one share, the host resumes the others' patches.  MFBD_EVAL_LANES
// (1..kMaxEvalLanes) overrides the default.
// Evaluation lanes of a batch: MFBD_EVAL_LANES, else 4 for FP32 / MIXED (2 % faster than 2)
// and for FP64 with the device's line search (256 config-2 patches: 263 / 268 patches/s at
// 3 / 4 lanes), 2 for FP64 with the host's (its OTF kernel fills the device; more lanes
// measured 7 % slower)
int eval_lanes(const mfbd_params& p) {
    const char* e = getenv("MFBD_EVAL_LANES");
    const bool deviceLS = p.lineRounds > 0 && p.linePhases && p.getstep == MFBD_GETSTEP_CONJUGATE;
    const int v = e ? atoi(e) : (p.precision == MFBD_FP64 && !deviceLS ? 2 : 4);
    return std::max(1, std::min(v, kMaxEvalLanes));
}

int check_ready(mfbd_ctx* c) {
    if (!c) return MFBD_EINVAL;
    if (!c->haveStructure || !c->haveBasis || !c->haveCons || c->nPatch == 0)
        return set_err(c, MFBD_ESTATE, "structure, basis, constraints and patches must be set first");
    return MFBD_OK;
}

void ensure_init(mfbd_ctx* c) {
    if (c->initDone || c->solved) return;
    std::vector<int> all(c->nPatch);
    for (int p = 0; p < c->nPatch; ++p) {
        all[p] = p;
        c->shifts[p].assign(2 * c->g.nImg, 0);
        c->tiltOff[p].assign(2 * c->g.nImg, 0.0);
    }
    reset_init(c, all);
    c->initDone = true;
}

}  // namespace

// ------------------------------------------------------------------- C-ABI ----
extern "C" {

int mfbd_ctx_create(const mfbd_geometry* geo, const mfbd_params* prm, int device, mfbd_ctx** out) {
    if (!geo || !prm || !out) return MFBD_EINVAL;
    *out = nullptr;
    const int P = geo->P;
    if (geo->s < 4 || P < 8 || geo->K < 1 || geo->K > 160 || geo->nObj < 1 || geo->nCh < geo->nObj ||
        geo->nImg < geo->nCh || geo->dataY < geo->s || geo->dataX < geo->s || geo->startY < 0 || geo->startX < 0 ||
        geo->startY + geo->s > geo->dataY || geo->startX + geo->s > geo->dataX || 2 * P < geo->s)
        return MFBD_EINVAL;
    if (prm->precision < MFBD_FP64 || prm->precision > MFBD_MIXED || prm->maxBatch < 1) return MFBD_EINVAL;
    auto* c = new mfbd_ctx();
    c->g = {geo->s, P, 2 * P, geo->K, geo->nObj, geo->nCh, geo->nImg, geo->dataY, geo->dataX, geo->startY, geo->startX};
    c->prm = *prm;
    if (c->prm.maxBatch < 1) c->prm.maxBatch = 1;
    if (c->prm.gradient != MFBD_GRAD_SCHEDULE_REFERENCE && c->prm.gradient != MFBD_GRAD_ADJOINT) {
        delete c;
        return MFBD_EINVAL;
    }
    // 0 (unset) falls to the conjugate gradient, as the reference's switch default does
    if (c->prm.getstep == 0) c->prm.getstep = MFBD_GETSTEP_CONJUGATE;
    if (c->prm.getstep < MFBD_GETSTEP_STEEPEST || c->prm.getstep > MFBD_GETSTEP_BFGS_INV || c->prm.lineRounds < 0) {
        delete c;
        return MFBD_EINVAL;
    }
    c->device = device;
    *out = c;
    return MFBD_OK;
}

int mfbd_ctx_destroy(mfbd_ctx* c) {
    delete c;
    return MFBD_OK;
}

const char* mfbd_last_error(const mfbd_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mfbd_set_structure(mfbd_ctx* c, const int* chObj, const int* imgCh, const double* objWeight,
                       const double* objWavelength, const double* freqCutoff) {
    if (!c || !chObj || !imgCh || !objWeight || !objWavelength || !freqCutoff) return set_err(c, MFBD_EINVAL, "null argument");
    Structure s;
    s.chObj.assign(chObj, chObj + c->g.nCh);
    s.imgCh.assign(imgCh, imgCh + c->g.nImg);
    s.weight.assign(objWeight, objWeight + c->g.nObj);
    s.wavelength.assign(objWavelength, objWavelength + c->g.nObj);
    s.freqCutoff.assign(freqCutoff, freqCutoff + c->g.nObj);
    s.nImgObj.assign(c->g.nObj, 0);
    for (int j = 0; j < c->g.nImg; ++j) {
        if (s.imgCh[j] < 0 || s.imgCh[j] >= c->g.nCh) return set_err(c, MFBD_EINVAL, "imgCh out of range");
        if (j && s.imgCh[j] < s.imgCh[j - 1]) return set_err(c, MFBD_EINVAL, "images must be ordered by channel");
        const int o = s.chObj[s.imgCh[j]];
        if (o < 0 || o >= c->g.nObj) return set_err(c, MFBD_EINVAL, "chObj out of range");
        s.imgObj.push_back(o);
        s.nImgObj[o]++;
    }
    c->st = s;
    c->haveStructure = true;
    return MFBD_OK;
}

int mfbd_set_basis(mfbd_ctx* c, const double* pupil, const double* modes, const int* tiltXY, const double* jac) {
    if (!c || !pupil || !modes || !tiltXY || !jac) return set_err(c, MFBD_EINVAL, "null argument");
    if (!c->haveStructure) return set_err(c, MFBD_ESTATE, "mfbd_set_structure must precede mfbd_set_basis");
    const size_t P2 = (size_t)c->g.P * c->g.P;
    c->pupil.assign(pupil, pupil + P2);
    c->modes.assign(modes, modes + (size_t)c->g.nObj * c->g.K * P2);
    c->tilt.assign(tiltXY, tiltXY + 2 * c->g.nObj);
    c->jac.assign(jac, jac + 8 * c->g.nObj);
    for (int o = 0; o < c->g.nObj; ++o)
        for (int t = 0; t < 2; ++t)
            if (c->tilt[2 * o + t] >= c->g.K) return set_err(c, MFBD_EINVAL, "tilt mode index out of range");
    const double nrm = otf_normalization(c);
    c->otfNorm.assign(c->g.nCh, nrm);
    c->st.graddiffStep = c->prm.graddiffStep;
    // Solver::init windows (solver.cpp:112-126)
    const int s = c->g.s;
    c->window.assign((size_t)s * s, 1.0);
    c->noiseWindow.assign((size_t)s * s, 1.0);
    const int apod = c->prm.apodSize < 0 ? s / 8 : c->prm.apodSize;
    if (apod) {
        apodize_in_place(c->window.data(), s, s, apod);
        apodize_in_place(c->noiseWindow.data(), s, s, apod / 2);
    }
    c->haveBasis = true;
    return MFBD_OK;
}

int mfbd_set_constraints(mfbd_ctx* c, int nPar, int nFree, int nnz, const int* rows, const int* cols,
                         const double* vals) {
    if (!c || !rows || !cols || !vals) return set_err(c, MFBD_EINVAL, "null argument");
    if (nPar != c->g.nImg * c->g.K || nFree < 1 || nFree > nPar || nnz < 1)
        return set_err(c, MFBD_EINVAL, "nullspace has the wrong number of entries");   // solver.cpp:102-110
    for (int e = 0; e < nnz; ++e)
        if (rows[e] < 0 || rows[e] >= nPar || cols[e] < 0 || cols[e] >= nFree)
            return set_err(c, MFBD_EINVAL, "nullspace entry out of range");
    c->nPar = nPar;
    c->nFree = nFree;
    c->ns.nFree = nFree;
    c->ns.row.assign(rows, rows + nnz);
    c->ns.col.assign(cols, cols + nnz);
    c->ns.val.assign(vals, vals + nnz);
    c->haveCons = true;
    return MFBD_OK;
}

int mfbd_load_patches(mfbd_ctx* c, int n, const float* data, const double* phiFixed, const double* resOffset,
                      const double* noiseFudge) {
    if (!c || n < 1 || !data || !phiFixed || !noiseFudge) return set_err(c, MFBD_EINVAL, "bad arguments");
    if (!c->haveBasis) return set_err(c, MFBD_ESTATE, "mfbd_set_basis must precede mfbd_load_patches");
    const size_t per = (size_t)c->g.nImg * c->g.dataY * c->g.dataX;
    c->nPatch = n;
    c->data.assign(data, data + per * n);
    c->phiFixed.assign(phiFixed, phiFixed + (size_t)c->g.nCh * c->g.P * c->g.P);
    if (resOffset) c->resOffset.assign(resOffset, resOffset + (size_t)n * c->g.nCh * 2);
    else c->resOffset.assign((size_t)n * c->g.nCh * 2, 0.0);
    c->noiseFudge.assign(noiseFudge, noiseFudge + c->g.nCh);
    c->shifts.assign(n, {});
    c->tiltOff.assign(n, {});
    c->noise.assign(n, {});
    c->gamma.assign(n, {});
    c->nModesP.clear();
    c->imgOn.clear();
    c->nsP.clear();
    c->initDone = false;
    c->solved = false;
    try {
        // device slots: the evaluator keeps every loaded patch resident
        std::string err;
        c->dev = make_device_evaluator(c->g, c->prm.precision, n, c->prm.maxBatch, c->device, err);
        if (!c->dev) return set_err(c, MFBD_EDEVICE, "device pipeline unavailable: " + err);
        c->dev->set_basis(c->pupil.data(), c->modes.data());
        Structure s = c->st;
        c->dev->set_structure(s, c->otfNorm);
        std::vector<double> phi;
        const size_t perPhi = (size_t)c->g.nCh * c->g.P * c->g.P;
        std::vector<double> all(perPhi * n);
        for (int p = 0; p < n; ++p) {
            phi_channels(c, p, phi);
            std::copy(phi.begin(), phi.end(), all.begin() + perPhi * p);
        }
        c->dev->load(0, n, data, all.data());
    } catch (const std::exception& e) {
        return set_err(c, MFBD_EDEVICE, e.what());
    }
    return MFBD_OK;
}

int mfbd_set_patch_shape(mfbd_ctx* c, int patch, int nModes, const unsigned char* imgOn, int nFree, int nnz,
                         const int* rows, const int* cols, const double* vals) {
    if (!c || !imgOn || !rows || !cols || !vals) return set_err(c, MFBD_EINVAL, "null argument");
    if (!c->dev || !c->haveCons) return set_err(c, MFBD_ESTATE, "constraints and patches must be set first");
    if (patch < 0 || patch >= c->nPatch) return set_err(c, MFBD_EINVAL, "patch index out of range");
    const int K = c->g.K, nImg = c->g.nImg, nPar = c->nPar;
    if (nModes < 1 || nModes > K) return set_err(c, MFBD_EINVAL, "nModes out of range");
    std::vector<int> perObj(c->g.nObj, 0);
    for (int j = 0; j < nImg; ++j) perObj[c->st.imgObj[j]] += imgOn[j] ? 1 : 0;
    for (int v : perObj)
        if (v == 0) return set_err(c, MFBD_EINVAL, "every object needs at least one image");
    if (nFree < 1 || nFree > nPar || nnz < 1) return set_err(c, MFBD_EINVAL, "nullspace has the wrong number of entries");
    for (int e = 0; e < nnz; ++e) {
        if (rows[e] < 0 || rows[e] >= nPar || cols[e] < 0 || cols[e] >= nFree)
            return set_err(c, MFBD_EINVAL, "nullspace entry out of range");
        if (!imgOn[rows[e] / K] || rows[e] % K >= nModes)
            return set_err(c, MFBD_EINVAL, "nullspace entry outside the patch's images and modes");
    }
    try {
        if (c->nModesP.empty()) {
            c->nModesP.assign(c->nPatch, K);
            c->imgOn.assign(c->nPatch, std::vector<unsigned char>(nImg, 1));
            c->nsP.assign(c->nPatch, NullSpace{});
        }
        c->nModesP[patch] = nModes;
        for (int j = 0; j < nImg; ++j) c->imgOn[patch][j] = imgOn[j] ? 1 : 0;
        NullSpace& ns = c->nsP[patch];
        ns.nFree = nFree;
        ns.row.assign(rows, rows + nnz);
        ns.col.assign(cols, cols + nnz);
        ns.val.assign(vals, vals + nnz);
        // the frames of a missing image are zero, so its transform D is zero too
        const size_t fr = (size_t)c->g.dataY * c->g.dataX;
        float* blk = &c->data[(size_t)patch * nImg * fr];
        for (int j = 0; j < nImg; ++j)
            if (!imgOn[j]) std::fill(blk + j * fr, blk + (j + 1) * fr, 0.0f);
        std::vector<double> phi;
        phi_channels(c, patch, phi);
        c->dev->load(patch, 1, blk, phi.data());
        c->dev->set_images_enabled(patch, c->imgOn[patch].data());
        c->initDone = false;
        c->solved = false;
    } catch (const std::exception& e) {
        return set_err(c, MFBD_EDEVICE, e.what());
    }
    return MFBD_OK;
}

int mfbd_init_patches(mfbd_ctx* c) {
    int rc = check_ready(c);
    if (rc) return rc;
    try {
        c->initDone = false;
        c->solved = false;
        ensure_init(c);
    } catch (const std::exception& e) {
        return set_err(c, MFBD_EDEVICE, e.what());
    }
    return MFBD_OK;
}

int mfbd_eval_metric(mfbd_ctx* c, int n, const int* patchIdx, const double* alpha, double* L) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (n < 1 || !patchIdx || !alpha || !L) return set_err(c, MFBD_EINVAL, "bad arguments");
    try {
        ensure_init(c);
        const size_t per = (size_t)c->g.nImg * c->g.K;
        std::vector<std::vector<double>> ae(n, std::vector<double>(per));
        std::vector<EvalReq> rq(n);
        for (int k = 0; k < n; ++k) {
            if (patchIdx[k] < 0 || patchIdx[k] >= c->nPatch) return set_err(c, MFBD_EINVAL, "patch index out of range");
            effective_alpha(c, patchIdx[k], alpha + per * k, ae[k].data());
            rq[k].slot = patchIdx[k];
            rq[k].alpha = ae[k].data();
        }
        c->dev->evaluate(rq);
        for (int k = 0; k < n; ++k) L[k] = rq[k].f;
    } catch (const std::exception& e) {
        return set_err(c, MFBD_EDEVICE, e.what());
    }
    return MFBD_OK;
}

int mfbd_eval_grad(mfbd_ctx* c, int n, const int* patchIdx, const double* alpha, int nEnabled, int method,
                   double* gAlpha) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (n < 1 || !patchIdx || !alpha || !gAlpha || method < 1 || method > 3 || nEnabled < 0 || nEnabled > c->g.K)
        return set_err(c, MFBD_EINVAL, "bad arguments");
    try {
        ensure_init(c);
        const size_t per = (size_t)c->g.nImg * c->g.K;
        std::vector<std::vector<double>> ae(n, std::vector<double>(per));
        std::vector<EvalReq> rq(n);
        for (int k = 0; k < n; ++k) {
            if (patchIdx[k] < 0 || patchIdx[k] >= c->nPatch) return set_err(c, MFBD_EINVAL, "patch index out of range");
            effective_alpha(c, patchIdx[k], alpha + per * k, ae[k].data());
            rq[k].slot = patchIdx[k];
            rq[k].alpha = ae[k].data();
            rq[k].wantGrad = true;
            rq[k].gradMethod = method;
            rq[k].nEnabled = nEnabled;
            rq[k].g = gAlpha + per * k;
        }
        c->dev->evaluate(rq);
    } catch (const std::exception& e) {
        return set_err(c, MFBD_EDEVICE, e.what());
    }
    return MFBD_OK;
}

int mfbd_restore(mfbd_ctx* c, int n, const int* patchIdx, const double* alpha, float* img) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (n < 1 || !patchIdx || !alpha || !img) return set_err(c, MFBD_EINVAL, "bad arguments");
    try {
        ensure_init(c);
        const size_t per = (size_t)c->g.nImg * c->g.K;
        for (int k = 0; k < n; ++k)
            if (patchIdx[k] < 0 || patchIdx[k] >= c->nPatch) return set_err(c, MFBD_EINVAL, "patch index out of range");
        std::vector<double> ae(per * n);
        std::vector<const double*> ap(n);
        for (int k = 0; k < n; ++k) {
            effective_alpha(c, patchIdx[k], alpha + per * k, &ae[per * k]);
            ap[k] = &ae[per * k];
        }
        const std::vector<double> nf = restore_noise(c, patchIdx, n), c2 = restore_cutoffs(c);
        c->dev->restore_objects_begin(n, patchIdx, ap.data(), nf.data(), c2.data(), !c->prm.noFilter, false);
        c->dev->restore_objects_end(img);
    } catch (const std::exception& e) {
        return set_err(c, MFBD_EDEVICE, e.what());
    }
    return MFBD_OK;
}

int mfbd_restore_outputs(mfbd_ctx* c, int n, const int* patchIdx, const double* alpha, int saveMask, float* psf,
                         float* cobj, float* res) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (n < 1 || !patchIdx || !alpha) return set_err(c, MFBD_EINVAL, "bad arguments");
    for (int k = 0; k < n; ++k)
        if (patchIdx[k] < 0 || patchIdx[k] >= c->nPatch) return set_err(c, MFBD_EINVAL, "patch index out of range");
    int mask = saveMask & (MFBD_SAVE_COBJ | MFBD_SAVE_PSF | MFBD_SAVE_PSF_AVG | MFBD_SAVE_RESIDUAL);
    if ((mask & MFBD_SAVE_PSF) && (mask & MFBD_SAVE_PSF_AVG)) mask &= ~MFBD_SAVE_PSF;   // config.cpp:497-500
    if (((mask & (MFBD_SAVE_PSF | MFBD_SAVE_PSF_AVG)) && !psf) || ((mask & MFBD_SAVE_COBJ) && !cobj) ||
        ((mask & MFBD_SAVE_RESIDUAL) && !res))
        return set_err(c, MFBD_EINVAL, "an output asked for by saveMask is null");
    if (!mask) return MFBD_OK;
    try {
        ensure_init(c);
        const int nImg = c->g.nImg, nObj = c->g.nObj;
        const size_t per = (size_t)nImg * c->g.K;
        const size_t N2 = (size_t)4 * c->g.P * c->g.P;
        std::vector<double> ae(per * n);
        std::vector<const double*> ap(n);
        std::vector<std::vector<ImgInit>> imgs(n);
        for (int k = 0; k < n; ++k) {
            const int p = patchIdx[k];
            effective_alpha(c, p, alpha + per * k, &ae[per * k]);
            ap[k] = &ae[per * k];
            for (int j = 0; j < nImg; ++j)
                if (img_on(c, p, j)) imgs[k].push_back({p, j, c->shifts[p][2 * j], c->shifts[p][2 * j + 1]});
        }
        std::vector<float> objFull;
        if (mask & (MFBD_SAVE_COBJ | MFBD_SAVE_RESIDUAL)) {
            // the restored objects on the whole grid (od.img before the crop)
            const std::vector<double> nf = restore_noise(c, patchIdx, n), c2 = restore_cutoffs(c);
            objFull.resize((size_t)n * nObj * N2);
            c->dev->restore_objects_begin(n, patchIdx, ap.data(), nf.data(), c2.data(), !c->prm.noFilter, true);
            c->dev->restore_objects_end(objFull.data());
        }
        // finished on the device in the caller's layout: PSF per image (SubImage::getPSF) or
        // summed in double per object, rounded, then / imgCount; the convolved object; the
        // float convolved object minus the frame; zero for images a patch lacks
        const int want = ((mask & MFBD_SAVE_PSF) ? DeviceEvaluator::kOutPsf : 0) |
                         ((mask & MFBD_SAVE_PSF_AVG) ? DeviceEvaluator::kOutPsfAvg : 0) |
                         ((mask & MFBD_SAVE_COBJ) ? DeviceEvaluator::kOutCobj : 0) |
                         ((mask & MFBD_SAVE_RESIDUAL) ? DeviceEvaluator::kOutRes : 0);
        c->dev->restore_extra(n, patchIdx, ap.data(), imgs, objFull, want, psf, cobj, res);
    } catch (const std::exception& e) {
        return set_err(c, MFBD_EDEVICE, e.what());
    }
    return MFBD_OK;
}

int mfbd_solve(mfbd_ctx* c, const double* alpha0, mfbd_result* results, double* alphaOut, float* imgOut) {
    int rc = check_ready(c);
    if (rc) return rc;
    try {
        const int K = c->g.K, nImg = c->g.nImg, nPar = c->nPar;
        const int batch = std::max(1, std::min(c->prm.maxBatch, c->nPatch));
        for (int b0 = 0; b0 < c->nPatch; b0 += batch) {
            const int nb = std::min(batch, c->nPatch - b0);
            std::vector<std::unique_ptr<PatchRun>> runs;
            std::vector<int> ids;
            for (int k = 0; k < nb; ++k) {
                auto r = std::make_unique<PatchRun>();
                r->c = c;
                r->p = b0 + k;
                r->ns = &ns_of(c, r->p);
                r->K = c->nModesP.empty() ? K : c->nModesP[r->p];
                const int nFree = r->ns->nFree;
                r->alpha.assign(nPar, 0.0);
                r->alphaOffset.assign(nPar, 0.0);
                if (alpha0) std::copy(alpha0 + (size_t)r->p * nPar, alpha0 + (size_t)(r->p + 1) * nPar, r->alphaOffset.begin());
                r->x.assign(nFree, 0.0);
                r->g.assign(nFree, 0.0);
                r->dx.assign(nFree, 0.0);
                r->x1.assign(nFree, 0.0);
                r->p_.assign(nFree, 0.0);
                r->prevG.assign(nFree, 0.0);
                r->beta0.assign(nFree, 0.0);
                r->ev.alphaEff = c->dev->alpha_staging() + (size_t)r->p * nPar;
                r->ev.gAlpha.assign(nPar, 0.0);
                c->shifts[r->p].assign(2 * nImg, 0);
                c->tiltOff[r->p].assign(2 * nImg, 0.0);
                ids.push_back(r->p);
                runs.push_back(std::move(r));
            }
            // shiftAndInit(alpha_offset, true): shifts from the warm start, then reset init
            for (auto& r : runs)
                for (int j = 0; j < nImg; ++j) adjust_shift(c, r->p, j, &r->alphaOffset[(size_t)j * K]);
            using clk = std::chrono::steady_clock;
            auto since = [](clk::time_point t0) {
                return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
            };
            c->hostStats[0].name = "host_evaluate";
            c->hostStats[1].name = "host_resume";
            c->hostStats[2].name = "host_reinit";
            c->hostStats[3].name = "host_reset_init";
            c->hostStats[4].name = "host_restore";
            c->hostStats[5].name = "host_restore_objects";
            c->hostStats[6].name = "host_line_search";
            {
                auto t0 = clk::now();
                reset_init(c, ids);
                c->hostStats[3].ms += since(t0);
                c->hostStats[3].launches++;
                c->hostStats[3].units += nb;
            }
            const int nLanes = std::max(1, std::min(eval_lanes(c->prm), nb));
            auto laneOf = [&](int k) { return (int)((long long)k * nLanes / nb); };
            // device line searches: per lane one captured round over the lane's patches
            std::vector<int> lsIdx(nb, -1);
            c->deviceLS = false;
            if (c->prm.lineRounds > 0 && c->prm.linePhases && c->prm.getstep == MFBD_GETSTEP_CONJUGATE) {
                c->deviceLS = true;
                for (int e = 0; e < nLanes && c->deviceLS; ++e) {
                    std::vector<int> slots;
                    for (int k = 0; k < nb; ++k)
                        if (laneOf(k) == e) {
                            lsIdx[k] = (int)slots.size();
                            slots.push_back(runs[k]->p);
                        }
                    c->deviceLS = c->dev->ls_begin(e, slots, c->prm.lineGraph != 0);
                }
                if (!c->deviceLS) c->dev->ls_clear();
            }
            std::vector<Task<int>> tasks;
            for (auto& r : runs) tasks.push_back(run_patch(*r));
            if (!c->pool) c->pool = std::make_unique<Pool>(
                (int)std::min<unsigned>(std::max(2u, std::thread::hardware_concurrency()), 16u) - 1);
            c->pool->run((int)tasks.size(), [&](int k) { tasks[k].h.resume(); });
            // The evaluation lanes take contiguous shares of the batch.  While the device
            // evaluates one share (and, on their own streams, the lanes' kernels overlap),
            // the host resumes another share's patches (their CG / line-search steps) and
            // enqueues their next requests.  Each patch still sees exactly its own sequence
            // of requests.
            std::vector<EvalReq> rq[kMaxEvalLanes];
            std::vector<PatchRun*> pend[kMaxEvalLanes];
            bool inflight[kMaxEvalLanes] = {};
            auto do_inits = [&](int e) {         // batched re-initialisations after a stage
                while (true) {
                    std::vector<PatchRun*> inits;
                    for (int k = 0; k < nb; ++k)
                        if (laneOf(k) == e && runs[k]->wantShiftInit) inits.push_back(runs[k].get());
                    if (inits.empty()) return;
                    auto t0 = clk::now();
                    shift_inits(c, inits, e);
                    c->hostStats[2].ms += since(t0);
                    c->hostStats[2].launches++;
                    c->hostStats[2].units += (long long)inits.size();
                    for (auto* r : inits) {
                        r->wantShiftInit = false;
                        auto h = r->initH;
                        h.resume();
                    }
                }
            };
            auto launch_lane = [&](int e) {
                pend[e].clear();
                for (int k = 0; k < nb; ++k)
                    if (laneOf(k) == e && runs[k]->ev.pending) pend[e].push_back(runs[k].get());
                if (pend[e].empty()) return false;
                rq[e].assign(pend[e].size(), EvalReq{});
                for (size_t k = 0; k < pend[e].size(); ++k) {
                    EvalSlot& ev = pend[e][k]->ev;
                    rq[e][k].slot = pend[e][k]->p;
                    rq[e][k].alpha = ev.alphaEff;
                    rq[e][k].wantGrad = ev.wantGrad;
                    rq[e][k].gradMethod = ev.method;
                    rq[e][k].nEnabled = ev.nEnabled;
                    rq[e][k].g = ev.gAlpha.data();
                    rq[e][k].line = ev.line;
                    rq[e][k].lineS = ev.lineS;
                    rq[e][k].saveBase = ev.saveBase;
                    rq[e][k].dAlpha = ev.line == 1 ? ev.dAlpha.data() : nullptr;
                }
                auto t0 = clk::now();
                c->dev->evaluate_begin(e, rq[e]);
                c->hostStats[0].ms += since(t0);
                c->hostStats[0].launches++;
                c->hostStats[0].units += (long long)pend[e].size();
                inflight[e] = true;
                return true;
            };
            auto finish_lane = [&](int e) {
                auto t0 = clk::now();
                c->dev->evaluate_end(e, rq[e]);
                c->hostStats[0].ms += since(t0);
                for (size_t k = 0; k < pend[e].size(); ++k) {
                    pend[e][k]->ev.f = rq[e][k].f;
                    pend[e][k]->ev.pending = false;
                }
                t0 = clk::now();
                c->pool->run((int)pend[e].size(), [&](int k) { pend[e][k]->ev.h.resume(); });
                c->hostStats[1].ms += since(t0);
                c->hostStats[1].launches++;
                inflight[e] = false;
            };
            // Device line searches: new lines start (their states and directions) and every
            // lane with lines in flight runs lineRounds rounds behind its evaluation; the lines
            // that ended resume their patches when the lane's rounds are read back.
            bool lsInflight[kMaxEvalLanes] = {};
            std::vector<LsResult> lsOut;
            auto launch_lines = [&](int e) {
                if (!c->deviceLS) return false;
                std::vector<LsStart> starts;
                bool any = false;
                for (int k = 0; k < nb; ++k) {
                    if (laneOf(k) != e) continue;
                    auto& d = runs[k]->dl;
                    if (d.pending) {
                        starts.push_back({lsIdx[k], d.fa, d.lambda, runs[k]->ev.dAlpha.data()});
                        d.pending = false;
                        d.active = true;
                    }
                    any |= d.active;
                }
                if (!any) return false;
                auto t0 = clk::now();
                c->dev->ls_start(e, starts);
                c->dev->ls_run(e, c->prm.lineRounds);
                c->hostStats[6].ms += since(t0);
                c->hostStats[6].launches++;
                c->hostStats[6].units += (long long)starts.size();
                lsInflight[e] = true;
                return true;
            };
            auto finish_lines = [&](int e) {
                auto t0 = clk::now();
                c->dev->ls_end(e, lsOut);
                c->hostStats[6].ms += since(t0);
                std::vector<PatchRun*> ended;
                for (int k = 0; k < nb; ++k) {
                    auto& d = runs[k]->dl;
                    if (laneOf(k) != e || !d.active) continue;
                    const LsResult& o = lsOut[lsIdx[k]];
                    if (!o.done) continue;
                    d.active = false;
                    d.step = o.step;
                    d.f = o.f;
                    d.probes = o.probes;
                    ended.push_back(runs[k].get());
                }
                t0 = clk::now();
                c->pool->run((int)ended.size(), [&](int k) { ended[k]->dl.h.resume(); });
                c->hostStats[1].ms += since(t0);
                lsInflight[e] = false;
            };
            // Restoration (Object::restorePatch) of the patches that have finished, while the
            // others still iterate: the sums, the filter, the inverse transform and the crop
            // run on the device's synchronous lane beside the evaluation lanes, in groups of
            // RESTORE_MIN..RESTORE_MAX patches, and the objects come back as floats.  A
            // patch's restoration depends on its own final alpha only, so grouping leaves it
            // unchanged.
            constexpr int RESTORE_MIN = 16, RESTORE_MAX = 32;
            const size_t objPer = (size_t)c->g.nObj * c->g.s * c->g.s;
            struct Restore {
                std::vector<int> ks;
                std::vector<double> ae, nf;
            };
            std::unique_ptr<Restore> onDevice;
            std::vector<char> queued(nb, 0);
            std::vector<float> rsOut;
            const std::vector<double> rsC2 = restore_cutoffs(c);
            auto device_done = [&](bool wait) {
                if (!onDevice || (!wait && !c->dev->restore_ready())) return;
                auto t0 = clk::now();
                rsOut.resize(onDevice->ks.size() * objPer);
                c->dev->restore_objects_end(rsOut.data());
                for (size_t i = 0; i < onDevice->ks.size(); ++i)
                    std::copy(rsOut.begin() + i * objPer, rsOut.begin() + (i + 1) * objPer,
                              imgOut + (size_t)runs[onDevice->ks[i]]->p * objPer);
                onDevice.reset();
                c->hostStats[5].ms += since(t0);
            };
            auto poll_restores = [&](bool flush) {
                if (!imgOut) return;
                device_done(flush);
                if (onDevice) return;
                std::vector<int> ks;
                for (int k = 0; k < nb && (int)ks.size() < RESTORE_MAX; ++k)
                    if (!queued[k] && tasks[k].h.done()) ks.push_back(k);
                if (ks.empty() || (!flush && (int)ks.size() < RESTORE_MIN)) return;
                auto t0 = clk::now();
                auto rs = std::make_unique<Restore>();
                rs->ks = ks;
                rs->ae.resize(ks.size() * (size_t)nPar);
                std::vector<const double*> ap(ks.size());
                std::vector<int> slots(ks.size());
                for (size_t i = 0; i < ks.size(); ++i) {
                    queued[ks[i]] = 1;
                    slots[i] = runs[ks[i]]->p;
                    effective_alpha(c, slots[i], runs[ks[i]]->alpha.data(), &rs->ae[i * nPar]);
                    ap[i] = &rs->ae[i * nPar];
                }
                rs->nf = restore_noise(c, slots.data(), (int)slots.size());
                c->dev->restore_objects_begin((int)ks.size(), slots.data(), ap.data(), rs->nf.data(), rsC2.data(),
                                              !c->prm.noFilter, false);
                onDevice = std::move(rs);
                c->hostStats[5].ms += since(t0);
                c->hostStats[5].launches++;
            };
            while (true) {
                bool any = false;
                for (int e = 0; e < nLanes; ++e) {
                    if (inflight[e]) {
                        finish_lane(e);
                        any = true;
                    }
                    if (lsInflight[e]) {
                        finish_lines(e);
                        any = true;
                    }
                    do_inits(e);
                    any |= launch_lane(e);
                    any |= launch_lines(e);
                    poll_restores(false);
                }
                if (!any) break;
            }
            if (c->deviceLS) c->dev->ls_clear();
            for (size_t k = 0; k < tasks.size(); ++k) {
                if (!tasks[k].h.done()) throw std::runtime_error("patch state machine stalled");
                if (tasks[k].h.promise().exc) std::rethrow_exception(tasks[k].h.promise().exc);
            }
            if (imgOut) {
                auto t0 = clk::now();
                while (true) {
                    poll_restores(true);
                    if (!onDevice) break;
                }
                c->hostStats[4].ms += since(t0);
                c->hostStats[4].launches++;
                c->hostStats[4].units += nb;
            }
            // results (Solver::run tail, solver.cpp:549-566); the objects were restored above
            for (int k = 0; k < nb; ++k) {
                auto& r = *runs[k];
                r.res.nF = r.nF;
                r.res.nDF = r.nDF;
                if (results) results[r.p] = r.res;
                if (alphaOut) std::copy(r.alpha.begin(), r.alpha.end(), alphaOut + (size_t)r.p * nPar);
            }
        }
        c->initDone = false;    // shifts moved during the solve: its final state is kept
        c->solved = true;
    } catch (const std::exception& e) {
        return set_err(c, MFBD_EDEVICE, e.what());
    }
    return MFBD_OK;
}

int mfbd_solve_group(int n, mfbd_ctx* const* ctxs, const double* const* alpha0, mfbd_result* const* results,
                     double* const* alphaOut, float* const* imgOut) {
    if (n < 1 || !ctxs || !results || !alphaOut) return MFBD_EINVAL;
    for (int k = 0; k < n; ++k) {
        if (!ctxs[k] || !results[k] || !alphaOut[k]) return MFBD_EINVAL;
        for (int k2 = 0; k2 < k; ++k2)
            if (ctxs[k2] == ctxs[k]) return set_err(ctxs[k], MFBD_EINVAL, "a context appears twice in the group");
        const int rc = check_ready(ctxs[k]);
        if (rc) return rc;
    }
    // one host thread per context: each drives its own streams and batches, so the
    // contexts' launches overlap on the device
    std::vector<int> rcs(n, MFBD_OK);
    std::vector<std::thread> th;
    for (int k = 1; k < n; ++k)
        th.emplace_back([&, k] {
            try {
                ctxs[k]->dev->bind_thread();
            } catch (const std::exception& e) {
                rcs[k] = set_err(ctxs[k], MFBD_EDEVICE, e.what());
                return;
            }
            rcs[k] = mfbd_solve(ctxs[k], alpha0 ? alpha0[k] : nullptr, results[k], alphaOut[k], imgOut ? imgOut[k] : nullptr);
        });
    rcs[0] = mfbd_solve(ctxs[0], alpha0 ? alpha0[0] : nullptr, results[0], alphaOut[0], imgOut ? imgOut[0] : nullptr);
    for (auto& t : th) t.join();
    for (int rc : rcs)
        if (rc) return rc;
    return MFBD_OK;
}

int mfbd_debug_get(mfbd_ctx* c, const char* what, int p, double* out, int* count) {
    if (!c || !what || !count) return MFBD_EINVAL;
    try {
        std::vector<double> v;
        if (!strcmp(what, "gamma_host")) v = c->gamma.at(p);
        else if (!strcmp(what, "noise_host")) v = c->noise.at(p);
        else if (!strcmp(what, "otfnorm")) v = c->otfNorm;
        else if (!strcmp(what, "window")) v = c->window;
        else if (!strcmp(what, "shifts")) v.assign(c->shifts.at(p).begin(), c->shifts.at(p).end());
        else if (!strcmp(what, "sizes") && c->dev) v = {(double)c->dev->nPS(), (double)c->dev->nOS(), (double)c->dev->nOSStored()};
        else if (!strcmp(what, "tiltoff")) v = c->tiltOff.at(p);
        else if (c->dev) c->dev->debug_get(what, p, v);
        *count = (int)v.size();
        if (out) std::copy(v.begin(), v.end(), out);
    } catch (const std::exception& e) {
        return set_err(c, MFBD_EDEVICE, e.what());
    }
    return MFBD_OK;
}

int mfbd_kernel_stats(mfbd_ctx* c, int cap, char (*names)[48], double* ms, int* launches, long long* units) {
    if (!c || !c->dev) return 0;
    auto st = c->dev->kernel_stats();
    for (auto& h : c->hostStats)
        if (h.launches) st.push_back(h);
    int n = 0;
    for (auto& s : st) {
        if (n >= cap) break;
        std::snprintf(names[n], 48, "%s", s.name.c_str());
        ms[n] = s.ms;
        launches[n] = s.launches;
        if (units) units[n] = s.units;
        n++;
    }
    return n;
}

}  // extern "C"
