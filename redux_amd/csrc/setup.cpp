// Per-job setup behind the C-ABI: pupil, supports, Zernike/KL mode sets, mode
// normalisation, tilt Jacobian and the constraint nullspace.  These run once per
// geometry (the reference caches them in GlobalData), on the host.
//
// Reference functions restated (ISP-SST/redux):
//   Pupil::calculatePupilSize     src/lib/image/pupil.cpp:78-100
//   makePupil / Grid::init        src/lib/image/utils.cpp:68-121, grid.cpp:45-67
//   Pupil::normalize/generateSupport  pupil.cpp:245-299
//   Zernike::NollToNM/getZernike  src/lib/image/zernike.cpp:69-80, 151-274
//   Zernike::calcCovariance       zernike.cpp:102-134
//   karhunenLoeveExpansion        zernike.cpp:367-479, PupilMode(KL) modes.cpp:158-205
//   ModeSet::getNorms/normalize/measureJacobian  src/lib/momfbd/modes.cpp:443-535
//   Constraints::init + nullspace src/lib/momfbd/constraints.cpp:85-198,302-356,429-469,536-622
#include "../../include/momfbd_hip.h"
#include "hostfft.hpp"
#include "modeplan.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

using namespace mfbd;

namespace {

thread_local std::string g_setup_err;

int fail(int code, const std::string& msg) {
    g_setup_err = msg;
    return code;
}

// Grid::init with origin (n/2 + 0.5) as a float, distances in long double
void grid(int n, std::vector<long double>& dist, std::vector<long double>& ang) {
    float mid = n / 2.0f + 0.5f;
    long double o = (long double)mid - 0.5L;
    dist.assign((size_t)n * n, 0);
    ang.assign((size_t)n * n, 0);
    for (int y = 0; y < n; ++y) {
        long double yd = y - o;
        for (int x = 0; x < n; ++x) {
            long double xd = x - o;
            if (yd != 0 || xd != 0) {
                dist[(size_t)y * n + x] = sqrtl(yd * yd + xd * xd);
                ang[(size_t)y * n + x] = atan2l(yd, xd);
            }
        }
    }
}

void noll_to_nm(int j, int& n, int& m) {
    n = 0;
    int len = 1;
    for (int i = 1; len < j; ++i) len += (n = i) + 1;
    int dl = n + 1 - len + j;
    m = 2 * ((dl + (n % 2)) / 2) + !(n % 2) - 1;
    if (j % 2) m = -m;
}

// Zernike radial polynomial R_n^m(r) via the Kintner three-term recurrence in n
// (stable for high orders, unlike the alternating factorial sum).
long double radial(int n, int m, long double r) {
    long double r2 = r * r;
    long double rm = powl(r, m);
    if (n == m) return rm;
    long double Rm2 = rm;                                   // R_m^m
    long double R = (m + 2) * rm * r2 - (m + 1) * rm;       // R_{m+2}^m
    for (int k = m + 4; k <= n; k += 2) {
        long double K1 = (long double)(k + m) * (k - m) * (k - 2) / 2.0L;
        long double K2 = 2.0L * k * (k - 1) * (k - 2);
        long double K3 = -(long double)m * m * (k - 1) - (long double)k * (k - 1) * (k - 2);
        long double K4 = -(long double)k * (k + m - 2) * (k - m - 2) / 2.0L;
        long double Rn = ((K2 * r2 + K3) * R + K4 * Rm2) / K1;
        Rm2 = R;
        R = Rn;
    }
    return R;
}

struct PupilData {
    std::vector<double> pupil;
    double area = 0;
};

PupilData make_pupil(int P, double radius, double inner) {
    PupilData pd;
    pd.pupil.assign((size_t)P * P, 0.0);
    std::vector<long double> D, A;
    grid(P, D, A);
    long double area = 0;
    bool hasInner = inner > 0;
    for (int x = 0; x < P; ++x)
        for (int y = 0; y < P; ++y) {
            size_t k = (size_t)y * P + x;
            long double Dd = D[k];
            long double aa = fabsl(A[k]);
            long double d = (aa > M_PI / 4 && aa < 3 * M_PI / 4) ? fabsl(1.0L / sinl(aa)) : fabsl(1.0L / cosl(aa));
            long double dir = Dd - inner;
            if (hasInner && dir < -d) continue;
            long double dor = Dd - radius;
            if (dor > d) continue;
            if (dor < -d && (!hasInner || dir > d)) {
                pd.pupil[k] = 1;
                area += 1;
                continue;
            }
            long double R = dor;
            d *= 2;
            if (hasInner && dor < -d) { R = dir; d = -d; }
            long double val = 0.5L - R / d;
            pd.pupil[k] = (double)val;
            area += val;
        }
    // Pupil::normalize: rescale to [0,1]
    double mn = *std::min_element(pd.pupil.begin(), pd.pupil.end());
    double mx = *std::max_element(pd.pupil.begin(), pd.pupil.end());
    if (mn != mx)
        for (auto& v : pd.pupil) v = (v - mn) * (1.0 / (mx - mn));
    pd.area = (double)area;
    return pd;
}

struct Support {
    std::vector<int> ps, pio, os;
    double area = 0;
};

Support pupil_support(int P, const double* pupil) {
    Support s;
    int N = 2 * P;
    std::vector<cd> otf((size_t)N * N, 0.0);
    for (int y = 0; y < P; ++y)
        for (int x = 0; x < P; ++x) {
            int cnt = y * P + x;
            double v = pupil[cnt];
            otf[(size_t)y * N + x] = v;
            if (v > 1e-9) {
                s.ps.push_back(cnt);
                s.pio.push_back(y * N + x + P / 2 + (P / 2) * N);
                s.area += v;
            }
        }
    fft2d(otf.data(), N, N, -1);
    for (auto& c : otf) c = std::norm(c);
    fft2d(otf.data(), N, N, +1);
    std::vector<cd> c2((size_t)N * N);
    reorder_into(otf.data(), N, N, c2.data());
    for (int i = 0; i < N * N; ++i)
        if (std::fabs(c2[i].real() / double(N * N)) > 1e-9) s.os.push_back(i);
    return s;
}

std::vector<double> zernike(int j, int P, float radius, float angle, bool normalize) {
    int n, m;
    noll_to_nm(j, n, m);
    std::vector<long double> D, A;
    grid(P, D, A);
    size_t P2 = (size_t)P * P;
    std::vector<long double> z(P2);
    long double rinv = 1.0L / radius;
    long double angr = (long double)angle * M_PI / 180.0L;
    int am = std::abs(m);
    for (size_t k = 0; k < P2; ++k) {
        long double r = (long double)std::min<double>(1.0, (double)(rinv * D[k]));
        long double rad = (long double)(double)radial(n, am, r);
        long double ang = (m < 0) ? sinl(am * (A[k] - angr)) : cosl(am * (A[k] - angr));
        z[k] = rad * (long double)(double)ang;
    }
    long double nrm;
    if (normalize) {
        PupilData pd = make_pupil(P, radius, 0.0);
        Support sp = pupil_support(P, pd.pupil.data());
        long double acc = 0;
        for (size_t k = 0; k < P2; ++k) acc += z[k] * z[k] * pd.pupil[k];
        nrm = sqrtl((long double)sp.area / acc);
    } else {
        nrm = m ? sqrtl(2.0L * (n + 1)) : sqrtl((long double)(n + 1));
    }
    std::vector<double> out(P2);
    for (size_t k = 0; k < P2; ++k) out[k] = (double)(z[k] * nrm);
    return out;
}

double lgam_sign(double x, double* sgn) {
    int s = 1;
    double v = lgamma_r(x, &s);
    *sgn = s;
    return v;
}

double cov_nm(int n1, int m1, int n2, int m2) {
    if (m1 != m2) return 0.0;
    double s1, s2, s3, s4;
    double k = std::pow(4.8 * std::exp(lgam_sign(6.0 / 5.0, &s1)), 5.0 / 6.0) *
               std::exp(lgam_sign(14.0 / 3.0, &s2) + 2.0 * lgam_sign(11.0 / 6.0, &s3)) /
               (std::pow(2.0, 8.0 / 3.0) * M_PI);
    k *= std::pow(-1.0, (double)((n1 + n2 - 2 * m1) / 2.0)) * std::sqrt((double)((n1 + 1) * (n2 + 1)));
    double g1 = lgam_sign(((double)(n1 + n2) - 5.0 / 3.0) / 2.0, &s1);
    double g2 = lgam_sign(((double)(n1 - n2) + 17.0 / 3.0) / 2.0, &s2);
    double g3 = lgam_sign(((double)(n2 - n1) + 17.0 / 3.0) / 2.0, &s3);
    double g4 = lgam_sign(((double)(n1 + n2) + 23.0 / 3.0) / 2.0, &s4);
    return k * std::exp(g1 - g2 - g3 - g4) * s1 * s2 * s3 * s4;
}

double noll_cov(int i, int j) {
    if (i < 2 || j < 2) return 0.0;
    if (i > j) std::swap(i, j);
    int in, im, jn, jm;
    noll_to_nm(i, in, im);
    if (im && ((i + j) % 2)) return 0.0;
    noll_to_nm(j, jn, jm);
    return cov_nm(in, std::abs(im), jn, std::abs(jm));
}

// cyclic Jacobi eigen-decomposition of a symmetric matrix (row-major n x n);
// returns eigenvalues descending with eigenvectors in the columns of V
void jacobi_eigen(std::vector<double> A, int n, std::vector<double>& w, std::vector<double>& V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += A[(size_t)p * n + q] * A[(size_t)p * n + q];
        if (off < 1e-60) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = A[(size_t)p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - s * vkq;
                    V[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return A[(size_t)a * n + a] > A[(size_t)b * n + b]; });
    w.resize(n);
    std::vector<double> V2((size_t)n * n);
    for (int c = 0; c < n; ++c) {
        w[c] = A[(size_t)order[c] * n + order[c]];
        for (int r = 0; r < n; ++r) V2[(size_t)r * n + c] = V[(size_t)r * n + order[c]];
    }
    V = V2;
}

struct KL {
    double covariance;
    std::vector<std::pair<int, double>> weights;
};

std::map<int, KL> kl_expansion(int first, int last) {
    std::map<int, int> mapping, reverse;
    for (int i = first; i <= last; ++i) mapping[i] = i;
    for (int i = first; i <= last - 1; ++i) {
        double cov1 = noll_cov(mapping[i], mapping[last]);
        for (int j = last - 1; j > i; --j) {
            double cov2 = noll_cov(mapping[i], mapping[j]);
            if (!cov2 && cov1) std::swap(mapping[j], mapping[j + 1]);
            else cov1 = cov2;
        }
    }
    for (auto& e : mapping) reverse[e.second] = e.first;
    std::map<int, double> values;
    int nBlocks = 1;
    for (int k = first + 1; k <= last; ++k) {
        values[k] = noll_cov(mapping[k - 1], mapping[k]);
        if (values[k] == 0) ++nBlocks;
    }
    struct Block { int fb, lb; std::vector<double> w, U; };
    std::vector<Block> blocks;
    int n = first;
    for (int b = 0; b < nBlocks; ++b) {
        Block bl;
        bl.fb = n++;
        while (n <= last && values.at(n)) ++n;
        bl.lb = n - 1;
        int sz = bl.lb - bl.fb + 1;
        std::vector<double> M((size_t)sz * sz);
        for (int a = 0; a < sz; ++a)
            for (int c = 0; c < sz; ++c) M[(size_t)a * sz + c] = noll_cov(mapping[bl.fb + a], mapping[bl.fb + c]);
        if (sz > 1) jacobi_eigen(M, sz, bl.w, bl.U);
        else { bl.w = {M[0]}; bl.U = {1.0}; }
        blocks.push_back(std::move(bl));
    }
    std::map<int, KL> out;
    for (int i = first; i <= last; ++i) {
        int im = reverse.at(i);
        const Block* bp = nullptr;
        for (auto& b : blocks) if (b.fb <= im && im <= b.lb) { bp = &b; break; }
        int sz = bp->lb - bp->fb + 1, col = im - bp->fb;
        KL kl;
        kl.covariance = bp->w[col];
        int sign = 1;
        double largest = 0;
        for (int m = 0; m < sz; ++m) {
            double c = bp->U[(size_t)m * sz + col];
            if (std::fabs(c) > largest) { largest = std::fabs(c); sign = c < 0 ? -1 : 1; }
        }
        for (int m = 0; m < sz; ++m) kl.weights.push_back({mapping.at(m + bp->fb), sign * bp->U[(size_t)m * sz + col]});
        out[i] = kl;
    }
    return out;
}

}  // namespace

namespace mfbd {

void noll_nm(int j, int& n, int& m) { noll_to_nm(j, n, m); }

std::vector<double> mode_pupil(int P, double radius, double& area) {
    PupilData pd = make_pupil(P, radius, 0.0);
    area = pupil_support(P, pd.pupil.data()).area;
    return pd.pupil;
}

// PupilMode selection (modes.cpp:158-205): tilts (Noll 2, 3) and the Zernike basis are
// Zernike polynomials, everything else in the KL basis a KL mode of [klMin, klMax]
int mode_plan(int nModes, const int* modeIds, int basis, int klMin, int klMax, double klCutoff, ModePlan& plan,
              std::string& err) {
    std::map<int, KL> kl;
    if (basis == MFBD_BASIS_KL) kl = kl_expansion(klMin, klMax);
    plan = ModePlan{};
    for (int k = 0; k < nModes; ++k) {
        const int j = modeIds[k];
        plan.noll.push_back(j);
        plan.terms.emplace_back();
        if (j == 2 || j == 3 || basis == MFBD_BASIS_ZERNIKE) {
            if (j == 2) plan.tilt[0] = k;
            if (j == 3) plan.tilt[1] = k;
            plan.kind.push_back(j < 2 ? ModePlan::CONSTANT : ModePlan::ZERNIKE);
            plan.atmRms.push_back(std::sqrt(noll_cov(j, j)));
        } else {
            auto it = kl.find(j);
            if (it == kl.end()) {
                err = "mfbd_make_modes: KL index outside [klMin, klMax]";
                return MFBD_EINVAL;
            }
            plan.kind.push_back(ModePlan::KL);
            for (auto& w : it->second.weights)
                if (std::fabs(w.second) >= klCutoff) plan.terms.back().push_back(w);
            plan.atmRms.push_back(std::sqrt(it->second.covariance));
        }
    }
    return MFBD_OK;
}

}  // namespace mfbd

extern "C" {

const char* mfbd_setup_last_error(void) { return g_setup_err.c_str(); }

int mfbd_pupil_size(double wavelength, int nPixels, double telescopeD, double arcSecsPerPixel,
                    double* freqCutoff, double* radius, int* pupilPixels) {
    static const int goodsizes[] = {16, 18, 20, 24, 25, 27, 30, 32, 36, 40, 45, 48, 50, 54, 60,
                                    64, 72, 75, 80, 81, 90, 96, 100, 108, 120, 125, 128, 135, 144};
    if (nPixels < 4 || wavelength <= 0 || telescopeD <= 0 || arcSecsPerPixel <= 0)
        return fail(MFBD_EINVAL, "mfbd_pupil_size: non-positive optics or nPixels < 4");
    double rpp = arcSecsPerPixel * (M_PI / (180.0 * 3600.0));
    double q = wavelength / (rpp * telescopeD);
    double fc = (double)nPixels / q;
    int npp = nPixels >> 2;
    double r = fc / 2.0;
    if (npp < r)
        for (int gs : goodsizes) {
            npp = std::max(gs, npp);
            if (npp >= r) break;
        }
    *freqCutoff = fc;
    *radius = r;
    *pupilPixels = npp << 1;
    return MFBD_OK;
}

int mfbd_make_pupil(int P, double radius, double coRadius, double* pupil, double* area) {
    if (P < 2 || !pupil) return fail(MFBD_EINVAL, "mfbd_make_pupil: bad size");
    PupilData pd = make_pupil(P, radius, coRadius);
    std::memcpy(pupil, pd.pupil.data(), sizeof(double) * P * P);
    if (area) *area = pd.area;
    return MFBD_OK;
}

int mfbd_pupil_support(int P, const double* pupil, int* nPS, int* ps, int* pio, int* nOS, int* os, double* area) {
    if (P < 2 || !pupil) return fail(MFBD_EINVAL, "mfbd_pupil_support: bad size");
    Support s = pupil_support(P, pupil);
    *nPS = (int)s.ps.size();
    *nOS = (int)s.os.size();
    if (ps) std::memcpy(ps, s.ps.data(), sizeof(int) * s.ps.size());
    if (pio) std::memcpy(pio, s.pio.data(), sizeof(int) * s.pio.size());
    if (os) std::memcpy(os, s.os.data(), sizeof(int) * s.os.size());
    if (area) *area = s.area;
    return MFBD_OK;
}

int mfbd_noll_covariance(int i, int j, double* cov) {
    *cov = noll_cov(i, j);
    return MFBD_OK;
}

int mfbd_make_modes(int P, double radius, double angle, int nModes, const int* modeIds, int basis,
                    int klMin, int klMax, double klCutoff, double* modes, double* atmRms, int* tiltXY) {
    if (P < 2 || nModes < 1 || !modeIds || !modes) return fail(MFBD_EINVAL, "mfbd_make_modes: bad arguments");
    size_t P2 = (size_t)P * P;
    ModePlan plan;
    std::string err;
    if (int rc = mode_plan(nModes, modeIds, basis, klMin, klMax, klCutoff, plan, err)) return fail(rc, err);
    std::map<int, std::vector<double>> zcache;
    auto Z = [&](int j) -> const std::vector<double>& {
        auto it = zcache.find(j);
        if (it == zcache.end()) it = zcache.emplace(j, zernike(j, P, (float)radius, (float)angle, true)).first;
        return it->second;
    };
    PupilData pd = make_pupil(P, radius, 0.0);
    Support sp = pupil_support(P, pd.pupil.data());
    tiltXY[0] = plan.tilt[0];
    tiltXY[1] = plan.tilt[1];
    for (int k = 0; k < nModes; ++k) {
        int j = plan.noll[k];
        double* out = modes + (size_t)k * P2;
        if (plan.kind[k] == ModePlan::CONSTANT) {
            for (size_t q = 0; q < P2; ++q) out[q] = j;
        } else if (plan.kind[k] == ModePlan::ZERNIKE) {
            std::memcpy(out, Z(j).data(), sizeof(double) * P2);
        } else {
            std::vector<double> m(P2, 0.0);
            for (auto& w : plan.terms[k]) {
                const auto& z = Z(w.first);
                for (size_t q = 0; q < P2; ++q) m[q] = m[q] + z[q] * w.second;
            }
            long double nrm = 0;
            for (size_t q = 0; q < P2; ++q)
                if (pd.pupil[q] > 0) nrm += (long double)m[q] * m[q] * pd.pupil[q];
            nrm = sqrtl((long double)sp.area / nrm);
            for (size_t q = 0; q < P2; ++q) out[q] = (double)(m[q] * nrm);
        }
        if (atmRms) atmRms[k] = plan.atmRms[k];
    }
    return MFBD_OK;
}

int mfbd_normalize_modes(int P, int K, double* modes, const double* pupil, double area, double scale) {
    size_t P2 = (size_t)P * P;
    for (int k = 0; k < K; ++k) {
        double* m = modes + (size_t)k * P2;
        double nrm = 0;
        for (size_t q = 0; q < P2; ++q)
            if (pupil[q] > 0) { double t = m[q] * pupil[q]; nrm += t * m[q]; }
        double norm = (double)sqrtl((long double)nrm / area);
        double s = scale / norm;
        for (size_t q = 0; q < P2; ++q) m[q] *= s;
    }
    return MFBD_OK;
}

int mfbd_measure_jacobian(int P, int K, const double* modes, const double* pupil, double scale,
                          int* tiltXY, double* J, double* Ji) {
    // fitPlane over pupil > 0: least squares z = a*y + b*x + c via 3x3 normal equations
    std::vector<int> idx;
    for (int q = 0; q < P * P; ++q) if (pupil[q] > 0) idx.push_back(q);
    struct Fit { double chi, a, b, c; int i; };
    std::vector<Fit> fits;
    double half = P / 2;
    for (int k = 0; k < K; ++k) {
        const double* m = modes + (size_t)k * P * P;
        double A[3][3] = {{0}}, r[3] = {0};
        for (int q : idx) {
            double v[3] = {double(q / P) - half, double(q % P) - half, 1.0};
            for (int a = 0; a < 3; ++a) {
                r[a] += v[a] * m[q];
                for (int b = 0; b < 3; ++b) A[a][b] += v[a] * v[b];
            }
        }
        // solve A x = r (Gaussian elimination with partial pivoting)
        double M[3][4];
        for (int a = 0; a < 3; ++a) { for (int b = 0; b < 3; ++b) M[a][b] = A[a][b]; M[a][3] = r[a]; }
        for (int c = 0; c < 3; ++c) {
            int p = c;
            for (int a = c + 1; a < 3; ++a) if (std::fabs(M[a][c]) > std::fabs(M[p][c])) p = a;
            for (int b = 0; b < 4; ++b) std::swap(M[c][b], M[p][b]);
            for (int a = c + 1; a < 3; ++a) {
                double f = M[a][c] / M[c][c];
                for (int b = c; b < 4; ++b) M[a][b] -= f * M[c][b];
            }
        }
        double x[3];
        for (int c = 2; c >= 0; --c) {
            double s = M[c][3];
            for (int b = c + 1; b < 3; ++b) s -= M[c][b] * x[b];
            x[c] = s / M[c][c];
        }
        double chi = 0;
        for (int q : idx) {
            double e = m[q] - (x[0] * (double(q / P) - half) + x[1] * (double(q % P) - half) + x[2]);
            chi += e * e;
        }
        fits.push_back({chi / idx.size(), x[0], x[1], x[2], k});
    }
    std::stable_sort(fits.begin(), fits.end(), [](const Fit& u, const Fit& v) {
        if (u.chi != v.chi) return u.chi < v.chi;
        if (u.a != v.a) return u.a < v.a;
        if (u.b != v.b) return u.b < v.b;
        return u.c < v.c;
    });
    double Jm[2][2] = {{0, 0}, {0, 0}};
    double chisq[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
    int t[2] = {-1, -1};
    int cnt = 0;
    for (auto& f : fits) {
        chisq[cnt] = f.chi;
        if (cnt == 2) break;
        Jm[0][cnt] = scale * f.b;
        Jm[1][cnt] = scale * f.a;
        t[cnt] = f.i;
        cnt++;
    }
    if (100 * chisq[0] < chisq[2] && 100 * chisq[1] < chisq[2]) {
        if (std::fabs(Jm[0][0]) / std::fabs(Jm[1][0]) < 1) {
            std::swap(t[0], t[1]);
            std::swap(Jm[0][0], Jm[0][1]);
            std::swap(Jm[1][0], Jm[1][1]);
        }
    }
    double invdet = 1 / (Jm[0][0] * Jm[1][1] - Jm[0][1] * Jm[1][0]);
    tiltXY[0] = t[0];
    tiltXY[1] = t[1];
    J[0] = Jm[0][0]; J[1] = Jm[0][1]; J[2] = Jm[1][0]; J[3] = Jm[1][1];
    Ji[0] = Jm[1][1] * invdet; Ji[1] = -Jm[1][0] * invdet; Ji[2] = -Jm[0][1] * invdet; Ji[3] = Jm[0][0] * invdet;
    return MFBD_OK;
}

// Constraints::init + groupConnectedVariables + blockify + calculateNullspace.
// Images are ordered object -> channel -> frame; wavefronts[img] is the image's
// wavefront id.  Output entries sorted by (row, col); capacity bounds the arrays.
int mfbd_make_constraints(int nObj, const int* nChPerObj, const int* nFramesPerCh, const int* wavefronts,
                          int K, int* nPar, int* nFree, int* nnz, int* rows, int* cols, double* vals,
                          int capacity) {
    int nImg = 0, nChTot = 0;
    for (int o = 0; o < nObj; ++o)
        for (int c = 0; c < nChPerObj[o]; ++c) nImg += nFramesPerCh[nChTot++];
    const int np = nImg * K;
    using Con = std::map<int, int>;
    std::vector<Con> cons;
    for (int mi = 0; mi < K; ++mi) {
        if (mi == 0 || mi == 1) {
            Con sum;
            for (int ic = 0; ic < nFramesPerCh[0]; ++ic) sum[ic * K + mi] = 1;
            cons.push_back(sum);
        }
        std::map<int, Con> seen;
        int img = 0;
        for (int o = 0, ch = 0; o < nObj; ++o)
            for (int c = 0; c < nChPerObj[o]; ++c, ++ch)
                for (int f = 0; f < nFramesPerCh[ch]; ++f, ++img) {
                    int wf = wavefronts[img];
                    int off = img * K + mi;
                    auto it = seen.find(wf);
                    if (it != seen.end()) {
                        Con c2 = it->second;
                        c2[off] = -1;
                        cons.push_back(c2);
                    } else {
                        seen[wf] = Con{{off, 1}};
                    }
                }
    }
    const int nf = np - (int)cons.size();
    // connected groups, in order of first constraint
    std::vector<char> used(cons.size(), 0);
    std::map<std::pair<int, int>, double> entries;
    std::set<int> usedRows, usedCols;
    int cOff = 0;
    for (size_t c0 = 0; c0 < cons.size(); ++c0) {
        if (used[c0]) continue;
        std::vector<int> g{(int)c0};
        used[c0] = 1;
        std::set<int> idx;
        for (auto& e : cons[c0]) idx.insert(e.first);
        bool changed = true;
        while (changed) {
            changed = false;
            for (size_t c = 0; c < cons.size(); ++c) {
                if (used[c]) continue;
                bool conn = false;
                for (auto& e : cons[c]) if (idx.count(e.first)) { conn = true; break; }
                if (conn) {
                    used[c] = 1;
                    g.push_back((int)c);
                    for (auto& e : cons[c]) idx.insert(e.first);
                    changed = true;
                    break;
                }
            }
        }
        // columns ordered by (entry count desc, index asc)
        std::vector<std::pair<int, int>> cnt;
        for (int i : idx) {
            int n = 0;
            for (int c : g) n += cons[c].count(i) ? 1 : 0;
            cnt.push_back({-n, i});
        }
        std::sort(cnt.begin(), cnt.end());
        std::vector<int> order;
        std::map<int, int> local;
        for (auto& e : cnt) { local[e.second] = (int)order.size(); order.push_back(e.second); }
        const int nP = (int)order.size(), nC = (int)g.size(), nF = nP - nC;
        // row order: the constraint with the most entries is the one the basis is built on
        std::vector<std::vector<int>> C(nC, std::vector<int>(nP, 0));
        for (int r = 0; r < nC; ++r)
            for (auto& e : cons[g[r]]) C[r][local[e.first]] = e.second;
        std::vector<std::vector<double>> ns(std::max(nF, 0), std::vector<double>(nP, 0.0));
        if (nF == 1) {
            for (int p = 0; p < nP; ++p) ns[0][p] = 1.0 / std::sqrt((double)nP);
        } else if (nF > 1) {
            int rowIndex = 0, maxc = 0;
            for (int r = 0; r < nC; ++r) {
                int n = 0;
                for (int p = 0; p < nP; ++p) n += C[r][p] != 0;
                if (n > maxc) { maxc = n; rowIndex = r; }
            }
            int bl = 1, basis = 0;
            while (bl < maxc) {
                int off = 0;
                while (off + bl < maxc) {
                    int rem = std::max(std::min(maxc - off - bl, bl), 0);
                    for (int i = 0; i < bl; ++i) {
                        ns[basis][off + i] = 1.0;
                        if (off + bl + i < maxc) ns[basis][off + bl + i] = -bl * 1.0 / rem;
                    }
                    basis++;
                    off += 2 * bl;
                }
                bl *= 2;
            }
            for (int i = 0; i < nF; ++i)
                for (int j = 0; j < maxc; ++j) {
                    double v = ns[i][j];
                    if (std::fabs(v) > 0)
                        for (int k = 0; k < nC; ++k) {
                            int v1 = C[k][j];
                            if (k != rowIndex && v1)
                                for (int p = 0; p < nP; ++p) {
                                    if (p == j) continue;
                                    int v2 = C[k][p];
                                    if (v2) ns[i][p] = -v1 * v / v2;
                                }
                        }
                }
            for (int i = 0; i < nF; ++i) {
                double sum = 0;
                for (int k = 0; k < nP; ++k) sum += ns[i][k] * ns[i][k];
                for (int j = i + 1; j < nF; ++j) {
                    double sp = 0;
                    for (int k = 0; k < nP; ++k) sp += ns[i][k] * ns[j][k];
                    if (std::fabs(sp) > 0) {
                        sp /= sum;
                        for (int k = 0; k < nP; ++k) ns[j][k] = ns[j][k] - sp * ns[i][k];
                    }
                }
                if (sum > 0 && sum != 1.0) {
                    double s = 1.0 / std::sqrt(sum);
                    for (int k = 0; k < nP; ++k) ns[i][k] = ns[i][k] * s;
                }
            }
        }
        for (int i = 0; i < nF; ++i)
            for (int p = 0; p < nP; ++p)
                if (std::fabs(ns[i][p]) > 1e-10) {
                    entries[{order[p], cOff + i}] = ns[i][p];
                    usedRows.insert(order[p]);
                    usedCols.insert(cOff + i);
                }
        cOff += std::max(nF, 0);
    }
    std::vector<int> u1, u2;
    for (int i = 0; i < np; ++i) if (!usedRows.count(i)) u1.push_back(i);
    for (int i = 0; i < nf; ++i) if (!usedCols.count(i)) u2.push_back(i);
    if (u1.size() != u2.size()) return fail(MFBD_EINVAL, "mfbd_make_constraints: inconsistent constraint set");
    for (size_t k = 0; k < u1.size(); ++k) entries[{u1[k], u2[k]}] = 1.0;
    *nPar = np;
    *nFree = nf;
    *nnz = (int)entries.size();
    if ((int)entries.size() > capacity) return fail(MFBD_ENOMEM, "mfbd_make_constraints: capacity too small");
    int k = 0;
    for (auto& e : entries) {
        rows[k] = e.first.first;
        cols[k] = e.first.second;
        vals[k] = e.second;
        ++k;
    }
    return MFBD_OK;
}

}  // extern "C"
