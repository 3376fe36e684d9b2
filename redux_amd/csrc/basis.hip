// Mode basis on the device: Zernike polynomials over the P x P pupil grid, their
// normalisation over the pupil, and the Karhunen-Loeve modes as weighted sums of them
// (Zernike::getZernike, src/lib/image/zernike.cpp:151-274; karhunenLoeveExpansion,
// zernike.cpp:367-479; PupilMode, src/lib/momfbd/modes.cpp:158-205).  The weights (the
// eigenvectors of the Noll covariance blocks) come from the host's mode plan; the P^2 x
// (number of Zernikes) polynomial evaluations, which grow with KL_MAX_MODE (2000 by
// default), run here in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/momfbd_hip.h"
#include "modeplan.hpp"

namespace {

#define BCHK(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e_)); \
    } while (0)

// z[i][p] for Noll index (n[i], m[i]) at pixel p of the P x P grid centred on (P/2, P/2)
// (Grid::init): the radial polynomial R_n^|m| at r = min(1, dist / radius) by the Kintner
// recurrence in n, times cos(|m| (theta - angle)) (m >= 0) or sin (m < 0)
__global__ void k_zernike(double* __restrict__ z, const int* __restrict__ nn, const int* __restrict__ mm, int nZ, int P,
                          double origin, double rinv, double angr) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (p >= P * P || i >= nZ) return;
    const double yd = p / P - origin, xd = p % P - origin;
    const bool centre = yd == 0 && xd == 0;
    const double dist = centre ? 0.0 : sqrt(yd * yd + xd * xd);
    const double theta = centre ? 0.0 : atan2(yd, xd);
    const double r = fmin(1.0, rinv * dist), r2 = r * r;
    const int n = nn[i], m = mm[i], am = abs(m);
    double rm = 1.0;
    for (int k = 0; k < am; ++k) rm *= r;
    double R = rm;
    if (n > am) {
        double Rm2 = rm;
        R = (am + 2) * rm * r2 - (am + 1) * rm;
        for (int k = am + 4; k <= n; k += 2) {
            const double K1 = (double)(k + am) * (k - am) * (k - 2) / 2.0;
            const double K2 = 2.0 * k * (k - 1) * (k - 2);
            const double K3 = -(double)am * am * (k - 1) - (double)k * (k - 1) * (k - 2);
            const double K4 = -(double)k * (k + am - 2) * (k - am - 2) / 2.0;
            const double Rn = ((K2 * r2 + K3) * R + K4 * Rm2) / K1;
            Rm2 = R;
            R = Rn;
        }
    }
    const double ang = m < 0 ? sin(am * (theta - angr)) : cos(am * (theta - angr));
    z[(size_t)i * P * P + p] = R * ang;
}

// deterministic block sum, blockDim = 256
__device__ double bsum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// row i scaled by sqrt(area / sum_p row^2 pupil) (the RMS over the pupil made one)
__global__ void __launch_bounds__(256) k_normalise(double* __restrict__ rows, const double* __restrict__ pupil, int P2,
                                                   double area) {
    __shared__ double sh[256];
    double* row = rows + (size_t)blockIdx.x * P2;
    double acc = 0;
    for (int p = threadIdx.x; p < P2; p += 256) acc += row[p] * row[p] * pupil[p];
    const double s = sqrt(area / bsum(acc, sh));
    for (int p = threadIdx.x; p < P2; p += 256) row[p] *= s;
}

// out[k][p] = sum_t w[t] z[zi[t]][p] over mode k's terms [first[k], first[k+1])
__global__ void k_combine(double* __restrict__ out, const double* __restrict__ z, const int* __restrict__ first,
                          const int* __restrict__ zi, const double* __restrict__ w, int P2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (p >= P2) return;
    double m = 0;
    for (int t = first[k]; t < first[k + 1]; ++t) m = m + z[(size_t)zi[t] * P2 + p] * w[t];
    out[(size_t)k * P2 + p] = m;
}

template <typename T>
struct Dev {
    T* p = nullptr;
    explicit Dev(size_t n) { BCHK(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T))); }
    ~Dev() { (void)hipFree(p); }
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
    void put(const std::vector<T>& h) { BCHK(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice)); }
};

thread_local std::string g_basis_err;

}  // namespace

extern "C" const char* mfbd_basis_last_error(void) { return g_basis_err.c_str(); }

extern "C" int mfbd_make_modes_device(int P, double radius, double angle, int nModes, const int* modeIds, int basis,
                                      int klMin, int klMax, double klCutoff, int device, double* modes,
                                      double* atmRms, int* tiltXY) {
    if (P < 2 || nModes < 1 || !modeIds || !modes || !tiltXY || radius <= 0) {
        g_basis_err = "mfbd_make_modes_device: bad arguments";
        return MFBD_EINVAL;
    }
    mfbd::ModePlan plan;
    if (int rc = mfbd::mode_plan(nModes, modeIds, basis, klMin, klMax, klCutoff, plan, g_basis_err)) return rc;
    try {
        BCHK(hipSetDevice(device));
        const int P2 = P * P;
        double area = 0;
        const std::vector<double> pupil = mfbd::mode_pupil(P, radius, area);
        // every Zernike the modes need, once
        std::map<int, int> zIndex;
        for (int k = 0; k < nModes; ++k) {
            if (plan.kind[k] == mfbd::ModePlan::ZERNIKE) zIndex.emplace(plan.noll[k], 0);
            for (auto& t : plan.terms[k]) zIndex.emplace(t.first, 0);
        }
        std::vector<int> nn, mm;
        for (auto& e : zIndex) {
            e.second = (int)nn.size();
            int n, m;
            mfbd::noll_nm(e.first, n, m);
            nn.push_back(n);
            mm.push_back(m);
        }
        const int nZ = (int)nn.size();
        if ((size_t)nZ * P2 * sizeof(double) > ((size_t)64 << 30)) throw std::runtime_error("basis too large");
        Dev<double> dZ((size_t)nZ * P2), dPup(P2), dOut((size_t)nModes * P2);
        Dev<int> dN(nZ), dM(nZ);
        dPup.put(pupil);
        dN.put(nn);
        dM.put(mm);
        const float mid = P / 2.0f + 0.5f;                  // Grid::init's origin, as a float
        const double origin = (double)mid - 0.5;
        const double rinv = 1.0 / (double)(float)radius, angr = (double)(float)angle * M_PI / 180.0;
        if (nZ) {
            hipLaunchKernelGGL(k_zernike, dim3((P2 + 255) / 256, nZ), dim3(256), 0, 0, dZ.p, dN.p, dM.p, nZ, P, origin,
                               rinv, angr);
            BCHK(hipGetLastError());
            hipLaunchKernelGGL(k_normalise, dim3(nZ), dim3(256), 0, 0, dZ.p, dPup.p, P2, area);
            BCHK(hipGetLastError());
        }
        // KL modes: weighted sums of the normalised Zernikes, then renormalised
        std::vector<int> first(1, 0), zi, klRows;
        std::vector<double> w;
        for (int k = 0; k < nModes; ++k) {
            if (plan.kind[k] != mfbd::ModePlan::KL) continue;
            klRows.push_back(k);
            for (auto& t : plan.terms[k]) {
                zi.push_back(zIndex.at(t.first));
                w.push_back(t.second);
            }
            first.push_back((int)zi.size());
        }
        std::vector<double> host((size_t)nModes * P2);
        if (!klRows.empty()) {
            const int nK = (int)klRows.size();
            Dev<int> dFirst(first.size()), dZi(zi.size());
            Dev<double> dW(w.size()), dKL((size_t)nK * P2);
            dFirst.put(first);
            if (!zi.empty()) {
                dZi.put(zi);
                dW.put(w);
            }
            hipLaunchKernelGGL(k_combine, dim3((P2 + 255) / 256, nK), dim3(256), 0, 0, dKL.p, dZ.p, dFirst.p, dZi.p, dW.p,
                               P2);
            BCHK(hipGetLastError());
            hipLaunchKernelGGL(k_normalise, dim3(nK), dim3(256), 0, 0, dKL.p, dPup.p, P2, area);
            BCHK(hipGetLastError());
            std::vector<double> kl((size_t)nK * P2);
            BCHK(hipMemcpy(kl.data(), dKL.p, kl.size() * sizeof(double), hipMemcpyDeviceToHost));
            for (int r = 0; r < nK; ++r) std::copy(kl.begin() + (size_t)r * P2, kl.begin() + (size_t)(r + 1) * P2,
                                                   host.begin() + (size_t)klRows[r] * P2);
        }
        std::vector<double> zs;
        if (nZ) {
            zs.resize((size_t)nZ * P2);
            BCHK(hipMemcpy(zs.data(), dZ.p, zs.size() * sizeof(double), hipMemcpyDeviceToHost));
        }
        for (int k = 0; k < nModes; ++k) {
            double* out = host.data() + (size_t)k * P2;
            if (plan.kind[k] == mfbd::ModePlan::CONSTANT) std::fill(out, out + P2, (double)plan.noll[k]);
            else if (plan.kind[k] == mfbd::ModePlan::ZERNIKE)
                std::copy(zs.begin() + (size_t)zIndex.at(plan.noll[k]) * P2,
                          zs.begin() + (size_t)(zIndex.at(plan.noll[k]) + 1) * P2, out);
        }
        std::copy(host.begin(), host.end(), modes);
        if (atmRms) std::copy(plan.atmRms.begin(), plan.atmRms.end(), atmRms);
        tiltXY[0] = plan.tilt[0];
        tiltXY[1] = plan.tilt[1];
    } catch (const std::exception& e) {
        g_basis_err = e.what();
        return MFBD_EDEVICE;
    }
    return MFBD_OK;
}
