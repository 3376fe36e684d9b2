// Mosaic of restored patches (Object::doMozaic, object.cpp:1160-1207): every patch
// weighted by the blend window of apodizeInPlace (utils.cpp:584-613) with margin =
// patchSize/8 and blend = (patchSize - 2 margin)/3, summed with its weights in FP64
// (mozaic / insertPatch, utils.hpp:389-502), divided where the weight sum exceeds 1e-6,
// then border rows and columns whose absolute sum is below 1e-15 trimmed (img_trim,
// utils.hpp:521-553).  One thread per output pixel gathers the patches that cover it, in
// patch order, so the sums have a fixed order (the reference's thread pool sums in whatever
// order its threads pop the patches).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/momfbd_hip.h"

namespace {

#define MCHK(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e_)); \
    } while (0)

// Output tile of one workgroup: MOS_TY rows of MOS_TX columns (one wavefront per row, so
// the patch reads and the image writes of a wavefront are 64 consecutive floats).
constexpr int MOS_TX = 64, MOS_TY = 4, MOS_NT = MOS_TX * MOS_TY;

// pixel (y, x) of the untrimmed mosaic: sum_i w(r, c) patch_i(r, c) / sum_i w(r, c) over
// the patches i with (y - posY_i, x - posX_i) = (r, c) inside their clipped extent.  The
// workgroup first lists, in patch order, the patches that touch its tile (MOS_NT candidates
// at a time, compacted by a ballot prefix), then every thread sums its pixel over the list,
// so each pixel sums its patches in the given order whatever the tiling.
__global__ __launch_bounds__(MOS_NT) void k_mosaic(float* __restrict__ out, const float* __restrict__ patches,
                                                   const double* __restrict__ w, const int* __restrict__ posY,
                                                   const int* __restrict__ posX, int nPatches, int pRows, int pCols,
                                                   int imgRows, int imgCols) {
#pragma clang fp contract(off)  // product and sum rounded separately, as the host sum
    __shared__ int list[MOS_NT];
    __shared__ int waveCount[MOS_NT / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int y0 = blockIdx.y * MOS_TY, x0 = blockIdx.x * MOS_TX;
    const int y = y0 + wave, x = x0 + lane;
    double acc = 0, sum = 0;
    for (int base = 0; base < nPatches; base += MOS_NT) {
        const int i = base + t;
        bool hit = false;
        if (i < nPatches) {
            const int py = posY[i], px = posX[i];
            // insertPatch clips a patch reaching the last row / column (posY + pRows >= imgRows)
            const int rows = py + pRows >= imgRows ? imgRows - py : pRows;
            const int cols = px + pCols >= imgCols ? imgCols - px : pCols;
            hit = py < y0 + MOS_TY && py + rows > y0 && px < x0 + MOS_TX && px + cols > x0;
        }
        const unsigned long long m = __ballot(hit);
        if (lane == 0) waveCount[wave] = __popcll(m);
        __syncthreads();
        int off = 0, total = 0;
        for (int v = 0; v < MOS_NT / 64; ++v) {
            off += v < wave ? waveCount[v] : 0;
            total += waveCount[v];
        }
        if (hit) list[off + __popcll(m & ((1ull << lane) - 1))] = i;
        __syncthreads();
        if (y < imgRows && x < imgCols)
            for (int j = 0; j < total; ++j) {
                const int k = list[j];
                const int py = posY[k], px = posX[k];
                const int rows = py + pRows >= imgRows ? imgRows - py : pRows;
                const int cols = px + pCols >= imgCols ? imgCols - px : pCols;
                const int r = y - py, c = x - px;
                if (r < 0 || r >= rows || c < 0 || c >= cols) continue;
                const double wt = w[(size_t)r * pCols + c];
                sum += wt;
                acc += wt * (double)patches[((size_t)k * pRows + r) * pCols + c];
            }
        __syncthreads();  // list and waveCount are rewritten by the next batch
    }
    if (y < imgRows && x < imgCols) out[(size_t)y * imgCols + x] = (float)(sum > 1E-6 ? acc / sum : acc);
}

// math::apodize (functions.cpp:39-50): data[i] = mean + amplitude cos(i pi/(n-1)), i >= 1
void apodize(double* data, int n, double target) {
    const double mean = 0.5 * (data[0] + target), amplitude = 0.5 * (data[0] - target);
    const double step = M_PI / (n - 1);
    for (int i = 1; i < n; ++i) data[i] = mean + amplitude * std::cos(i * step);
}

// apodizeInPlace of an all-ones nRows x nCols array (utils.cpp:584-613)
std::vector<double> blend_window(int nRows, int nCols, int blend, int margin) {
    int rowBlend = blend, colBlend = blend;
    const int rowMargin = margin, colMargin = margin;
    if (rowBlend + rowMargin > nCols / 2) rowBlend = nCols / 2 - rowMargin;
    if (colBlend + colMargin > nRows / 2) colBlend = nRows / 2 - colMargin;
    std::vector<double> w((size_t)nRows * nCols, 1.0);
    const int sz = std::max(rowBlend + rowMargin, colBlend + colMargin) + 2;
    std::vector<double> tmp(sz, 0.0);
    apodize(tmp.data() + rowMargin, rowBlend + 2, 1.0);
    for (int c = 0; c < nCols; ++c)
        for (int r = 0; r < rowBlend + rowMargin; ++r) {
            w[(size_t)r * nCols + c] *= tmp[r + 1];
            w[(size_t)(nRows - r - 1) * nCols + c] *= tmp[r + 1];
        }
    std::fill(tmp.begin(), tmp.end(), 0.0);
    apodize(tmp.data() + colMargin, colBlend + 2, 1.0);
    for (int r = 0; r < nRows; ++r)
        for (int c = 0; c < colBlend + colMargin; ++c) {
            w[(size_t)r * nCols + c] *= tmp[c + 1];
            w[(size_t)r * nCols + nCols - c - 1] *= tmp[c + 1];
        }
    return w;
}

// The blend window of an n x n patch is separable: apodizeInPlace multiplies each pixel of
// the all-ones array by one row taper (rows within blend + margin of the top or bottom edge)
// and then one column taper, and the two edge ranges of either taper never overlap (blend +
// margin <= n/2).  So w[r][c] = 1.0 * rowTaper[r] * colTaper[c] rounds exactly as the host
// window does.
__global__ __launch_bounds__(256) void k_blend_window(double* __restrict__ w, const double* __restrict__ rowTaper,
                                                      const double* __restrict__ colTaper, int nRows, int nCols) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nRows * nCols) return;
    const int r = (int)(i / nCols), c = (int)(i % nCols);
    w[i] = rowTaper[r] * colTaper[c];
}

// the row and column tapers of blend_window (the factors apodizeInPlace applies), written to
// caller storage of nRows and nCols doubles
void blend_tapers(int nRows, int nCols, int blend, int margin, double* rowTaper, double* colTaper, double* tmp) {
    int rowBlend = blend, colBlend = blend;
    const int rowMargin = margin, colMargin = margin;
    if (rowBlend + rowMargin > nCols / 2) rowBlend = nCols / 2 - rowMargin;
    if (colBlend + colMargin > nRows / 2) colBlend = nRows / 2 - colMargin;
    const int sz = std::max(rowBlend + rowMargin, colBlend + colMargin) + 2;
    for (int r = 0; r < nRows; ++r) rowTaper[r] = 1.0;
    for (int c = 0; c < nCols; ++c) colTaper[c] = 1.0;
    std::fill(tmp, tmp + sz, 0.0);
    apodize(tmp + rowMargin, rowBlend + 2, 1.0);
    for (int r = 0; r < rowBlend + rowMargin; ++r) {
        rowTaper[r] *= tmp[r + 1];
        rowTaper[nRows - r - 1] *= tmp[r + 1];
    }
    std::fill(tmp, tmp + sz, 0.0);
    apodize(tmp + colMargin, colBlend + 2, 1.0);
    for (int c = 0; c < colBlend + colMargin; ++c) {
        colTaper[c] *= tmp[c + 1];
        colTaper[nCols - c - 1] *= tmp[c + 1];
    }
}

// img_trim (utils.hpp:521-553) of the untrimmed rows x cols mosaic held in img: drops border
// columns and rows whose absolute sum is below 1e-15, moves the kept window to the front of
// img and returns its extent and first row / column.  The sums run in the same order as the
// allocating path's column and row sums.
void trim_in_place(float* img, int rows, int cols, int clip, int* outRows, int* outCols, int* r0o, int* c0o) {
    int r0 = 0, r1 = rows - 1, c0 = 0, c1 = cols - 1;
    if (clip) {
        const float thr = 1E-15f;
        auto colSum = [&](int c) {
            double v = 0;
            for (int r = 0; r < rows; ++r) v += std::fabs(img[(size_t)r * cols + c]);
            return v;
        };
        auto rowSum = [&](int r) {
            double v = 0;
            for (int c = 0; c < cols; ++c) v += std::fabs(img[(size_t)r * cols + c]);
            return v;
        };
        while (c0 < c1 && colSum(c0) < thr) ++c0;
        while (c1 > c0 && colSum(c1) < thr) --c1;
        while (r0 < r1 && rowSum(r0) < thr) ++r0;
        while (r1 > r0 && rowSum(r1) < thr) --r1;
    }
    const int nr = r1 - r0 + 1, nc = c1 - c0 + 1;
    for (int r = 0; r < nr; ++r)          // destinations never pass their sources
        std::memmove(img + (size_t)r * nc, img + (size_t)(r + r0) * cols + c0, nc * sizeof(float));
    *outRows = nr;
    *outCols = nc;
    if (r0o) *r0o = r0;
    if (c0o) *c0o = c0;
}

// Fixed device scratch of the allocation-free mosaic (mfbd_mosaic_noalloc): tapers, window,
// positions, patches and the untrimmed mosaic, each region 256-byte aligned.  Every device
// the library runs on holds one copy until the process ends; host code reaches it only
// through the per-device address hipGetSymbolAddress returns.
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr int kNoAllocMaxSide = 2048;
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex scratchMtx;
void* scratchPtr[kMaxDevices] = {};

// the device address of `scratch` on the current device, looked up on the first call there
hipError_t scratch_address(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lock(scratchMtx);
    if (!scratchPtr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratchPtr[device] = p;
    }
    *out = static_cast<unsigned char*>(scratchPtr[device]);
    return hipSuccess;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

template <typename T>
struct DevBuf {
    T* p = nullptr;
    explicit DevBuf(size_t n) { MCHK(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T))); }
    ~DevBuf() { (void)hipFree(p); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

}  // namespace

extern "C" int mfbd_mosaic_size(int nPatches, int patchSize, const int* posY, const int* posX, int* imgRows,
                                int* imgCols) {
    if (nPatches < 1 || patchSize < 1 || !posY || !posX || !imgRows || !imgCols) return MFBD_EINVAL;
    int maxY = 0, maxX = 0;
    for (int i = 0; i < nPatches; ++i) {
        if (posY[i] < 0 || posX[i] < 0) return MFBD_EINVAL;
        maxY = std::max(maxY, posY[i]);
        maxX = std::max(maxX, posX[i]);
    }
    *imgRows = maxY + patchSize + 1;                      // doMozaic's extent before trimming
    *imgCols = maxX + patchSize + 1;
    return MFBD_OK;
}

extern "C" int mfbd_mosaic(int nPatches, int patchSize, const float* patches, const int* posY, const int* posX,
                           int clip, int device, float* img, int* imgRows, int* imgCols, int* firstRow, int* firstCol) {
    int rows = 0, cols = 0;
    int rc = mfbd_mosaic_size(nPatches, patchSize, posY, posX, &rows, &cols);
    if (rc) return rc;
    if (!patches || !img || (rows + MOS_TY - 1) / MOS_TY > 65535) return MFBD_EINVAL;
    try {
        MCHK(hipSetDevice(device));
        const int margin = patchSize / 8, blend = (patchSize - 2 * margin) / 3;
        const std::vector<double> w = blend_window(patchSize, patchSize, blend, margin);
        const size_t nPix = (size_t)rows * cols, pSz = (size_t)patchSize * patchSize;
        DevBuf<float> dP((size_t)nPatches * pSz), dOut(nPix);
        DevBuf<double> dW(pSz);
        DevBuf<int> dY(nPatches), dX(nPatches);
        MCHK(hipMemcpy(dP.p, patches, (size_t)nPatches * pSz * sizeof(float), hipMemcpyHostToDevice));
        MCHK(hipMemcpy(dW.p, w.data(), pSz * sizeof(double), hipMemcpyHostToDevice));
        MCHK(hipMemcpy(dY.p, posY, nPatches * sizeof(int), hipMemcpyHostToDevice));
        MCHK(hipMemcpy(dX.p, posX, nPatches * sizeof(int), hipMemcpyHostToDevice));
        const dim3 grid((unsigned)((cols + MOS_TX - 1) / MOS_TX), (unsigned)((rows + MOS_TY - 1) / MOS_TY));
        hipLaunchKernelGGL(k_mosaic, grid, dim3(MOS_NT), 0, 0, dOut.p, dP.p, dW.p, dY.p, dX.p, nPatches, patchSize,
                           patchSize, rows, cols);
        MCHK(hipGetLastError());
        std::vector<float> full(nPix);
        MCHK(hipMemcpy(full.data(), dOut.p, nPix * sizeof(float), hipMemcpyDeviceToHost));
        // img_trim: drop border columns / rows whose absolute sum is below 1e-15
        int r0 = 0, r1 = rows - 1, c0 = 0, c1 = cols - 1;
        if (clip) {
            std::vector<double> colSum(cols, 0.0), rowSum(rows, 0.0);
            for (int c = 0; c < cols; ++c)
                for (int r = 0; r < rows; ++r) {
                    const double v = std::fabs(full[(size_t)r * cols + c]);
                    colSum[c] += v;
                    rowSum[r] += v;
                }
            const float thr = 1E-15f;
            while (c0 < c1 && colSum[c0] < thr) ++c0;
            while (c1 > c0 && colSum[c1] < thr) --c1;
            while (r0 < r1 && rowSum[r0] < thr) ++r0;
            while (r1 > r0 && rowSum[r1] < thr) --r1;
        }
        const int outRows = r1 - r0 + 1, outCols = c1 - c0 + 1;
        for (int r = 0; r < outRows; ++r)
            std::memcpy(img + (size_t)r * outCols, full.data() + (size_t)(r + r0) * cols + c0, outCols * sizeof(float));
        *imgRows = outRows;
        *imgCols = outCols;
        if (firstRow) *firstRow = r0;
        if (firstCol) *firstCol = c0;
    } catch (const std::exception&) {
        return MFBD_EDEVICE;
    }
    return MFBD_OK;
}

extern "C" int mfbd_mosaic_noalloc(int nPatches, int patchSize, const float* patches, const int* posY,
                                   const int* posX, int clip, int device, float* img, int* imgRows, int* imgCols,
                                   int* firstRow, int* firstCol) {
    int rows = 0, cols = 0;
    int rc = mfbd_mosaic_size(nPatches, patchSize, posY, posX, &rows, &cols);
    if (rc) return rc;
    if (!patches || !img || (rows + MOS_TY - 1) / MOS_TY > 65535 || device < 0 || device >= kMaxDevices)
        return MFBD_EINVAL;
    if (patchSize > kNoAllocMaxSide) return MFBD_ESCRATCH;
    const size_t pSz = (size_t)patchSize * patchSize, nPix = (size_t)rows * cols;
    const size_t oRow = 0, oCol = oRow + align256(patchSize * sizeof(double));
    const size_t oW = oCol + align256(patchSize * sizeof(double));
    const size_t oY = oW + align256(pSz * sizeof(double));
    const size_t oX = oY + align256(nPatches * sizeof(int));
    const size_t oP = oX + align256(nPatches * sizeof(int));
    const size_t oOut = oP + align256((size_t)nPatches * pSz * sizeof(float));
    if (oOut + nPix * sizeof(float) > kScratchBytes) return MFBD_ESCRATCH;
    const int margin = patchSize / 8, blend = (patchSize - 2 * margin) / 3;
    double rowTaper[kNoAllocMaxSide], colTaper[kNoAllocMaxSide], tmp[kNoAllocMaxSide + 2];
    blend_tapers(patchSize, patchSize, blend, margin, rowTaper, colTaper, tmp);
    if (hipSetDevice(device) != hipSuccess) return MFBD_EDEVICE;
    unsigned char* base = nullptr;
    if (scratch_address(device, &base) != hipSuccess) return MFBD_EDEVICE;
    double* dRow = reinterpret_cast<double*>(base + oRow);
    double* dCol = reinterpret_cast<double*>(base + oCol);
    double* dW = reinterpret_cast<double*>(base + oW);
    int* dY = reinterpret_cast<int*>(base + oY);
    int* dX = reinterpret_cast<int*>(base + oX);
    float* dP = reinterpret_cast<float*>(base + oP);
    float* dOut = reinterpret_cast<float*>(base + oOut);
    if (hipMemcpy(dRow, rowTaper, patchSize * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dCol, colTaper, patchSize * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dY, posY, nPatches * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dX, posX, nPatches * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dP, patches, (size_t)nPatches * pSz * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return MFBD_EDEVICE;
    hipLaunchKernelGGL(k_blend_window, dim3((unsigned)((pSz + 255) / 256)), dim3(256), 0, 0, dW, dRow, dCol, patchSize,
                       patchSize);
    if (hipGetLastError() != hipSuccess) return MFBD_EDEVICE;
    const dim3 grid((unsigned)((cols + MOS_TX - 1) / MOS_TX), (unsigned)((rows + MOS_TY - 1) / MOS_TY));
    hipLaunchKernelGGL(k_mosaic, grid, dim3(MOS_NT), 0, 0, dOut, dP, dW, dY, dX, nPatches, patchSize, patchSize, rows,
                       cols);
    if (hipGetLastError() != hipSuccess) return MFBD_EDEVICE;
    if (hipMemcpy(img, dOut, nPix * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return MFBD_EDEVICE;
    trim_in_place(img, rows, cols, clip, imgRows, imgCols, firstRow, firstCol);
    return MFBD_OK;
}
