// Device pipeline of the MOMFBD patch solver for gfx950 (MI355X).
//
// Per evaluation of a batch of requests (patch slot + alpha):
//   k_phase_mfma   phi = phi_channel + B alpha           (mode expansion, subimage.cpp:509-544)
//   k_field        a = otfNorm * pupil * exp(i phi) on the OTF grid  (subimage.cpp:586-612)
//   FFT2 / |.|^2 / IFFT2 / k_extract   S = reorder(IFFT(|FFT(a)|^2)/N^2)  (fouriertransform.hpp:280-286)
//   k_pq           P = sum conj(D) S, Q = gamma + sum |S|^2 and metric partials
//                  (subimage.cpp:226-234, object.cpp:459-462, 508-518, 597-615)
//   k_metric       L = sum_o w_o L_o / patchSize^2        (solver.cpp:763-817)
// Gradient (Vogel / adjoint, subimage.cpp:312-390):
//   k_build_g, IFFT2, k_save_re, k_field(de-centred), IFFT2, k_mul_re, FFT2, k_extract_v, k_grad
// Finite-difference gradient (subimage.cpp:261-292): k_phase_pert + OTF chain + k_metric_change.
//
// FFTs: batched 2-D transforms as a row pass and a column pass; each workgroup
// stages a tile of lines in LDS and runs a self-sorting (Stockham) radix-16/8
// transform with the radix butterflies held in registers.  Twiddles come from a
// table computed on the host in long double.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "evaluator.hpp"
#include "line_search.hpp"

namespace mfbd {

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " #x); \
    } while (0)

template <typename T>
struct alignas(2 * sizeof(T)) cx {   // aligned as a unit: one 64/128-bit LDS or global access
    T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }

__device__ __forceinline__ void sincos_t(double a, double* s, double* c) { sincos(a, s, c); }
__device__ __forceinline__ void sincos_t(float a, float* s, float* c) { sincosf(a, s, c); }

// ---------------------------------------------------------------- FFT core ----

constexpr int bitrev(int i, int R) {
    int r = 0;
    for (int b = 1; b < R; b <<= 1) { r = (r << 1) | (i & 1); i >>= 1; }
    return r;
}

// 16th roots of unity exp(-2 pi i k / 16), k = 0..15
__device__ __forceinline__ double root_c(int k) {
    constexpr double c[16] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                              0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                              -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
                              0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
    return c[k & 15];
}
__device__ __forceinline__ double root_s(int k) { return -root_c((k + 12) & 15); }   // -sin(2 pi k/16)

// a * exp(DIR 2 pi i k / 16) for a compile-time k in [0, 16): the multiplications by 0
// and 1 of k = 0, 4 and the common factor of k = 2, 6 are written out, since IEEE
// arithmetic does not let the compiler fold them
template <typename T, int DIR>
__device__ __forceinline__ cx<T> rot16(cx<T> a, int k) {
    constexpr T h = (T)0.70710678118654752440;
    switch (k) {
        case 0: return a;
        case 4: return DIR < 0 ? cx<T>{a.y, -a.x} : cx<T>{-a.y, a.x};
        case 2: return DIR < 0 ? cx<T>{(a.x + a.y) * h, (a.y - a.x) * h} : cx<T>{(a.x - a.y) * h, (a.y + a.x) * h};
        case 6: return DIR < 0 ? cx<T>{(a.y - a.x) * h, -(a.x + a.y) * h} : cx<T>{-(a.x + a.y) * h, (a.x - a.y) * h};
        case 8: return cx<T>{-a.x, -a.y};
        default: {
            const cx<T> w = {(T)root_c(k), (T)(DIR < 0 ? root_s(k) : -root_s(k))};
            return cmul(a, w);
        }
    }
}

// 4-point DFT in place on v[i0], v[i0+s], v[i0+2s], v[i0+3s]: additions and a
// multiplication by -i (forward) / +i (backward) only
template <typename T, int DIR>
__device__ __forceinline__ void dft4(cx<T>* v, int i0, int s) {
    const cx<T> a = v[i0], b = v[i0 + s], c = v[i0 + 2 * s], d = v[i0 + 3 * s];
    const cx<T> t0 = cadd(a, c), t1 = csub(a, c), t2 = cadd(b, d), t3 = csub(b, d);
    const cx<T> t3r = DIR < 0 ? cx<T>{t3.y, -t3.x} : cx<T>{-t3.y, t3.x};   // t3 * (-+i)
    v[i0] = cadd(t0, t2);
    v[i0 + s] = cadd(t1, t3r);
    v[i0 + 2 * s] = csub(t0, t2);
    v[i0 + 3 * s] = csub(t1, t3r);
}

// 16-point DFT as 4 x 4: DFT4 of the stride-4 subsequences, twiddles W16^(m k1),
// DFT4 across m; output in natural order
template <typename T, int DIR>
__device__ __forceinline__ void dft16(cx<T>* v) {
#pragma unroll
    for (int m = 0; m < 4; ++m) dft4<T, DIR>(v, m, 4);         // Y_m[k1] at v[m + 4 k1]
#pragma unroll
    for (int m = 1; m < 4; ++m)
#pragma unroll
        for (int k1 = 1; k1 < 4; ++k1) v[m + 4 * k1] = rot16<T, DIR>(v[m + 4 * k1], m * k1);
    cx<T> w[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
#pragma unroll
        for (int m = 0; m < 4; ++m) w[4 * k1 + m] = v[m + 4 * k1];
        dft4<T, DIR>(w, 4 * k1, 1);                             // X[k1 + 4 k2] = w[4 k1 + k2]
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = w[4 * k1 + k2];
}

// in-register DFT of R points (R power of two <= 16), DIR = -1 forward, +1 backward
template <typename T, int R, int DIR>
__device__ __forceinline__ void dft_reg(cx<T>* v) {
    if constexpr (R == 16) {
        dft16<T, DIR>(v);
        return;
    }
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const int j = bitrev(i, R);
        if (i < j) { cx<T> t = v[i]; v[i] = v[j]; v[j] = t; }
    }
#pragma unroll
    for (int L = 2; L <= R; L <<= 1) {
#pragma unroll
        for (int i = 0; i < R; i += L) {
#pragma unroll
            for (int j = 0; j < L / 2; ++j) {
                const cx<T> t = rot16<T, DIR>(v[i + j + L / 2], j * (16 / L));
                v[i + j + L / 2] = csub(v[i + j], t);
                v[i + j] = cadd(v[i + j], t);
            }
        }
    }
}

// One Stockham pass of radix R over NL lines of length N held in LDS.
// Element k of line l is at buf[l*ls + k*es].  NS = span already transformed.
// MAP 0: consecutive threads take consecutive butterflies of one line (lines
// contiguous in LDS); MAP 1: consecutive threads take consecutive lines (lines
// interleaved in LDS) -- each keeps neighbouring lanes on neighbouring banks.
template <typename T, int N, int R, int NS, int DIR, int NL, int NT, int MAP = 0>
__device__ __forceinline__ void spass(cx<T>* buf, int ls, int es, const cx<T>* __restrict__ tw) {
    constexpr int NB = NL * (N / R);
    constexpr int BT = (NB + NT - 1) / NT;
    cx<T> v[BT][R];
#pragma unroll
    for (int i = 0; i < BT; ++i) {
        const int b = threadIdx.x + i * NT;
        if (NB % NT == 0 || b < NB) {
            const int line = MAP ? b % NL : b / (N / R), j = MAP ? b / NL : b % (N / R);
            const cx<T>* L = buf + line * ls;
            const int k = j % NS;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                cx<T> x = L[(j + r * (N / R)) * es];
                if (NS > 1 && r > 0) {
                    cx<T> w = tw[r * k * (N / (NS * R))];
                    if (DIR > 0) w.y = -w.y;
                    x = cmul(x, w);
                }
                v[i][r] = x;
            }
            dft_reg<T, R, DIR>(v[i]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < BT; ++i) {
        const int b = threadIdx.x + i * NT;
        if (NB % NT == 0 || b < NB) {
            const int line = MAP ? b % NL : b / (N / R), j = MAP ? b / NL : b % (N / R);
            cx<T>* L = buf + line * ls;
            const int k = j % NS;
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) L[(base + r * NS) * es] = v[i][r];
        }
    }
    __syncthreads();
}

template <typename T, int N, int DIR, int NL, int NT, int MAP = 0>
__device__ __forceinline__ void fft_lines(cx<T>* buf, int ls, int es, const cx<T>* __restrict__ tw) {
    if constexpr (N == 64) {
        spass<T, N, 8, 1, DIR, NL, NT, MAP>(buf, ls, es, tw);
        spass<T, N, 8, 8, DIR, NL, NT, MAP>(buf, ls, es, tw);
    } else if constexpr (N == 128) {
        spass<T, N, 16, 1, DIR, NL, NT, MAP>(buf, ls, es, tw);
        spass<T, N, 8, 16, DIR, NL, NT, MAP>(buf, ls, es, tw);
    } else if constexpr (N == 256) {
        spass<T, N, 16, 1, DIR, NL, NT, MAP>(buf, ls, es, tw);
        spass<T, N, 16, 16, DIR, NL, NT, MAP>(buf, ls, es, tw);
    } else {
        static_assert(N == 64 || N == 128 || N == 256, "unsupported OTF grid");
    }
}

// Row pass: each block transforms LR consecutive rows of one N x N tile in place.
template <typename T, int N, int DIR, int LR>
__global__ void __launch_bounds__(LR * N / 16) k_fft_rows(cx<T>* __restrict__ data, const cx<T>* __restrict__ tw) {
    constexpr int NT = LR * N / 16;
    __shared__ cx<T> lds[LR * N];
    const int chunks = N / LR;
    const size_t tile = blockIdx.x / chunks, chunk = blockIdx.x % chunks;
    cx<T>* g = data + tile * N * N + chunk * LR * N;
    for (int e = threadIdx.x; e < LR * N; e += NT) lds[e] = g[e];
    __syncthreads();
    fft_lines<T, N, DIR, LR, NT>(lds, N, 1, tw);
    for (int e = threadIdx.x; e < LR * N; e += NT) g[e] = lds[e];
}

// Column pass: each block transforms LC consecutive columns of one tile in place.
template <typename T, int N, int DIR, int LC>
__global__ void __launch_bounds__(LC * N / 16) k_fft_cols(cx<T>* __restrict__ data, const cx<T>* __restrict__ tw) {
    constexpr int NT = LC * N / 16;
    __shared__ cx<T> lds[LC * N];
    const int chunks = N / LC;
    const size_t tile = blockIdx.x / chunks, chunk = blockIdx.x % chunks;
    cx<T>* g = data + tile * N * N + chunk * LC;
    for (int e = threadIdx.x; e < LC * N; e += NT) {
        const int r = e / LC, c = e % LC;
        lds[e] = g[(size_t)r * N + c];
    }
    __syncthreads();
    fft_lines<T, N, DIR, LC, NT, 1>(lds, 1, LC, tw);
    for (int e = threadIdx.x; e < LC * N; e += NT) {
        const int r = e / LC, c = e % LC;
        g[(size_t)r * N + c] = lds[e];
    }
}

// The multi-pass OTF (grids too large for one tile in LDS) in four passes over the tile
// workspace, the elementwise stages folded into the transform passes and the real and
// Hermitian structure used:
//   k_field_rows     pupil field a = mag exp(i phi) built in LDS, forward row FFTs; only
//                    the rows the pupil touches (rowNZ) exist, the others are zero
//   k_cols_power     forward column FFTs (zero rows not read), |.|^2 to a real array
//   k_rows_inv_real  inverse row FFTs of the real rows, two per complex transform; the
//                    Hermitian rows are written in columns [0, N/2] only
//   k_cols_extract   inverse column FFTs of the columns [0, N/2] (where the stored half
//                    of the support lies), its points scaled by 1/N^2 written to S
//                    (invOs: stored index h >= 0, else not stored)
// Images a patch lacks (imgScale 0) skip every pass; k_cols_extract writes their zero OTF.
template <typename T, int N, int LR>
__global__ void __launch_bounds__(LR * N / 16) k_field_rows(cx<T>* __restrict__ data, const T* __restrict__ phi,
                                                            const int* __restrict__ inv, const T* __restrict__ pupil,
                                                            const int* __restrict__ qImg, const T* __restrict__ imgScale,
                                                            const unsigned char* __restrict__ rowNZ, int nPS,
                                                            const cx<T>* __restrict__ tw) {
    constexpr int NT = LR * N / 16;
    __shared__ cx<T> lds[LR * N];
    const int chunks = N / LR;
    const size_t tile = blockIdx.x / chunks, chunk = blockIdx.x % chunks;
    const T sc = imgScale[qImg[tile]];
    if (sc == T(0)) return;
    bool any = false;
    for (int l = 0; l < LR; ++l) any |= rowNZ[chunk * LR + l] != 0;
    if (!any) return;
    const T* ph = phi + tile * nPS;
    const int p0 = (int)chunk * LR * N;
    for (int e = threadIdx.x; e < LR * N; e += NT) {
        const int idx = inv[p0 + e];
        cx<T> v = {T(0), T(0)};
        if (idx >= 0) {
            T sn, cs;
            sincos_t(ph[idx], &sn, &cs);
            const T mag = pupil[idx] * sc;
            v = {mag * cs, mag * sn};
        }
        lds[e] = v;
    }
    __syncthreads();
    fft_lines<T, N, -1, LR, NT>(lds, N, 1, tw);
    cx<T>* g = data + tile * N * N + p0;
    for (int e = threadIdx.x; e < LR * N; e += NT) g[e] = lds[e];
}

template <typename T, int N, int LC>
__global__ void __launch_bounds__(LC * N / 16) k_cols_power(T* __restrict__ power, const cx<T>* __restrict__ data,
                                                            const unsigned char* __restrict__ rowNZ,
                                                            const int* __restrict__ qImg, const T* __restrict__ imgScale,
                                                            const cx<T>* __restrict__ tw) {
    constexpr int NT = LC * N / 16;
    __shared__ cx<T> lds[LC * N];
    const int chunks = N / LC;
    const size_t tile = blockIdx.x / chunks, chunk = blockIdx.x % chunks;
    if (imgScale[qImg[tile]] == T(0)) return;
    const cx<T>* g = data + tile * N * N + chunk * LC;
    for (int e = threadIdx.x; e < LC * N; e += NT) {
        const int r = e / LC, c = e % LC;
        lds[e] = rowNZ[r] ? g[(size_t)r * N + c] : cx<T>{T(0), T(0)};
    }
    __syncthreads();
    fft_lines<T, N, -1, LC, NT, 1>(lds, 1, LC, tw);
    T* pw = power + tile * N * N + chunk * LC;
    for (int e = threadIdx.x; e < LC * N; e += NT) {
        const int r = e / LC, c = e % LC;
        const cx<T> v = lds[e];
        pw[(size_t)r * N + c] = v.x * v.x + v.y * v.y;
    }
}

// LR complex lines per block, each the pair of real rows (2l, 2l+1) of the block's 2 LR rows:
// z = x1 + i x2, Z = IFFT(z), X1(k) = (Z(k) + Z(N-k)*)/2, X2(k) = (Z(k) - Z(N-k)*)/(2i)
template <typename T, int N, int LR>
__global__ void __launch_bounds__(LR * N / 16) k_rows_inv_real(cx<T>* __restrict__ data, const T* __restrict__ power,
                                                               const int* __restrict__ qImg,
                                                               const T* __restrict__ imgScale,
                                                               const cx<T>* __restrict__ tw) {
    constexpr int NT = LR * N / 16, H = N / 2 + 1;
    __shared__ cx<T> lds[LR * N];
    const int chunks = N / (2 * LR);
    const size_t tile = blockIdx.x / chunks, chunk = blockIdx.x % chunks;
    if (imgScale[qImg[tile]] == T(0)) return;
    const size_t r0 = chunk * 2 * LR;
    const T* pw = power + tile * N * N + r0 * N;
    for (int e = threadIdx.x; e < LR * N; e += NT) {
        const int l = e / N, n = e % N;
        lds[e] = {pw[(size_t)(2 * l) * N + n], pw[(size_t)(2 * l + 1) * N + n]};
    }
    __syncthreads();
    fft_lines<T, N, +1, LR, NT>(lds, N, 1, tw);
    cx<T>* g = data + tile * N * N + r0 * N;
    for (int e = threadIdx.x; e < LR * H; e += NT) {
        const int l = e / H, k = e % H;
        const cx<T> zk = lds[l * N + k], zm = lds[l * N + (N - k) % N];
        g[(size_t)(2 * l) * N + k] = {T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)};
        g[(size_t)(2 * l + 1) * N + k] = {T(0.5) * (zk.y + zm.y), T(0.5) * (zm.x - zk.x)};
    }
}

// columns [0, (N/2)/LC * LC + LC) in `chunks` blocks of LC per tile; those beyond N/2 read as
// zero (no stored point lies there)
template <typename T, int N, int LC>
__global__ void __launch_bounds__(LC * N / 16) k_cols_extract(cx<T>* __restrict__ S, const cx<T>* __restrict__ data,
                                                              const int* __restrict__ invOs, int nOS, T scale,
                                                              const int* __restrict__ qImg, const T* __restrict__ imgScale,
                                                              int chunks, const cx<T>* __restrict__ tw) {
    constexpr int NT = LC * N / 16;
    __shared__ cx<T> lds[LC * N];
    const size_t tile = blockIdx.x / chunks, chunk = blockIdx.x % chunks;
    const int c0 = (int)chunk * LC;
    cx<T>* So = S + tile * nOS;
    if (imgScale[qImg[tile]] == T(0)) {       // an image the patch lacks: zero OTF
        for (int e = threadIdx.x; e < LC * N; e += NT) {
            const int h = invOs[(e / LC) * N + c0 + e % LC];
            if (h >= 0) So[h] = {T(0), T(0)};
        }
        return;
    }
    const cx<T>* g = data + tile * N * N + c0;
    for (int e = threadIdx.x; e < LC * N; e += NT) {
        const int r = e / LC, c = e % LC;
        lds[e] = c0 + c <= N / 2 ? g[(size_t)r * N + c] : cx<T>{T(0), T(0)};
    }
    __syncthreads();
    fft_lines<T, N, +1, LC, NT, 1>(lds, 1, LC, tw);
    for (int e = threadIdx.x; e < LC * N; e += NT) {
        const int r = e / LC, c = e % LC;
        const int h = invOs[r * N + c0 + c];
        if (h >= 0) {
            const cx<T> v = lds[e];
            So[h] = {v.x * scale, v.y * scale};
        }
    }
}

// ------------------------------------------------------------ elementwise ----

struct QMap {             // virtual image q -> source
    const int* slot;      // patch slot
    const int* img;       // image index within the patch
    const int* arow;      // alpha row
};

// Mode expansion as a GEMM on the matrix cores: phi = phiCh + A B with A = alpha rows
// (FP64 in memory, converted on load) and B = modes [K][nPS] of the image's object,
// K stepped by 4 through v_mfma_{f32,f64}_16x16x4.  Every wave is independent (no LDS,
// no barrier): it owns 16 images x 64 pupil points (4 accumulator tiles), issues all
// alpha and mode fragment loads of a chunk of k-steps at once and then runs the MFMAs.
// Lane layout of the 16x16x4 MFMAs (measured on gfx950): A[i][k] from lane i + 16k,
// B[k][j] from lane j + 16k; C[i][j] is in lane j + 16(i/4), register i%4 for f32, and
// in lane j + 16(i%4), register i/4 for f64.  Rows of different objects in one wave
// take one pass per object with the other rows' A zeroed, so every element is the same
// k-ordered sum whatever shares its tile.
template <typename T>
struct MfmaVec;
template <>
struct MfmaVec<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    __device__ static v4 mac(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static int row(int lk, int v) { return 4 * lk + v; }     // accumulator row of (lane/16, register)
};
template <>
struct MfmaVec<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    __device__ static v4 mac(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int row(int lk, int v) { return lk + 4 * v; }
};

template <typename T>
struct PhaseChunk { static constexpr int value = 2; };         // k-steps whose fragments are loaded together
template <>
struct PhaseChunk<double> { static constexpr int value = 1; };
// Per 256-patch config-2 step (accumulators in VGPRs): FP64 98 ms at one k-step per chunk,
// 100 at 2, 105 at 3, 111 at 5 (fewer registers, more waves per SIMD to cover the fragment
// loads); FP32 45 ms at 7, 46 at 4, 48 at 13 on its own, but at 2 (62 registers) it fits in
// the 64 per SIMD lane that k_otf_wave leaves, beside the other lanes' transforms: the FP32
// step 473 -> 478 patches/s

// Row q of a phase GEMM goes to phi + q nPS, or with an output table to phi + qOut[q] nPS
// (qOut[q] >= 0) or phi1 + (-1 - qOut[q]) nPS; a channel row qPhRow[q] < 0 adds nothing.
template <typename T>
__device__ __forceinline__ T* phase_out_row(T* phi, T* phi1, const int* qOut, int q, int nPS) {
    if (!qOut) return phi + (size_t)q * nPS;
    const int r = qOut[q];
    return r >= 0 ? phi + (size_t)r * nPS : phi1 + (size_t)(-1 - r) * nPS;
}

template <typename T>
__global__ void __launch_bounds__(256) k_phase_mfma(T* __restrict__ phi, const T* __restrict__ B,
                                                    const T* __restrict__ phiCh, const double* __restrict__ alpha,
                                                    const int* __restrict__ qArow, const int* __restrict__ qPhRow,
                                                    const int* __restrict__ qObj, int nQ, int K, int nPS,
                                                    T* __restrict__ phi1 = nullptr, const int* __restrict__ qOut = nullptr,
                                                    const int* __restrict__ qOut2 = nullptr) {
    using V = typename MfmaVec<T>::v4;
    constexpr int KS = PhaseChunk<T>::value;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int q0 = blockIdx.y * 64 + 16 * w, c0 = blockIdx.x * 64;
    if (q0 >= nQ) return;                                    // whole wave exits together
    // this lane's A row and the 4 output rows it holds
    const int qa = min(q0 + li, nQ - 1);
    const int objA = (q0 + li < nQ) ? qObj[qa] : INT_MAX;
    const double* arow = alpha + (size_t)qArow[qa] * K;
    V acc[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int q = min(q0 + MfmaVec<T>::row(lk, v), nQ - 1);
        const int pr = qPhRow[q];
        const T* pc = phiCh + (size_t)(pr >= 0 ? pr : 0) * nPS;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int col = c0 + 16 * c + li;
            acc[c][v] = (pr >= 0 && col < nPS) ? pc[col] : T(0);
        }
    }
    // distinct objects among the wave's 16 rows, ascending (wave-uniform)
    int o = INT_MAX;
    for (int r = 0; r < 16; ++r) o = min(o, __shfl(objA, r));
    while (o != INT_MAX) {
        const T* Bo = B + (size_t)o * K * nPS;
        const bool mine = objA == o;
        for (int s0 = 0; 4 * s0 < K; s0 += KS) {
            T a[KS], b[4][KS];
#pragma unroll
            for (int t = 0; t < KS; ++t) {
                const int k = 4 * (s0 + t) + lk;
                a[t] = (mine && k < K) ? (T)arow[k] : T(0);
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int col = c0 + 16 * c + li;
                    b[c][t] = (k < K && col < nPS) ? Bo[(size_t)k * nPS + col] : T(0);
                }
            }
#pragma unroll
            for (int t = 0; t < KS; ++t)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[c] = MfmaVec<T>::mac(a[t], b[c][t], acc[c]);
        }
        int next = INT_MAX;
        for (int r = 0; r < 16; ++r) {
            const int x = __shfl(objA, r);
            if (x > o) next = min(next, x);
        }
        o = next;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int q = q0 + MfmaVec<T>::row(lk, v), col = c0 + 16 * c + li;
            if (q < nQ && col < nPS) {
                phase_out_row(phi, phi1, qOut, q, nPS)[col] = acc[c][v];
                if (qOut2 && qOut2[q] >= 0) phi1[(size_t)qOut2[q] * nPS + col] = acc[c][v];
            }
        }
    }
}

// Mixed-precision mode expansion: phi = phiCh + A B with the basis in BF16 and FP32
// accumulation on v_mfma_f32_16x16x32_bf16 (gfx950).  Bt holds the basis transposed,
// [obj][nPS][Kp] BF16 with Kp a multiple of 32, so the 8 consecutive modes a lane
// feeds per k-step are one 16-byte load.  alpha enters as two BF16 parts, hi =
// bf16(a) and lo = bf16(a - hi), i.e. about 16 significant bits, at two MFMAs per
// k-step.  Lane layout (measured on gfx950): A[i][k] from lane i + 16(k/8), element
// k%8; B[k][j] from lane j + 16(k/8), element k%8; C as the f32 16x16x4 form.  One wave
// = 16 images x 64 pupil points, independent of the other waves.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__global__ void __launch_bounds__(256) k_phase_bf16(float* __restrict__ phi, const bf16x8* __restrict__ Bt,
                                                    const float* __restrict__ phiCh, const double* __restrict__ alpha,
                                                    const int* __restrict__ qArow, const int* __restrict__ qPhRow,
                                                    const int* __restrict__ qObj, int nQ, int K, int Kp, int nPS,
                                                    float* __restrict__ phi1 = nullptr,
                                                    const int* __restrict__ qOut = nullptr,
                                                    const int* __restrict__ qOut2 = nullptr) {
    using V = MfmaVec<float>::v4;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int q0 = blockIdx.y * 64 + 16 * w, c0 = blockIdx.x * 64;
    if (q0 >= nQ) return;                                    // whole wave exits together
    const int qa = min(q0 + li, nQ - 1);
    const int objA = (q0 + li < nQ) ? qObj[qa] : INT_MAX;
    const double* arow = alpha + (size_t)qArow[qa] * K;
    V acc[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int q = min(q0 + MfmaVec<float>::row(lk, v), nQ - 1);
        const int pr = qPhRow[q];
        const float* pc = phiCh + (size_t)(pr >= 0 ? pr : 0) * nPS;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int col = c0 + 16 * c + li;
            acc[c][v] = (pr >= 0 && col < nPS) ? pc[col] : 0.f;
        }
    }
    int o = INT_MAX;
    for (int r = 0; r < 16; ++r) o = min(o, __shfl(objA, r));
    while (o != INT_MAX) {
        const bf16x8* Bo = Bt + (size_t)o * nPS * (Kp / 8);
        const bool mine = objA == o;
        for (int k0 = 0; k0 < Kp; k0 += 32) {
            bf16x8 ahi, alo;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = k0 + 8 * lk + e;
                const float a = (mine && k < K) ? (float)arow[k] : 0.f;
                const __bf16 h = (__bf16)a;
                ahi[e] = h;
                alo[e] = (__bf16)(a - (float)h);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int col = c0 + 16 * c + li;
                bf16x8 b = {};
                if (col < nPS) b = Bo[(size_t)col * (Kp / 8) + k0 / 8 + lk];
                acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, b, acc[c], 0, 0, 0);
                acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, b, acc[c], 0, 0, 0);
            }
        }
        int next = INT_MAX;
        for (int r = 0; r < 16; ++r) {
            const int x = __shfl(objA, r);
            if (x > o) next = min(next, x);
        }
        o = next;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int q = q0 + MfmaVec<float>::row(lk, v), col = c0 + 16 * c + li;
            if (q < nQ && col < nPS) {
                phase_out_row(phi, phi1, qOut, q, nPS)[col] = acc[c][v];
                if (qOut2 && qOut2[q] >= 0) phi1[(size_t)qOut2[q] * nPS + col] = acc[c][v];
            }
        }
    }
}

// The line of patch r takes the metric f of its last probe and runs to its next probe,
// whose step goes to its images' qS (s = -stp lambda) with qLine = its line row; a line
// that has ended (or an idle patch) marks its images -2, which every kernel of a round
// skips, and clears active[r].
__device__ void ls_next(LsState* __restrict__ st, int r, double f, int* __restrict__ qLine, double* __restrict__ qS,
                        int* __restrict__ active, const int* __restrict__ slot, int nImg) {
    LsState S = st[r];
    int on = 0;
    double stp = 0.0;
    if (S.pc >= 0) on = ls_advance(S, f, &stp);
    S.nProbe += on;
    st[r] = S;
    active[r] = on;
    const double sv = -stp * S.lambda;
    for (int j = 0; j < nImg; ++j) {
        qLine[r * nImg + j] = on ? slot[r] * nImg + j : -2;
        qS[r * nImg + j] = on ? sv : 0.0;
    }
}

// New lines: state idx[i] starts at metric fa with lambda (its probes count from 0) and
// takes its first step, the probe at 1e-21 / lambda
__global__ void k_ls_init(LsState* __restrict__ st, const int* __restrict__ idx, const double* __restrict__ fl, int n,
                          int* __restrict__ qLine, double* __restrict__ qS, int* __restrict__ active,
                          const int* __restrict__ slot, int nImg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    LsState S = {};
    S.pc = 0;
    S.fa0 = fl[2 * i];
    S.lambda = fl[2 * i + 1];
    st[idx[i]] = S;
    ls_next(st, idx[i], 0.0, qLine, qS, active, slot, nImg);
}

// Phases of line-search probes without the mode GEMM (Solver::my_precalc / metricAt,
// solver.cpp:246-327): phi[q] = base[row] + s * dir[row] for the line rows i < nL (q = lq[i],
// row = lrow[i], the slot's line: base phases at the line's start, dir = B dAlpha).
template <typename T>
__global__ void k_line_phase(T* __restrict__ phi, const T* __restrict__ line, const int* __restrict__ lq,
                             const int* __restrict__ lrow, const double* __restrict__ qS, int nL, int nPS,
                             size_t dirOff, const int* __restrict__ qLine = nullptr) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nL * nPS) return;
    const int i = (int)(e / nPS), k = (int)(e % nPS), q = lq[i];
    if (qLine && qLine[q] < 0) return;          // an image outside this round's probes
    const size_t src = (size_t)lrow[i] * nPS + k;
    phi[(size_t)q * nPS + k] = line[src] + (T)qS[q] * line[dirOff + src];
}

// phiOut[q'] = phi[base(q')] + step * B[obj][mode(q')]   (finite-difference probes)
template <typename T>
__global__ void k_phase_pert(T* __restrict__ out, const T* __restrict__ phi, const T* __restrict__ B,
                             const int* __restrict__ base, const int* __restrict__ mode, const int* __restrict__ obj,
                             const T* __restrict__ step, int nQ, int K, int nPS) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nQ * nPS) return;
    const int q = e / nPS, i = e % nPS;
    out[e] = phi[(size_t)base[q] * nPS + i] + step[q] * B[((size_t)obj[q] * K + mode[q]) * nPS + i];
}

// tile[q][e] = scale[q] * pupil * exp(i phi) where inv[e] names the pupil pixel (or -1)
template <typename T>
__global__ void k_field(cx<T>* __restrict__ tiles, const T* __restrict__ phi, const int* __restrict__ inv,
                        const T* __restrict__ pupil, const int* __restrict__ qImg, const T* __restrict__ imgScale,
                        const int* __restrict__ qMap, int nQ, int N2, int nPS) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nQ * N2) return;
    const int qt = e / N2, p = e % N2;
    const int q = qMap ? qMap[qt] : qt;      // tile index -> virtual image
    const int idx = inv[p];
    cx<T> v = {T(0), T(0)};
    if (idx >= 0) {
        T s, c;
        sincos_t(phi[(size_t)q * nPS + idx], &s, &c);
        const T mag = pupil[idx] * (imgScale ? imgScale[qImg[q]] : T(1));
        v = {mag * c, mag * s};
    }
    tiles[e] = v;
}

template <typename T>
__global__ void k_power(cx<T>* __restrict__ tiles, size_t n) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    cx<T> v = tiles[e];
    tiles[e] = {v.x * v.x + v.y * v.y, T(0)};
}

// S[q][u] = tile[q][osNat[u]] * scale
template <typename T>
__global__ void k_extract(cx<T>* __restrict__ S, const cx<T>* __restrict__ tiles, const int* __restrict__ osNat,
                          int nQ, int nOS, int N2, T scale) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nQ * nOS) return;
    const int q = e / nOS, u = e % nOS;
    cx<T> v = tiles[(size_t)q * N2 + osNat[u]];
    S[e] = {v.x * scale, v.y * scale};
}

// Block reduction of a double (deterministic tree), blockDim = 256
__device__ double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// The metric of request r from its chunks' partial sums, L = sum_o w_o (sum_chunks
// partial) / s^2, in a fixed order; ld(i) reads partial i
template <typename LD>
__device__ __forceinline__ double metric_of(int r, const double* __restrict__ w, int nObj, int nChunk, double invS2,
                                            LD ld) {
    double sum = 0;
    for (int o = 0; o < nObj; ++o) {
        double m = 0;
        for (int c = 0; c < nChunk; ++c) m += ld(((size_t)r * nObj + o) * nChunk + c);
        if (w[o] != 0.0) sum += w[o] * (m * invS2);
    }
    return sum;
}

// P, Q on the stored half of the OTF support for (request r, object o, chunk of 256
// points) plus the chunk's metric partial sum(wt (ftSum - |P|^2/Q)), wt = 2 for a point
// standing for itself and its mirror.  Images are summed per
// channel in groups of 30 (Solver::metric task split, solver.cpp:756-797).
template <typename T>
__global__ void __launch_bounds__(256) k_pq(double* __restrict__ Pout, double* __restrict__ Qout,
                                            double* __restrict__ partial, const cx<T>* __restrict__ S,
                                            const cx<T>* __restrict__ D, const double* __restrict__ ftSum,
                                            const double* __restrict__ gamma, const int* __restrict__ reqSlot,
                                            const int* __restrict__ chObj, const int* __restrict__ chFirst,
                                            const int* __restrict__ chCount, const double* __restrict__ wt,
                                            const T* __restrict__ imgScale, int nCh, int nObj, int nImg, int nOS,
                                            int useGamma, const int* __restrict__ active = nullptr) {
    __shared__ double sh[256];
    const int r = blockIdx.z, o = blockIdx.y;
    if (active && !active[r]) return;            // a request not probed this round (the whole block)
    const int u = blockIdx.x * 256 + threadIdx.x;
    const int slot = reqSlot[r];
    double pr = 0, pi = 0, q = 0, part = 0;
    if (u < nOS) {
        q = useGamma ? gamma[(size_t)slot * nObj + o] : 0.0;
        for (int c = 0; c < nCh; ++c) {
            if (chObj[c] != o) continue;
            const int f0 = chFirst[c], n = chCount[c];
            for (int b = 0; b < n; b += 30) {
                const int e = min(b + 30, n);
                double tr = 0, ti = 0, tq = 0;
                // four images' S and D loaded together, then summed in image order; an
                // image the patch lacks (scale 0) is neither read nor summed.  Four keeps the
                // kernel at 58 registers, which fit beside a running k_otf_wave workgroup
                constexpr int G = 4;
                for (int j0 = f0 + b; j0 < f0 + e; j0 += G) {
                    cx<T> sv[G], dv[G];
                    bool on[G];
#pragma unroll
                    for (int i = 0; i < G; ++i) {
                        const int j = j0 + i;
                        on[i] = j < f0 + e && imgScale[(size_t)slot * nImg + j] != T(0);
                        sv[i] = on[i] ? S[((size_t)r * nImg + j) * nOS + u] : cx<T>{T(0), T(0)};
                        dv[i] = on[i] ? D[((size_t)slot * nImg + j) * nOS + u] : cx<T>{T(0), T(0)};
                    }
#pragma unroll
                    for (int i = 0; i < G; ++i) {
                        if (!on[i]) continue;
                        const double sx = sv[i].x, sy = sv[i].y, dx = dv[i].x, dy = dv[i].y;
                        tq += sx * sx + sy * sy;
                        tr += dx * sx + dy * sy;      // conj(d) * s
                        ti += dx * sy - dy * sx;
                    }
                }
                q += tq;
                pr += tr;
                pi += ti;
            }
        }
        const size_t k = ((size_t)r * nObj + o) * nOS + u;
        if (Pout) {             // a line-search round needs the metric partials only
            Pout[2 * k] = pr;
            Pout[2 * k + 1] = pi;
            Qout[k] = q;
        }
        part = wt[u] * (ftSum[((size_t)slot * nObj + o) * nOS + u] - (pr * pr + pi * pi) / q);
    }
    const double s = block_sum(part, sh);
    if (threadIdx.x == 0) partial[((size_t)r * nObj + o) * gridDim.x + blockIdx.x] = s;
}

// L[r] = sum_o w_o * (sum_chunks partial) / s^2
__global__ void k_metric(double* __restrict__ L, const double* __restrict__ partial, const double* __restrict__ w,
                         int nReq, int nObj, int nChunk, double invS2) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nReq) return;
    L[r] = metric_of(r, w, nObj, nChunk, invS2, [&](size_t i) { return partial[i]; });
}

// The end of a device line-search round, one 64-thread workgroup per patch: each probed
// patch's metric (the sum k_metric forms, from the partials staged in LDS by the whole
// workgroup) and its line moved to the next probe (ls_next on a copy of the state staged
// the same way); patches not probed return at once
__global__ void __launch_bounds__(64) k_ls_metric(LsState* __restrict__ st, double* __restrict__ L,
                                                  const double* __restrict__ partial, const double* __restrict__ w,
                                                  int nObj, int nChunk, double invS2, int* __restrict__ qLine,
                                                  double* __restrict__ qS, int* __restrict__ active,
                                                  const int* __restrict__ slot, int nImg) {
    constexpr int kMaxPart = 512;
    constexpr int kWords = (int)(sizeof(LsState) / sizeof(int));
    static_assert(sizeof(LsState) % sizeof(int) == 0, "LsState copies as ints");
    __shared__ double part[kMaxPart];
    __shared__ LsState S;
    const int r = blockIdx.x, np = nObj * nChunk;
    if (!active[r]) return;
    const bool staged = np <= kMaxPart;
    if (staged)
        for (int i = threadIdx.x; i < np; i += 64) part[i] = partial[(size_t)r * np + i];
    int* sw = reinterpret_cast<int*>(&S);
    const int* gw = reinterpret_cast<const int*>(st + r);
    for (int i = threadIdx.x; i < kWords; i += 64) sw[i] = gw[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        const double f = staged ? metric_of(r, w, nObj, nChunk, invS2, [&](size_t i) { return part[i - (size_t)r * np]; })
                                : metric_of(r, w, nObj, nChunk, invS2, [&](size_t i) { return partial[i]; });
        L[r] = f;
        ls_next(&S, 0, f, qLine + (size_t)r * nImg, qS + (size_t)r * nImg, active + r, slot + r, nImg);
    }
    __syncthreads();
    int* ow = reinterpret_cast<int*>(st + r);
    for (int i = threadIdx.x; i < kWords; i += 64) ow[i] = sw[i];
}

// G' (de-centred): tile[q][e] = (PQ D - PS S)/QS on the support (the mirror half as
// the conjugate of the stored half), else 0
template <typename T>
__global__ void k_build_g(cx<T>* __restrict__ tiles, const cx<T>* __restrict__ S, const cx<T>* __restrict__ D,
                          const double* __restrict__ P, const double* __restrict__ Q, const int* __restrict__ invOs,
                          QMap qm, const int* __restrict__ qReq, const int* __restrict__ imgObj,
                          const int* __restrict__ gIdx, int nQ, int N2, int nOS, int nImg, int nObj) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nQ * N2) return;
    const int q = gIdx[e / N2], p = e % N2;
    const int code = invOs[p];                 // h: stored point, -2-h: mirror of h, -1: off the support
    const int u = code >= 0 ? code : -2 - code;
    cx<T> v = {T(0), T(0)};
    if (code != -1) {
        const int img = qm.img[q], r = qReq[q], o = imgObj[img];
        const size_t k = ((size_t)r * nObj + o) * nOS + u;
        const double pr = P[2 * k], pi = P[2 * k + 1], qq = Q[k];
        const double pqr = pr * qq, pqi = pi * qq, ps = pr * pr + pi * pi, qs = qq * qq;
        const cx<T> d = D[((size_t)qm.slot[q] * nImg + img) * nOS + u];
        const cx<T> s = S[((size_t)r * nImg + img) * nOS + u];
        const double gr = (pqr * d.x - pqi * d.y) - ps * s.x;
        const double gi = (pqr * d.y + pqi * d.x) - ps * s.y;
        v = {(T)(gr / qs), (T)((code >= 0 ? gi : -gi) / qs)};   // G(-u) = G(u)*
    }
    tiles[e] = v;
}

template <typename T>
__global__ void k_save_re(T* __restrict__ out, const cx<T>* __restrict__ tiles, size_t n, T scale) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = tiles[e].x * scale;
}

template <typename T>
__global__ void k_mul_re(cx<T>* __restrict__ tiles, const T* __restrict__ re, size_t n, T scale) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const T f = re[e] * scale;
    cx<T> v = tiles[e];
    tiles[e] = {v.x * f, v.y * f};
}

// v[q][i] = Im(conj(PF_i) t[pde_i]) * (pupil_i for the reference form, 1 for the adjoint)
template <typename T>
__global__ void k_extract_v(T* __restrict__ v, const cx<T>* __restrict__ tiles, const T* __restrict__ phi,
                            const int* __restrict__ pde, const T* __restrict__ pupil, const int* __restrict__ gIdx,
                            int nQ, int nPS, int N2, int extraPupil) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nQ * nPS) return;
    const int q = e / nPS, i = e % nPS;
    T s, c;
    sincos_t(phi[(size_t)gIdx[q] * nPS + i], &s, &c);
    const T pu = pupil[i];
    const cx<T> t = tiles[(size_t)q * N2 + pde[i]];
    const T pfr = pu * c, pfi = pu * s;
    const T im = pfr * t.y - pfi * t.x;     // Im(conj(pf) t)
    v[e] = extraPupil ? im * pu : im;
}

// g[gIdx[r]][m] = scale[r] * sum_i v[r][i] * B[obj][m][i] for m < nEnabled, 0 otherwise
// (Solver::gradient's mode projection, solver.cpp:860-894), as a GEMM on the matrix
// cores in two passes.  k_grad_mfma: workgroup (16 gradient images, pixel slice s)
// computes all modes in tiles of 64; the pixels are cut into runs of GR_RUN, run c
// belongs to slice c % GR_SLICES and, inside the workgroup, to wave (c / GR_SLICES) % 4.
// Lane (i, k) of the 16x16x4 MFMA holds pixels [run + k*GR_RUN/4, +GR_RUN/4) of row /
// mode i, so each k-step feeds 4 pixels and each row is read in contiguous runs.
// Partial sums leave the MFMA accumulators for FP64 after every run; the 4 waves'
// partials are added in a fixed order and written to part[s][r][m].  k_grad_sum adds
// the slices in a fixed order and applies the scale.
constexpr int GR_SLICES = 8;
template <typename T>
struct GradRun { static constexpr int value = 32; };
template <>
struct GradRun<double> { static constexpr int value = 16; };

template <typename T>
__global__ void __launch_bounds__(256) k_grad_mfma(double* __restrict__ part, const T* __restrict__ v,
                                                   const T* __restrict__ B, QMap qm,
                                                   const int* __restrict__ imgObj, const int* __restrict__ gIdx, int nG,
                                                   int K, int Kt, int nPS, int nEnabled) {
    using V = typename MfmaVec<T>::v4;
    constexpr int RUN = GradRun<T>::value, PL = RUN / 4;      // pixels per run / per lane
    __shared__ double red[4][16][65];
    __shared__ int sobj[16];
    const int g0 = blockIdx.x * 16, slice = blockIdx.y;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    if (threadIdx.x < 16)
        sobj[threadIdx.x] = (g0 + (int)threadIdx.x < nG) ? imgObj[qm.img[gIdx[g0 + threadIdx.x]]] : INT_MAX;
    __syncthreads();
    const T* vr = v + (size_t)min(g0 + li, nG - 1) * nPS;
    const int nRun = (nPS + RUN - 1) / RUN;
    for (int mt = 0; mt < Kt; mt += 64) {
        double accd[4][4];
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int e = 0; e < 4; ++e) accd[c][e] = 0.0;
        int o = INT_MAX;
        for (int r = 0; r < 16; ++r) o = min(o, sobj[r]);
        while (o != INT_MAX) {
            const T* Bo = B + (size_t)o * K * nPS;
            const bool mine = sobj[li] == o;
            for (int run = slice + GR_SLICES * w; run < nRun; run += 4 * GR_SLICES) {
                const int px0 = run * RUN + lk * PL;
                T a[PL], b[4][PL];
#pragma unroll
                for (int t = 0; t < PL; ++t) {
                    const int px = px0 + t;
                    a[t] = (mine && px < nPS) ? vr[px] : T(0);
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int m = mt + 16 * c + li;
#pragma unroll
                    for (int t = 0; t < PL; ++t) {
                        const int px = px0 + t;
                        b[c][t] = (m < nEnabled && px < nPS) ? Bo[(size_t)m * nPS + px] : T(0);
                    }
                }
                V acc[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[c] = V{0, 0, 0, 0};
#pragma unroll
                for (int t = 0; t < PL; ++t)
#pragma unroll
                    for (int c = 0; c < 4; ++c) acc[c] = MfmaVec<T>::mac(a[t], b[c][t], acc[c]);
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int e = 0; e < 4; ++e) accd[c][e] += (double)acc[c][e];
            }
            int next = INT_MAX;
            for (int r = 0; r < 16; ++r)
                if (sobj[r] > o) next = min(next, sobj[r]);
            o = next;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int e = 0; e < 4; ++e) red[w][MfmaVec<T>::row(lk, e)][16 * c + li] = accd[c][e];
        __syncthreads();
        for (int t = threadIdx.x; t < 16 * 64; t += 256) {
            const int r = t / 64, mm = t % 64, m = mt + mm;
            if (g0 + r < nG && m < Kt)
                part[((size_t)slice * nG + g0 + r) * Kt + m] =
                    ((red[0][r][mm] + red[1][r][mm]) + red[2][r][mm]) + red[3][r][mm];
        }
        __syncthreads();
    }
}

// Row r of g is gradient image r (the order of gIdx); nEnabled[r] is the enabled-mode
// count of its request (each request has its own; part holds the columns of the
// batch's largest count)
__global__ void k_grad_sum(double* __restrict__ g, const double* __restrict__ part, const double* __restrict__ scale,
                           const int* __restrict__ nEnabled, int nG, int K, int Kt) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nG * K) return;
    const int r = e / K, m = e % K;
    const bool on = m < nEnabled[r];
    double sum = 0.0;
    if (on)
        for (int sl = 0; sl < GR_SLICES; ++sl) sum += part[((size_t)sl * nG + r) * Kt + m];
    g[(size_t)r * K + m] = on ? scale[r] * sum : 0.0;
}

// dl = -sum_u (q dn - dq |p|^2) / (q (q + dq)) / N^2 for one perturbed image
// (SubImage::metricChange, subimage.cpp:261-278)
template <typename T>
__global__ void __launch_bounds__(256) k_metric_change(double* __restrict__ out, const cx<T>* __restrict__ Snew,
                                                       const cx<T>* __restrict__ Sold, const cx<T>* __restrict__ D,
                                                       const double* __restrict__ P, const double* __restrict__ Q,
                                                       const int* __restrict__ baseQ, QMap qm,
                                                       const int* __restrict__ qReq, const int* __restrict__ imgObj,
                                                       const double* __restrict__ wt, int nOS, int nImg, int nObj,
                                                       double invN2) {
    __shared__ double sh[256];
    const int q = blockIdx.x;          // perturbed probe
    const int b = baseQ[q];            // the unperturbed image
    const int img = qm.img[b], r = qReq[b], o = imgObj[img];
    double dl = 0;
    for (int u = threadIdx.x; u < nOS; u += 256) {
        const cx<T> sn = Snew[(size_t)q * nOS + u];
        const cx<T> so = Sold[((size_t)r * nImg + img) * nOS + u];
        const cx<T> d = D[((size_t)qm.slot[b] * nImg + img) * nOS + u];
        const size_t k = ((size_t)r * nObj + o) * nOS + u;
        const double pr = P[2 * k], pi = P[2 * k + 1], qq = Q[k];
        const double dsr = (double)sn.x - (double)so.x, dsi = (double)sn.y - (double)so.y;
        const double dpr = d.x * dsr + d.y * dsi, dpi = d.x * dsi - d.y * dsr;   // conj(d) * ds
        const double dq = 2.0 * (so.x * dsr + so.y * dsi) + (dsr * dsr + dsi * dsi);
        const double dn = 2.0 * (dpr * pr + dpi * pi) + (dpr * dpr + dpi * dpi);
        dl -= wt[u] * ((qq * dn - dq * (pr * pr + pi * pi)) / (qq * (qq + dq)));
    }
    const double s = block_sum(dl, sh);
    if (threadIdx.x == 0) out[q] = s * invN2;
}

// ------------------------------------------------------- preprocessing ----
// SubImage::getWindowedImg + padding of SubImage::initialize (subimage.cpp:106-209):
// cut s x s at the shifted start, (x - mean) * window + mean, pad to N x N with the
// windowed mean.  One block per image.
template <typename T>
__global__ void __launch_bounds__(256) k_window(cx<T>* __restrict__ tiles, const float* __restrict__ raw,
                                                const double* __restrict__ window, const int* __restrict__ slot,
                                                const int* __restrict__ img, const int* __restrict__ sx,
                                                const int* __restrict__ sy, int nImg, int dataY, int dataX,
                                                int startY, int startX, int s, int N) {
    __shared__ double sh[256];
    const int q = blockIdx.x;
    const float* blk = raw + ((size_t)slot[q] * nImg + img[q]) * dataY * dataX;
    const int y0 = startY + sy[q], x0 = startX + sx[q];
    double acc = 0;
    for (int e = threadIdx.x; e < s * s; e += 256) acc += (double)blk[(size_t)(y0 + e / s) * dataX + x0 + e % s];
    const double avg = block_sum(acc, sh) / (double)(s * s);
    acc = 0;
    for (int e = threadIdx.x; e < s * s; e += 256) {
        const double v = ((double)blk[(size_t)(y0 + e / s) * dataX + x0 + e % s] - avg) * window[e] + avg;
        acc += v;
    }
    const double mean = block_sum(acc, sh) / (double)(s * s);
    const int off = (N - s) / 2;
    cx<T>* t = tiles + (size_t)q * N * N;
    for (int e = threadIdx.x; e < N * N; e += 256) {
        const int y = e / N - off, x = e % N - off;
        double v = mean;
        if (y >= 0 && y < s && x >= 0 && x < s)
            v = ((double)blk[(size_t)(y0 + y) * dataX + x0 + x] - avg) * window[y * s + x] + avg;
        t[e] = {(T)v, T(0)};
    }
}

// ----------------------------------------- optional restoration outputs ----
// tiles[q][p] = OTF of image img[q] at natural index p: the stored value (invOs code h),
// the conjugate of the stored mirror (-2-h), or 0 off the support -- the de-centred OTF
// that SubImage::getPSF and FourierTransform::convolve transform (subimage.cpp:615-634,
// fouriertransform.hpp:309-324)
template <typename T>
__global__ void k_otf_scatter(cx<T>* __restrict__ tiles, const cx<T>* __restrict__ S, const int* __restrict__ invOs,
                              const int* __restrict__ img, int nQ, int nOS, int N2) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nQ * N2) return;
    const int q = e / N2, code = invOs[e % N2];
    cx<T> v = {T(0), T(0)};
    if (code != -1) {
        v = S[(size_t)img[q] * nOS + (code >= 0 ? code : -2 - code)];
        if (code < 0) v.y = -v.y;
    }
    tiles[e] = v;
}

// tiles[q][p] = (F[obj[q]][p] * nrm) * tiles[q][p]: the normalised transform of the
// object times the OTF (FourierTransform::convolve's product)
template <typename T>
__global__ void k_mul_obj(cx<T>* __restrict__ tiles, const cx<T>* __restrict__ F, const int* __restrict__ obj,
                          int nQ, int N2, T nrm) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nQ * N2) return;
    const int q = e / N2;
    const cx<T> f = F[(size_t)obj[q] * N2 + e % N2];
    const T fr = f.x * nrm, fi = f.y * nrm;
    const cx<T> t = tiles[e];
    tiles[e] = {fr * t.x - fi * t.y, fr * t.y + fi * t.x};
}

// ---- the restored objects after the sums (Object::restorePatch, object.cpp:298-331) ----
// R restorations (patch, object) at once on the N x N grid in the centred layout (DC at
// (N/2, N/2)); code[k] of a grid point: 2h for the stored support point h, 2h + 1 for the
// mirror of h (P(-u) = P(u)*), -1 off the support.  Every step keeps the reference's
// operation order (no contraction), the connected region aside, which is a set.

// ao = conj(P) / Q where Q != 0 (safeDivide), dn = Q, and the filter's first power
// pt = |ao|^2 Q (ScharmerFilter, utils.cpp:341-343)
__global__ void k_rs_build(cx<double>* __restrict__ ao, double* __restrict__ dn, double* __restrict__ pt,
                           const double* __restrict__ P, const double* __restrict__ Q, const int* __restrict__ code,
                           int R, int nOS, int N2) {
#pragma clang fp contract(off)
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)R * N2) return;
    const int r = (int)(e / N2), c = code[e % N2];
    cx<double> a{0.0, 0.0};
    double q = 0.0;
    if (c >= 0) {
        const size_t i = (size_t)r * nOS + (c >> 1);
        const double pr = P[2 * i], pi = (c & 1) ? -P[2 * i + 1] : P[2 * i + 1];
        q = Q[i];
        if (fabs(q) > 0) a = {pr / q, -pi / q};
    }
    ao[e] = a;
    dn[e] = q;
    pt[e] = (a.x * a.x + a.y * a.y) * q;
}

// image::smooth (utils.cpp:304-328): the mean over the half-open window [y - n, y + n) x
// [x - n, x + n) clipped to the grid, summed row by row as the reference sums it; with nf
// the result v becomes nf[r] / v (the noise-to-power ratio, utils.cpp:344-345)
__global__ void k_rs_box(double* __restrict__ out, const double* __restrict__ in, const double* __restrict__ nf, int R,
                         int N, int n) {
#pragma clang fp contract(off)
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N2 = (size_t)N * N;
    if (e >= (size_t)R * N2) return;
    const int r = (int)(e / N2), k = (int)(e % N2), y = k / N, x = k % N;
    const double* g = in + (size_t)r * N2;
    const int yl = max(y - n, 0), yh = min(y + n, N), xl = max(x - n, 0), xh = min(x + n, N);
    double acc = 0;
    for (int yy = yl; yy < yh; ++yy)
        for (int xx = xl; xx < xh; ++xx) acc += g[(size_t)yy * N + xx];
    double v = acc / ((yh - yl) * (xh - xl));
    if (nf) v = nf[r] / v;
    out[e] = v;
}

// The filter weights (utils.cpp:346-365): 1 - (v(-u) + v(u)) / 2, clipped at 0, set to 0
// below 0.2 and to at most 1, the centre 1
__global__ void k_rs_weights(double* __restrict__ rr, const double* __restrict__ v, int R, int N) {
#pragma clang fp contract(off)
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N2 = (size_t)N * N;
    if (e >= (size_t)R * N2) return;
    const int k = (int)(e % N2), y = k / N, x = k % N;
    const double* g = v + (e - k);
    double w = 1.0 - 0.5 * (g[(size_t)((N - y) % N) * N + (N - x) % N] + g[k]);
    w = w < 0.0 ? 0.0 : w;
    if (w < 0.2) w = 0.0;
    if (w > 1.0) w = 1.0;
    if (y == N / 2 && x == N / 2) w = 1.0;
    rr[e] = w;
}

// connectedRegion from the centre (utils.cpp:49-55, 294-300): the 4-connected set of
// points with a positive weight that holds the centre, one workgroup per restoration,
// marks in LDS (0 outside, 1 a candidate, 2 in the region) grown in place until a sweep
// adds nothing (each sweep either adds a point or ends, so at most N^2 + 1 sweeps); the
// weights outside the region are multiplied by 0.
template <int N>
__global__ void __launch_bounds__(1024) k_rs_region(double* __restrict__ out, const double* __restrict__ rr) {
    constexpr int N2 = N * N;
    __shared__ unsigned char m[N2];
    __shared__ int changed;
    const double* w = rr + (size_t)blockIdx.x * N2;
    for (int e = threadIdx.x; e < N2; e += 1024) m[e] = w[e] > 0.0 ? 1 : 0;
    __syncthreads();
    if (threadIdx.x == 0) m[(N / 2) * N + N / 2] = 2;
    while (true) {
        __syncthreads();
        if (threadIdx.x == 0) changed = 0;
        __syncthreads();
        int ch = 0;
        for (int e = threadIdx.x; e < N2; e += 1024) {
            if (m[e] != 1) continue;
            const int y = e / N, x = e % N;
            if ((y + 1 < N && m[e + N] == 2) || (y > 0 && m[e - N] == 2) || (x + 1 < N && m[e + 1] == 2) ||
                (x > 0 && m[e - 1] == 2)) {
                m[e] = 2;
                ch = 1;
            }
        }
        if (ch) changed = 1;
        __syncthreads();
        if (!changed) break;
    }
    double* o = out + (size_t)blockIdx.x * N2;
    for (int e = threadIdx.x; e < N2; e += 1024) o[e] = w[e] * (double)(m[e] == 2 ? 1 : 0);
}

// data = ao rs inside the cut-off radius, 0 outside (utils.cpp:372-389), or ao unfiltered,
// then the quadrant swap to the natural layout (reorder) for the inverse transform
__global__ void k_rs_apply(cx<double>* __restrict__ t, const cx<double>* __restrict__ ao, const double* __restrict__ rs,
                           const double* __restrict__ c2, int R, int N, int nObj) {
#pragma clang fp contract(off)
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N2 = (size_t)N * N;
    if (e >= (size_t)R * N2) return;
    const int r = (int)(e / N2), k = (int)(e % N2), y = k / N, x = k % N, H = N / 2;
    cx<double> a = ao[e];
    if (rs) {
        if ((double)((y - H) * (y - H) + (x - H) * (x - H)) > c2[r % nObj]) a = {0.0, 0.0};
        else a = {a.x * rs[e], a.y * rs[e]};
    }
    t[(size_t)r * N2 + (size_t)((y + H) % N) * N + (x + H) % N] = a;
}

// the real part after the inverse transform, normalised: the centred s x s crop
// (object.cpp:329-331, 393-397), or the whole grid as the float od.img holds it (the real
// part rounded to float before the normalisation)
__global__ void k_rs_crop(float* __restrict__ out, const cx<double>* __restrict__ t, int R, int N, int s, int full) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N2 = (size_t)N * N;
    if (full) {
        if (e >= (size_t)R * N2) return;
        out[e] = (float)((double)(float)t[e].x * (1.0 / (double)N2));
        return;
    }
    const size_t s2 = (size_t)s * s;
    if (e >= (size_t)R * s2) return;
    const int r = (int)(e / s2), k = (int)(e % s2), off = (N - s) / 2;
    out[e] = (float)(t[(size_t)r * N2 + (size_t)(k / s + off) * N + k % s + off].x / ((double)N * N));
}

// out[q][y][x] = Re tiles[q][(y + off + sh) % N][(x + off + sh) % N] * scale: the centred
// s x s of each tile, re-centred by sh = N/2 for a PSF (the OTF's REORDER_IMG flag)
template <typename T>
__global__ void k_crop_re(double* __restrict__ out, const cx<T>* __restrict__ tiles, int nQ, int N, int s, int sh,
                          double scale) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t s2 = (size_t)s * s;
    if (e >= (size_t)nQ * s2) return;
    const int q = e / s2, r = e % s2, off = (N - s) / 2;
    const int y = (r / s + off + sh) % N, x = (r % s + off + sh) % N;
    out[e] = (double)tiles[((size_t)q * N + y) * N + x].x * scale;
}

// Optional outputs as floats in the caller's layout, image q of patch r = tile[q] / nObj
// at row r nImg + img[q]: the row's value, or with b the float value minus b (a residual)
__global__ void k_out_rows(float* __restrict__ out, const double* __restrict__ a, const double* __restrict__ b,
                           const int* __restrict__ tile, const int* __restrict__ img, int nI, int nObj, int nImg,
                           int s2) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nI * s2) return;
    const int q = (int)(e / s2), k = (int)(e % s2), r = tile[q] / nObj;
    const double v = a[e];
    out[((size_t)r * nImg + img[q]) * s2 + k] = b ? (float)((double)(float)v - b[e]) : (float)v;
}

// The PSFs averaged per object (Object::restorePatch's PSF_AVG, object.cpp:347-364): the
// patch's images of the object summed in image order in double, rounded to float, then
// divided by their count; 0 for an object without images
__global__ void k_psf_avg(float* __restrict__ out, const double* __restrict__ psf, const int* __restrict__ first,
                          const int* __restrict__ img, const int* __restrict__ imgObj, int m, int nObj, int s2) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)m * nObj * s2) return;
    const int k = (int)(e % s2), ro = (int)(e / s2), r = ro / nObj, o = ro % nObj;
    double acc = 0;
    int cnt = 0;
    for (int q = first[r]; q < first[r + 1]; ++q)
        if (imgObj[img[q]] == o) {
            acc += psf[(size_t)q * s2 + k];
            ++cnt;
        }
    out[e] = cnt ? (float)((double)(float)acc * (1.0 / cnt)) : 0.0f;
}

// D[slot][img][u] = tile[q][osNat[u]]
template <typename T>
__global__ void k_store_d(cx<T>* __restrict__ D, const cx<T>* __restrict__ tiles, const int* __restrict__ osNat,
                          const int* __restrict__ slot, const int* __restrict__ img, int nQ, int nOS, int nImg, int N2) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nQ * nOS) return;
    const int q = e / nOS, u = e % nOS;
    D[((size_t)slot[q] * nImg + img[q]) * nOS + u] = tiles[(size_t)q * N2 + osNat[u]];
}

// ftSum[slot][o][u] = sum over the object's images of |D|^2 (Object::addToFT, object.cpp:471-478)
template <typename T>
__global__ void k_ftsum(double* __restrict__ ftSum, const cx<T>* __restrict__ D, const int* __restrict__ slotObj,
                        const int* __restrict__ imgObj, int nPairs, int nImg, int nObj, int nOS) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nPairs * nOS) return;
    const int k = e / nOS, u = e % nOS;
    const int slot = slotObj[2 * k], o = slotObj[2 * k + 1];
    double acc = 0;
    for (int j = 0; j < nImg; ++j) {
        if (imgObj[j] != o) continue;
        const cx<T> d = D[((size_t)slot * nImg + j) * nOS + u];
        acc = acc + ((double)d.x * d.x + (double)d.y * d.y);
    }
    ftSum[((size_t)slot * nObj + o) * nOS + u] = acc;
}

// Per-frame statistics and noise level of SubImage::initialize with doReset
// (subimage.cpp:155-187): windowed image stats, then FourierTransform::noise
// (fouriertransform.cpp:448-495, mask = s/6+1, limit = s^2/4) of
// (img - mean) * noiseWindow.  The s x s DFT is evaluated directly (any s <= 128,
// the patch side need not be a power of two), one workgroup per frame, FP64.
__global__ void __launch_bounds__(256) k_noise(double* __restrict__ noise, double* __restrict__ stddev,
                                               const float* __restrict__ raw, const double* __restrict__ window,
                                               const double* __restrict__ nwindow, const int* __restrict__ slot,
                                               const int* __restrict__ img, const int* __restrict__ sx,
                                               const int* __restrict__ sy, int nImg, int dataY, int dataX,
                                               int startY, int startX, int s) {
    extern __shared__ double nsh[];
    double* X = nsh;                      // s*s real
    double* twr = X + s * s;              // cos(2 pi k/s)
    double* twi = twr + s;                // -sin(2 pi k/s)
    double* Zr = twi + s;                 // two columns of Z = X W
    double* Zi = Zr + 2 * s;
    __shared__ double red[256];
    const int q = blockIdx.x;
    const float* blk = raw + ((size_t)slot[q] * nImg + img[q]) * dataY * dataX;
    const int y0 = startY + sy[q], x0 = startX + sx[q];
    const int s2 = s * s;
    double acc = 0;
    for (int e = threadIdx.x; e < s2; e += 256) {
        const double v = blk[(size_t)(y0 + e / s) * dataX + x0 + e % s];
        X[e] = v;
        acc += v;
    }
    for (int k = threadIdx.x; k < s; k += 256) {
        double sn, cs;
        sincospi(2.0 * k / s, &sn, &cs);
        twr[k] = cs;
        twi[k] = -sn;
    }
    const double avg = block_sum(acc, red) / s2;
    acc = 0;
    for (int e = threadIdx.x; e < s2; e += 256) {
        const double v = (X[e] - avg) * window[e] + avg;
        X[e] = v;
        acc += v;
    }
    const double mean = block_sum(acc, red) / s2;
    acc = 0;
    for (int e = threadIdx.x; e < s2; e += 256) acc += X[e] * X[e];
    const double rms = block_sum(acc, red) / s2;
    for (int e = threadIdx.x; e < s2; e += 256) X[e] = (X[e] - mean) * nwindow[e];
    __syncthreads();
    const int mask = s / 6 + 1;
    const double limits = s2 / 4.0;
    double np = 0;
    long cnt = 0;
    for (int v0 = 0; v0 < s; v0 += 2) {               // column frequencies v0, v0+1
        for (int t = threadIdx.x; t < 2 * s; t += 256) {
            const int c = t / s, y = t % s, v = v0 + c;
            double zr = 0, zi = 0;
            if (v < s)
                for (int x = 0; x < s; ++x) {
                    const int k = (x * v) % s;
                    zr += X[y * s + x] * twr[k];
                    zi += X[y * s + x] * twi[k];
                }
            Zr[c * s + y] = zr;
            Zi[c * s + y] = zi;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < 2 * s; t += 256) {
            const int c = t / s, u = t % s, v = v0 + c;
            if (v >= s || v < mask || v >= s - mask + 1 || u < mask || u >= s - mask + 1) continue;
            const double xx = (double)min(v, s - v) * min(v, s - v);
            const double yy = min(u, s - u);
            if (xx + yy * yy < limits) continue;
            double fr = 0, fi = 0;
            for (int y = 0; y < s; ++y) {
                const int k = (u * y) % s;
                const double ar = twr[k], ai = twi[k], br = Zr[c * s + y], bi = Zi[c * s + y];
                fr += ar * br - ai * bi;
                fi += ar * bi + ai * br;
            }
            np += fr * fr + fi * fi;
            cnt++;
        }
        __syncthreads();
    }
    const double tot = block_sum(np, red);
    const double n = block_sum((double)cnt, red);
    if (threadIdx.x == 0) {
        noise[q] = n > 0 ? sqrt(tot / (n * s2)) : tot;
        stddev[q] = sqrt(rms - mean * mean);
    }
}

// --------------------------------------------------------- fused (LDS) ----
// One workgroup per image with the whole N x N tile resident in LDS (FP32, N <= 128:
// 128 x 129 complex floats = 132 KB of gfx950's 160 KB).  Rows are padded by one
// element so row and column passes both walk consecutive banks (MAP 1).
// First Stockham pass (NS = 1) of lines whose non-zero elements are known: the
// middle [N/4, 3N/4) (ENDS = 0) or the two ends [0, N/4) and [3N/4, N) (ENDS = 1).
// Element j + r N/R lies in the middle exactly for r in [R/4, 3R/4), so only the
// non-zero elements are read and the rest of the butterfly input is the constant zero.
template <typename T, int N, int R, int NL, int NT, int MAP, int DIR, int ENDS>
__device__ __forceinline__ void spass_first_pruned(cx<T>* buf, int ls, int es) {
    constexpr int NB = NL * (N / R);
    constexpr int BT = (NB + NT - 1) / NT;
    cx<T> v[BT][R];
#pragma unroll
    for (int i = 0; i < BT; ++i) {
        const int b = threadIdx.x + i * NT;
        if (NB % NT == 0 || b < NB) {
            const int line = MAP ? b % NL : b / (N / R), j = MAP ? b / NL : b % (N / R);
            const cx<T>* L = buf + line * ls;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const bool mid = r >= R / 4 && r < 3 * R / 4;
                v[i][r] = (mid != (ENDS != 0)) ? L[(j + r * (N / R)) * es] : cx<T>{T(0), T(0)};
            }
            dft_reg<T, R, DIR>(v[i]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < BT; ++i) {
        const int b = threadIdx.x + i * NT;
        if (NB % NT == 0 || b < NB) {
            const int line = MAP ? b % NL : b / (N / R), j = MAP ? b / NL : b % (N / R);
            cx<T>* L = buf + line * ls;
#pragma unroll
            for (int r = 0; r < R; ++r) L[(j * R + r) * es] = v[i][r];
        }
    }
    __syncthreads();
}

template <typename T, int N, int R, int NL, int NT, int MAP>
__device__ __forceinline__ void spass_first_mid(cx<T>* buf, int ls, int es) {
    spass_first_pruned<T, N, R, NL, NT, MAP, -1, 0>(buf, ls, es);
}

// Last forward Stockham pass (NS = N/R) of NL lines that writes, instead of the
// transform F, the packed power spectrum pk[line][y] = |F(y)|^2 + i |F(y + N/2)|^2
// for y < N/2 (both outputs of one butterfly: y = j + r NS and y + N/2 = j + (r + R/2) NS).
template <typename T, int N, int R, int NL, int NT, int MAP>
__device__ __forceinline__ void spass_last_power(cx<T>* buf, int ls, int es, const cx<T>* __restrict__ tw,
                                                 cx<T>* pk, int pls, int pes) {
    constexpr int NS = N / R;
    constexpr int NB = NL * (N / R);
    constexpr int BT = (NB + NT - 1) / NT;
    cx<T> v[BT][R];
#pragma unroll
    for (int i = 0; i < BT; ++i) {
        const int b = threadIdx.x + i * NT;
        if (NB % NT == 0 || b < NB) {
            const int line = MAP ? b % NL : b / (N / R), j = MAP ? b / NL : b % (N / R);
            const cx<T>* L = buf + line * ls;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                cx<T> x = L[(j + r * (N / R)) * es];
                if (r > 0) x = cmul(x, tw[r * j * (N / (NS * R))]);
                v[i][r] = x;
            }
            dft_reg<T, R, -1>(v[i]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < BT; ++i) {
        const int b = threadIdx.x + i * NT;
        if (NB % NT == 0 || b < NB) {
            const int line = MAP ? b % NL : b / (N / R), j = MAP ? b / NL : b % (N / R);
            cx<T>* L = pk + line * pls;
#pragma unroll
            for (int r = 0; r < R / 2; ++r) {
                const cx<T> a = v[i][r], c = v[i][r + R / 2];
                L[(j + r * NS) * pes] = {a.x * a.x + a.y * a.y, c.x * c.x + c.y * c.y};
            }
        }
    }
    __syncthreads();
}

template <int N>
struct Tile {
    static constexpr int LD = N + 1;                 // padded row stride (complex elements)
    static constexpr int SIZE = N * LD;
    __device__ static int at(int y, int x) { return y * LD + x; }
};

template <typename T, int N, int DIR, int NT>
__device__ __forceinline__ void tile_fft2(cx<T>* t, const cx<T>* __restrict__ tw, int rowFirst, int nRows) {
    // rows [rowFirst, rowFirst + nRows) (others known zero), then all columns
    if (nRows == N) fft_lines<T, N, DIR, N, NT, 1>(t, Tile<N>::LD, 1, tw);
    else fft_lines<T, N, DIR, N / 2, NT, 1>(t + rowFirst * Tile<N>::LD, Tile<N>::LD, 1, tw);
    fft_lines<T, N, DIR, N, NT, 1>(t, 1, Tile<N>::LD, tw);
}

// S[q][h] = reorder(IFFT(|FFT(a)|^2))/N^2 on the stored half of the OTF support,
// a = scale * pupil * exp(i phi) with the pupil centred at (N/2, N/2) (pupilInOTF,
// pupil.cpp:261), so a is non-zero only in rows and columns [N/4, 3N/4).
//   1. write the P x P pupil region (zeros off the pupil support); nothing else is read
//   2. forward row FFTs of the N/2 pupil rows, first pass pruned to the pupil columns
//   3. forward column FFTs, first pass pruned to the pupil rows; the last pass writes
//      the real |F|^2 packed two rows per complex row: pk[y] = |F|^2(y) + i |F|^2(y+N/2)
//   4. inverse FFTs of the N/2 packed rows; each splits into the half spectra of its
//      two real rows: A(k) = (Z(k) + Z(-k)*)/2, B(k) = (Z(k) - Z(-k)*)/2i, k <= N/2
//   5. inverse column FFTs of the N/2+1 half-spectrum columns; the 2-D result is
//      Hermitian, so these columns hold every stored support point (hread: index into
//      the half array, -1-index for a point read as the conjugate of its mirror).
// The per-image body of k_otf_fused: S of image q from its phases `ph` (this thread's
// pupil points); a separate function, so that the persistent loop around it does not keep
// loop-invariant loads in registers
template <int IP>
struct PhaseSet {               // one thread's pupil-point phases of an image, by value
    float v[IP];
};

template <int N, int NT>
__device__ __noinline__ void otf_image(cx<float>* __restrict__ S, PhaseSet<(N / 2 * N / 2 + NT - 1) / NT> ph,
                                       const int* __restrict__ invPs,
                                       const float* __restrict__ pupil, float sc, const int* __restrict__ hread,
                                       const cx<float>* tw, cx<float>* t, cx<float>* hs, int nOS, int q) {
    constexpr int P = N / 2, LD = Tile<N>::LD, HC = N / 2 + 1;
    constexpr int RF = N == 64 ? 8 : 16;                     // radix of the first pass (fft_lines)
    constexpr int RL = N == 256 ? 16 : 8;                    // radix of the last pass
    constexpr int IP = (P * P + NT - 1) / NT;
#pragma unroll
    for (int i = 0; i < IP; ++i) {
        const int e = threadIdx.x + i * NT;
        if (e >= P * P) continue;
        const int k = invPs[e];
        cx<float> val = {0.f, 0.f};
        if (k >= 0) {
            float sn, cs;
            sincosf(ph.v[i], &sn, &cs);
            const float m = pupil[k] * sc;
            val = {m * cs, m * sn};
        }
        t[(e / P + N / 4) * LD + e % P + N / 4] = val;
    }
    __syncthreads();
    // 2. rows [N/4, 3N/4)
    spass_first_mid<float, N, RF, N / 2, NT, 1>(t + (N / 4) * LD, LD, 1);
    spass<float, N, RL, RF, -1, N / 2, NT, 1>(t + (N / 4) * LD, LD, 1, tw);
    // 3. all columns; the packed power spectrum goes to rows [0, N/2) of the tile
    spass_first_mid<float, N, RF, N, NT, 1>(t, 1, LD);
    spass_last_power<float, N, RL, N, NT, 1>(t, 1, LD, tw, t, 1, LD);
    // 4. inverse row FFTs of the packed rows
    fft_lines<float, N, +1, N / 2, NT, 1>(t, LD, 1, tw);
    // 5. inverse column FFTs of the N/2+1 half-spectrum columns; the first pass forms its
    //    inputs, the half spectra A_y(k) = (Z_y(k) + Z_y*(-k))/2 (rows y < N/2) and
    //    B_y(k) = (Z_y(k) - Z_y*(-k))/2i (rows y + N/2), from the packed rows as it reads them
    {
        constexpr int NB = HC * (N / RF), BT = (NB + NT - 1) / NT;
        cx<float> v[BT][RF];
#pragma unroll
        for (int i = 0; i < BT; ++i) {
            const int b = threadIdx.x + i * NT;
            if (NB % NT == 0 || b < NB) {
                const int k = b % HC, j = b / HC;
#pragma unroll
                for (int r = 0; r < RF; ++r) {
                    const int y = j + r * (N / RF), yy = y % (N / 2);
                    const cx<float> z = t[yy * LD + k], zm = t[yy * LD + (N - k) % N];
                    v[i][r] = y < N / 2 ? cx<float>{0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y)}
                                        : cx<float>{0.5f * (z.y + zm.y), -0.5f * (z.x - zm.x)};
                }
                dft_reg<float, RF, +1>(v[i]);
            }
        }
#pragma unroll
        for (int i = 0; i < BT; ++i) {
            const int b = threadIdx.x + i * NT;
            if (NB % NT == 0 || b < NB) {
                const int k = b % HC, j = b / HC;
#pragma unroll
                for (int r = 0; r < RF; ++r) hs[(j * RF + r) * HC + k] = v[i][r];
            }
        }
        __syncthreads();
    }
    spass<float, N, RL, RF, +1, HC, NT, 1>(hs, 1, HC, tw);
    constexpr float inv = 1.0f / (N * N);
    cx<float>* Sq = S + (size_t)q * nOS;
    for (int u = threadIdx.x; u < nOS; u += NT) {
        const int code = hread[u];
        const cx<float> v = hs[code >= 0 ? code : -1 - code];
        Sq[u] = {v.x * inv, (code >= 0 ? v.y : -v.y) * inv};
    }
}

// With line tables (qLine[q] >= 0: a line-search probe, Solver::metricAt, solver.cpp:294-327)
// image q's phases are line[row] + s line[dirOff + row] (row = qLine[q], s = qLineS[q]),
// the slot's line base and direction, read in place of phi.
template <int N, int NT>
__global__ void __launch_bounds__(NT) k_otf_fused(cx<float>* __restrict__ S, const float* __restrict__ phi,
                                                  const int* __restrict__ invPs, const float* __restrict__ pupil,
                                                  const int* __restrict__ qImg, const float* __restrict__ imgScale,
                                                  const int* __restrict__ hread, const cx<float>* __restrict__ twg,
                                                  int nPS, int nOS, int nQ, const float* __restrict__ line = nullptr,
                                                  const int* __restrict__ qLine = nullptr,
                                                  const double* __restrict__ qLineS = nullptr, size_t dirOff = 0) {
    constexpr int P = N / 2, LD = Tile<N>::LD, HC = N / 2 + 1;
    constexpr int IP = (P * P + NT - 1) / NT;                // pupil-grid points per thread
    extern __shared__ cx<float> lds_dyn[];
    cx<float>* t = lds_dyn;                                  // N x LD tile (natural layout)
    cx<float>* hs = lds_dyn + (N / 2) * LD;                  // N x HC half spectra, element (y, k) at y*HC + k
    cx<float>* tw = hs + N * HC;                             // twiddles
    for (int e = threadIdx.x; e < N; e += NT) tw[e] = twg[e];
    // persistent: the next image's phases load while the current one is transformed (a line
    // probe's base and direction separately, combined when the image starts)
    float ph[IP], pd[IP];
    auto fetch = [&](int qq) {
        const int lr = (qq < nQ && qLine) ? qLine[qq] : -1;
#pragma unroll
        for (int i = 0; i < IP; ++i) {
            const int e = threadIdx.x + i * NT;
            const int k = e < P * P ? invPs[e] : -1;
            if (lr < 0) {                        // -1: phases from phi; -2: an image not probed
                ph[i] = (qq < nQ && k >= 0 && lr == -1) ? phi[(size_t)qq * nPS + k] : 0.f;
                pd[i] = 0.f;
            } else {
                const size_t src = (size_t)lr * nPS + k;
                ph[i] = k >= 0 ? line[src] : 0.f;
                pd[i] = k >= 0 ? line[dirOff + src] : 0.f;
            }
        }
    };
    fetch(blockIdx.x);
    for (int q = blockIdx.x; q < nQ; q += gridDim.x) {
        if (qLine && qLine[q] == -2) {           // not probed this round: no S
            fetch(q + gridDim.x);
            continue;
        }
        const float sc = imgScale[qImg[q]];
        PhaseSet<IP> cur;
        if (qLine && qLine[q] >= 0) {
            const float ls = (float)qLineS[q];
#pragma unroll
            for (int i = 0; i < IP; ++i) cur.v[i] = ph[i] + ls * pd[i];
        } else {
#pragma unroll
            for (int i = 0; i < IP; ++i) cur.v[i] = ph[i];
        }
        fetch(q + gridDim.x);
        if (sc == 0.f) {        // an image the patch lacks: its OTF is zero
            for (int u = threadIdx.x; u < nOS; u += NT) S[(size_t)q * nOS + u] = {0.f, 0.f};
            continue;
        }
        otf_image<N, NT>(S, cur, invPs, pupil, sc, hread, tw, t, hs, nOS, q);
        __syncthreads();        // the next image's field overwrites the half spectra
    }
}

// ------------------------------------------------ wave-local FP32 OTF (N = 128) ----
// The transform of k_otf_fused with every 128-point line done by 8 lanes of one wave:
// pass 1 is a 16-point DFT of the stride-8 subsequence l (lane l of the line), pass 2 two
// 8-point DFTs of its twiddled outputs k1 = m and m + 8 (lane m): the butterflies,
// twiddles and order of the two Stockham passes of fft_lines.  The pass-1 outputs go
// through slots in the line's own LDS storage, which only lanes of the same wave touch, so
// the two passes of a line need the wave's own LDS order and no workgroup barrier.
// Waves 0-7 ("row waves") and 8-15 ("column waves") share an image's work, four barriers
// per image:
//   A    (row waves) the next image's phases into LDS (cur); forward transforms of the 64
//        pupil rows, pupil row y (grid rows [32, 96)) in LDS row y + 32, its field there
//   D    (column waves, beside A) the previous image's inverse transforms of the
//        half-spectrum columns k in [1, 63], their inputs A_y(k) = (Z_y(k) + Z_y*(-k))/2,
//        B_y(k) = (Z_y(k) - Z_y*(-k))/2i formed as they are read, and of the real columns 0
//        and 64 as the one complex line C0 + i C64, split afterwards through a side area;
//        line k's pass-1 slots and results are the two Z columns k and 128 - k it alone reads
//   out  (all) S of the previous image: each stored support point read from where D left it
//        (hsrc), consecutive lanes writing consecutive points
//   B    (all) forward column transforms, the first pass reading the pupil rows only, and
//        the packed power spectrum pk[y] = |F|^2(y) + i |F|^2(y + 64) in rows [0, 64)
//   C    (row waves) inverse transforms of the packed rows: Z_y in rows [0, 64);
//        (column waves) the field a = scale pupil exp(i phi) of the next image from cur into
//        rows [64, 128)
// D of one image (rows [0, 64) and the side area) runs beside A of the next (rows
// [64, 128)).  Rows are 130 elements apart and the slots are permuted so that every LDS
// access is free of bank conflicts.  The row waves store nothing to global memory, so their
// phase loads never wait behind stores (one in-order counter tracks a wave's loads and
// stores).  On config-2 patches a launch takes 18 % less time than k_otf_fused's 17
// barrier-separated passes per image.
namespace wl {
constexpr int N = 128, LD = 130;
constexpr int SIDE = N * LD;              // the split of the combined line of columns 0 and 64
constexpr int TW = SIDE + N;              // twiddles of pass 2, W^(r k1) at 16 r + k1 (r < 8, k1 < 16)
constexpr int CUR = TW + N;               // the next image's phases on the P x P pupil grid (floats)
constexpr int SIZE = CUR + N * N / 8;     // complex floats of LDS
// slot of the pass-1 output k1 of lane l of a line.  LDS serves a wave's 64-bit accesses
// sixteen lanes at a time over 32 four-byte banks, so the sixteen elements must differ
// modulo 16.  Rows (lane s + 8 l of the wave holds lane l of row s; a quarter-wave is rows
// 0-7 at two neighbouring l) sit 2 (mod 16) apart and take the parity of (l + k1) from the
// slot; columns (lane l + 8 s holds lane l of column s; a quarter-wave is l = 0-7 of two
// neighbouring columns) take 2 (l + k1 mod 8) from it
__device__ __forceinline__ int slot(int l, int k1) { return 8 * k1 + ((l + k1) & 7); }
__device__ __forceinline__ int slot_row(int l, int k1) { return slot(l, k1); }
__device__ __forceinline__ int slot_col(int l, int k1) { return slot(l, k1); }
// a value the compiler cannot see through: the lane indices are passed through it at the
// start of every phase, so the LDS addresses derived from them are computed in the phase
// that uses them instead of being hoisted out of the image loop and held in registers
__device__ __forceinline__ int opq(int x) {
    asm volatile("" : "+v"(x));
    return x;
}
// a wave's LDS operations run in order on the hardware; this keeps the compiler from
// moving one lane's accesses across another lane's
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// pass 2 of lane m: o[p][k2] = X(k1 + 16 k2), k1 = m + 8 p, from the pass-1 outputs
// Y_r(k1) at at(r, k1), r = 0..7 (spass with R = 8, NS = 16); tw2 holds W^(r k1) at
// 16 r + k1, so that lanes of consecutive k1 read consecutive twiddles (the N-entry table
// W^e read at e = r k1 put up to 8 of them on one bank)
template <int DIR, typename AT>
__device__ __forceinline__ void pass2(cx<float> (&o)[2][8], int m, AT at, const cx<float>* tw2) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const int k1 = m + 8 * p;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            cx<float> x = at(r, p, k1);
            if (r > 0) {
                cx<float> w = tw2[16 * r + k1];
                if (DIR > 0) w.y = -w.y;
                x = cmul(x, w);
            }
            o[p][r] = x;
        }
        dft_reg<float, 8, DIR>(o[p]);
    }
}
// one row (lane l = rb of row base): pass 1 from v into the slots, then (row_pass2) pass 2,
// natural order back into the row
template <int DIR>
__device__ __forceinline__ void row_pass1(cx<float>* row, int rb, cx<float> (&v)[16]) {
    dft_reg<float, 16, DIR>(v);
    wave_sync();
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) row[slot_row(rb, k1)] = v[k1];
}
template <int DIR>
__device__ __forceinline__ void row_pass2(cx<float>* row, int rb, const cx<float>* tw) {
    wave_sync();
    cx<float> o[2][8];
    pass2<DIR>(o, rb, [&](int r, int, int k1) { return row[slot_row(r, k1)]; }, tw);
    wave_sync();
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int k2 = 0; k2 < 8; ++k2) row[rb + 8 * p + 16 * k2] = o[p][k2];
}
template <int DIR>
__device__ __forceinline__ void row_line(cx<float>* row, int rb, cx<float> (&v)[16], const cx<float>* tw) {
    row_pass1<DIR>(row, rb, v);
    row_pass2<DIR>(row, rb, tw);
}
}  // namespace wl

// At most 112 registers per lane (amdgpu_num_vgpr counts half the unified VGPR/AGPR file on
// gfx950, so 56 here): the 16 waves of the persistent workgroup then leave 64 registers per
// lane of each SIMD, where a workgroup of k_pq (58) from another evaluation lane runs beside
// the transforms; the unconstrained allocation (113) leaves 32 and k_pq waits for the grid
template <int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_num_vgpr(56))) k_otf_wave(cx<float>* __restrict__ S, const float* __restrict__ phi,
                                                 const float* __restrict__ pupilG,
                                                 const int* __restrict__ qImg, const float* __restrict__ imgScale,
                                                 const int* __restrict__ hsrc, const int* __restrict__ psd,
                                                 const cx<float>* __restrict__ twg,
                                                 int nPS, int nOS, int nQ, const float* __restrict__ line = nullptr,
                                                 const int* __restrict__ qLine = nullptr,
                                                 const double* __restrict__ qLineS = nullptr, size_t dirOff = 0) {
    static_assert(NT == 1024, "16 waves of 8 lines");
    constexpr int N = wl::N, LD = wl::LD, P = N / 2;
    extern __shared__ cx<float> lds_dyn[];
    cx<float>* t = lds_dyn;
    cx<float>* tw = lds_dyn + wl::TW;
    for (int e = threadIdx.x; e < N; e += NT) tw[e] = twg[(e >> 4) * (e & 15)];     // W^(r k1), e = 16 r + k1
    const bool rowWave = threadIdx.x < NT / 2;
    // images with a transform: probed (qLine != -2) and present (scale != 0)
    auto real = [&](int qq) { return !(qLine && qLine[qq] == -2) && imgScale[qImg[qq]] != 0.f; };
    auto next = [&](int qq) {
        do qq += gridDim.x;
        while (qq < nQ && !real(qq));
        return qq;
    };
    for (int q = blockIdx.x; q < nQ; q += gridDim.x)        // an image the patch lacks: zero OTF
        if (!(qLine && qLine[q] == -2) && imgScale[qImg[q]] == 0.f)
            for (int u = threadIdx.x; u < nOS; u += NT) S[(size_t)q * nOS + u] = {0.f, 0.f};
    // The field of one image ahead: the row waves load its phases on the pupil support (point
    // k = threadIdx.x + 512 i) at the start of A (fetch) and put them (a line probe's base +
    // s direction) on the pupil grid in LDS at its end (stage: cur; points off the support stay
    // 0); in C the column waves make a = scale pupil exp(i phi) of grid point e = j + 512 i
    // (j = threadIdx.x - 512, i = 0..7) in LDS row 64 + e / 64, column 32 + e % 64 (field).
    float* cur = reinterpret_cast<float*>(lds_dyn + wl::CUR);
    for (int e = threadIdx.x; e < P * P; e += NT) cur[e] = 0.f;
    constexpr int NI = 7;                        // support points per lane: nPS <= 3584
    float ph[NI], pd[NI];
    int pe[NI];
    int fq = nQ;                                 // the image fetched
    auto fetch = [&](int qq) {
        fq = qq;
        if (!rowWave || qq >= nQ) return;
        const int lr = qLine ? qLine[qq] : -1;
        const float* src = lr < 0 ? phi + (size_t)qq * nPS : line + (size_t)lr * nPS;
        const int j = wl::opq(threadIdx.x);
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int k = j + NT / 2 * i, kc = k < nPS ? k : 0;
            pe[i] = k < nPS ? psd[kc] : -1;
            ph[i] = src[kc];
            pd[i] = lr < 0 ? 0.f : src[dirOff + kc];
        }
    };
    auto stage = [&]() {
        if (!rowWave || fq >= nQ) return;
        const bool probe = qLine && qLine[fq] >= 0;
        const float ls = probe ? (float)qLineS[fq] : 0.f;
#pragma unroll
        for (int i = 0; i < NI; ++i)
            if (pe[i] >= 0) cur[pe[i]] = probe ? ph[i] + ls * pd[i] : ph[i];
    };
    auto field = [&](int qq) {
        if (rowWave || qq >= nQ) return;
        const float sc = imgScale[qImg[qq]];
        const int j = wl::opq(threadIdx.x - NT / 2);
        float amp[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) amp[i] = pupilG[j + NT / 2 * i];
#pragma unroll
        for (int i = 0; i < 8; ++i) {       // off the support amp = 0 and cur = 0: a = 0
            const int e = j + NT / 2 * i;
            float sn, cs;
            sincosf(cur[e], &sn, &cs);
            const float m = amp[i] * sc;
            t[(N / 2 + e / P) * LD + N / 4 + e % P] = {m * cs, m * sn};
        }
    };
    int q = blockIdx.x;
    if (q < nQ && !real(q)) q = next(q);
    fetch(q);
    __syncthreads();                             // twiddles, cur cleared
    stage();
    __syncthreads();
    field(q);
    __syncthreads();                             // the first field
    // D (column waves): inverse transform of half-spectrum line k = 8 (w - 8) + rb, lane ra of
    // the line, line 0 as C0 + i C64; the results, scaled by 1/N^2, into the line's own Z slots
    // (two of line 0 into the side area), where out finds them
    auto halfCols = [&]() {
        {
            cx<float> v[16];
            const int tid = wl::opq(threadIdx.x), w = tid >> 6, ra = tid & 7, rb = (tid >> 3) & 7;
            const int k = 8 * (w - 8) + rb, kk = k ? N - k : N / 2;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int yy = ra + 8 * (r & 7);               // row of Z; y = ra + 8 r
                const cx<float> z = t[yy * LD + k], zm = t[yy * LD + kk];
                if (k == 0)
                    v[r] = r < 8 ? cx<float>{z.x, zm.x} : cx<float>{z.y, zm.y};
                else
                    v[r] = r < 8 ? cx<float>{0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y)}
                                 : cx<float>{0.5f * (z.y + zm.y), -0.5f * (z.x - zm.x)};
            }
            dft_reg<float, 16, +1>(v);
            // slot pos: row pos of column k (pos < 64), row pos - 64 of column kk
            wl::wave_sync();
#pragma unroll
            for (int k1 = 0; k1 < 16; ++k1) {
                const int pos = wl::slot_col(ra, k1);
                t[(pos & (N / 2 - 1)) * LD + (k1 < 8 ? k : kk)] = v[k1];
            }
            wl::wave_sync();
            cx<float> o[2][8];
            wl::pass2<+1>(o, ra, [&](int r, int p, int k1) {
                return t[(wl::slot_col(r, k1) & (N / 2 - 1)) * LD + (p == 0 ? k : kk)];
            }, tw);
            constexpr float inv = 1.0f / (N * N);
            // X(y), y = ra + 8 p + 16 k2: row y of column k (y < 64), row y - 64 of column kk
            if (k == 0) {      // split X = T0 + i T64: T0(y) = (X(y) + X*(-y))/2, T64(y) = (X(y) - X*(-y))/2i
                // rows y < 64: T0(y) in column 0, T64(y) in column 64; y = 64: side[0], side[1]
                cx<float>* sd = t + wl::SIDE;
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int k2 = 0; k2 < 8; ++k2) sd[ra + 8 * p + 16 * k2] = o[p][k2];
                wl::wave_sync();
                cx<float> m0{}, m64{};
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int k2 = 0; k2 < 8; ++k2) {
                        const int y = ra + 8 * p + 16 * k2;
                        const cx<float> x = o[p][k2], xm = sd[(N - y) & (N - 1)];
                        const cx<float> a0 = {0.5f * (x.x + xm.x) * inv, 0.5f * (x.y - xm.y) * inv};
                        const cx<float> a64 = {0.5f * (x.y + xm.y) * inv, -0.5f * (x.x - xm.x) * inv};
                        if (k2 < 4) {
                            t[y * LD] = a0;
                            t[y * LD + N / 2] = a64;
                        } else if (p == 0 && k2 == 4) {    // y = 64 when ra = 0
                            m0 = a0;
                            m64 = a64;
                        }
                    }
                wl::wave_sync();
                if (ra == 0) {
                    sd[0] = m0;
                    sd[1] = m64;
                }
            } else {
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int k2 = 0; k2 < 8; ++k2) {
                        const int y = ra + 8 * p + 16 * k2;      // y < 64 exactly for k2 < 4
                        t[(y & (N / 2 - 1)) * LD + (k2 < 4 ? k : kk)] = {o[p][k2].x * inv, o[p][k2].y * inv};
                    }
            }
        }
    };
    // out (all waves): S of image qo from the LDS places hsrc names, in the order of the stored
    // points (consecutive lanes, consecutive points)
    auto out = [&](int qo) {
        cx<float>* Sq = S + (size_t)qo * nOS;
        constexpr int NU = (N * (N / 2 + 1) + NT - 1) / NT;      // nOS <= N (N/2 + 1)
        int src[NU];
#pragma unroll
        for (int i = 0; i < NU; ++i) {
            const int u = threadIdx.x + NT * i;
            src[i] = u < nOS ? hsrc[u] : -1;
        }
#pragma unroll
        for (int i = 0; i < NU; ++i)
            if (src[i] >= 0) Sq[threadIdx.x + NT * i] = t[src[i]];
    };
    int qp = -1;                                 // the image whose D and out are pending
    for (; q < nQ;) {
        const int qn = next(q);
        cx<float> v[16];
        // A: forward transform of pupil row 8 w + ra (LDS row 64 + 8 w + ra), lane rb of the row,
        // beside D of the previous image (rows [0, 64) and the side area); then out of that image
        if (rowWave) {
            fetch(qn);                           // in flight during the transform
            const int tid = wl::opq(threadIdx.x), w = tid >> 6, ra = tid & 7, rb = (tid >> 3) & 7;
            cx<float>* row = t + (N / 2 + 8 * w + ra) * LD;
#pragma unroll
            for (int r = 0; r < 16; ++r) v[r] = (r >= 4 && r < 12) ? row[rb + 8 * r] : cx<float>{0.f, 0.f};
            wl::row_line<-1>(row, rb, v, tw);
            stage();                             // cur of this image was read in the last C
        }
        __syncthreads();
        if (qp >= 0) out(qp);
        __syncthreads();
        // B: forward transform of column 8 w + rb, lane ra of the column; the power spectrum
        {
            const int tid = wl::opq(threadIdx.x), w = tid >> 6, ra = tid & 7, rb = (tid >> 3) & 7;
            cx<float>* col = t + 8 * w + rb;
#pragma unroll
            for (int r = 0; r < 16; ++r)
                v[r] = (r >= 4 && r < 12) ? col[(ra + 8 * r + N / 4) * LD] : cx<float>{0.f, 0.f};
            dft_reg<float, 16, -1>(v);
            wl::wave_sync();
#pragma unroll
            for (int k1 = 0; k1 < 16; ++k1) col[wl::slot_col(ra, k1) * LD] = v[k1];
            wl::wave_sync();
            cx<float> o[2][8];
            wl::pass2<-1>(o, ra, [&](int r, int, int k1) { return col[wl::slot_col(r, k1) * LD]; }, tw);
            wl::wave_sync();
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int k2 = 0; k2 < 4; ++k2) {
                    const cx<float> a = o[p][k2], c = o[p][k2 + 4];
                    col[(ra + 8 * p + 16 * k2) * LD] = {a.x * a.x + a.y * a.y, c.x * c.x + c.y * c.y};
                }
        }
        __syncthreads();
        // C: inverse transform of packed row 8 w + ra, lane rb of the row; the next field
        if (rowWave) {
            const int tid = wl::opq(threadIdx.x), w = tid >> 6, ra = tid & 7, rb = (tid >> 3) & 7;
            cx<float>* row = t + (8 * w + ra) * LD;
#pragma unroll
            for (int r = 0; r < 16; ++r) v[r] = row[rb + 8 * r];
            wl::row_line<+1>(row, rb, v, tw);
        } else {
            field(qn);
        }
        __syncthreads();
        if (!rowWave) halfCols();                // beside the next image's A
        qp = q;
        q = qn;
    }
    if (qp >= 0) {
        __syncthreads();
        out(qp);
    }
}

// One Stockham pass like spass, over NL lines whose element e of line l sits at
// buf[addr(l, e)] (any layout: the caller's functor).
template <typename T, int N, int R, int NS, int DIR, int NL, int NT, typename F>
__device__ __forceinline__ void spass_at(cx<T>* buf, F addr, const cx<T>* __restrict__ tw) {
    constexpr int NB = NL * (N / R);
    constexpr int BT = (NB + NT - 1) / NT;
    cx<T> v[BT][R];
#pragma unroll
    for (int i = 0; i < BT; ++i) {
        const int b = threadIdx.x + i * NT;
        if (NB % NT == 0 || b < NB) {
            const int line = b % NL, j = b / NL, k = j % NS;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                cx<T> x = buf[addr(line, j + r * (N / R))];
                if (NS > 1 && r > 0) {
                    cx<T> w = tw[r * k * (N / (NS * R))];
                    if (DIR > 0) w.y = -w.y;
                    x = cmul(x, w);
                }
                v[i][r] = x;
            }
            dft_reg<T, R, DIR>(v[i]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < BT; ++i) {
        const int b = threadIdx.x + i * NT;
        if (NB % NT == 0 || b < NB) {
            const int line = b % NL, j = b / NL, k = j % NS;
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[addr(line, base + r * NS)] = v[i][r];
        }
    }
    __syncthreads();
}

// FP64 OTF with every intermediate of an image in one P x (N+1) complex-double LDS
// buffer (P = N/2; 132 KB at N = 128, where the whole N x N tile would need 256 KB),
// plus the N twiddles.  Persistent: each workgroup loops over images q = blockIdx.x,
// + gridDim.x, ..., keeping its pupil indices and amplitudes in registers and loading
// the next image's phases while the current image is transformed.  The steps of
// k_otf_fused, arranged to work in place:
//   1-2. forward row FFTs of the P pupil rows (grid rows [N/4, 3N/4)): the first pass
//      reads only the pupil columns and forms a = scale pupil exp(i phi) on the fly
//   3. forward column FFTs (first pass pruned to the pupil rows) and the packed power
//      spectrum pk[y][x] = |F|^2(y, x) + i |F|^2(y + N/2, x), y < N/2.  A column's
//      transform has N values between its two passes but its input and output have N/2,
//      so the columns go in two halves: the inputs of both are read into registers,
//      the first-pass outputs of one half take the whole buffer (column c as buffer row
//      c), and the left half's power values wait in registers until both are done
//   4. inverse row FFTs of the packed rows: Z_y = IFFT(pk[y]) holds the half spectra
//      of the real rows y and y + N/2, A(k) = (Z(k) + Z*(-k))/2, B(k) = (Z(k) - Z*(-k))/2i
//   5. inverse FFTs over y of the half-spectrum columns.  Line k (0 < k < N/2) is stored
//      in slot k of rows y < N/2 and slot N-k of rows y - N/2; line 0 is the complex
//      combination X = C0 + i C_{N/2} of the two real columns, slots 0 and N/2.  The
//      first pass forms its inputs A_y(k), B_y(k) (line 0: A_y(0) + i A_y(N/2),
//      B_y(0) + i B_y(N/2)) from Z_y(k) and Z_y(N-k) as it reads them: the two values
//      one butterfly needs from row y are exactly that pair
//   6. the stored support points: hread codes name (y, k) of the N x (N/2+1) half
//      array; columns 0 and N/2 separate as C0 = (X(y) + X*(-y))/2,
//      C_{N/2} = (X(y) - X*(-y))/2i
//   LINE: line-search probes take their phases as base + s * direction from the line
//   buffers (qLine[q] >= 0: the line's row, qLineS[q] = s; the directions dirOff further),
//   the other images from phi, as k_line_phase would have written them
template <int N, int NT, bool LINE>
__global__ void __launch_bounds__(NT) k_otf64_fused(cx<double>* __restrict__ S, const double* __restrict__ phi,
                                                    const int* __restrict__ invPs, const double* __restrict__ pupil,
                                                    const int* __restrict__ qImg, const double* __restrict__ imgScale,
                                                    const int* __restrict__ hread, const cx<double>* __restrict__ twg,
                                                    int nPS, int nOS, int nQ, const int* __restrict__ qSkip = nullptr,
                                                    const double* __restrict__ line = nullptr,
                                                    const int* __restrict__ qLine = nullptr,
                                                    const double* __restrict__ qLineS = nullptr, size_t dirOff = 0) {
    constexpr int P = N / 2, LD = N + 1, HC = N / 2 + 1;
    constexpr int RF = N == 64 ? 8 : 16, RL = 8;             // radices of the two passes (N = RF RL)
    static_assert(RF * RL == N, "two-pass transforms only (N = 64, 128)");
    constexpr int NBR = P * (N / RF);                        // first-pass row butterflies
    constexpr int NB1 = N * (N / RF);                        // first-pass column butterflies, both halves
    constexpr int NB2 = P * (N / RL);                        // second-pass butterflies of one half
    constexpr int IR = NBR / NT, IH = NB1 / 2 / NT, I2 = NB2 / NT;
    static_assert(NBR % NT == 0 && NB1 / 2 % NT == 0 && NB2 % NT == 0, "thread count must divide the butterflies");
    extern __shared__ cx<double> lds64[];
    cx<double>* t = lds64;
    cx<double>* tw = lds64 + P * LD;
    for (int e = threadIdx.x; e < N; e += NT) tw[e] = twg[e];
    // row butterfly (line, j) reads pupil row `line`, columns j + r N/RF, r < RF/2
    int kk[IR][RF / 2];
    double amp[IR][RF / 2], ph[IR][RF / 2];
    double pd[LINE ? IR : 1][RF / 2];                        // LINE: the probe's direction
    double ls = 0.0;                                         // LINE: its step (0: phases from phi)
#pragma unroll
    for (int i = 0; i < IR; ++i) {
        const int b = threadIdx.x + i * NT, row = b % P, j = b / P;
#pragma unroll
        for (int r = 0; r < RF / 2; ++r) {
            kk[i][r] = invPs[row * P + j + r * (N / RF)];
            amp[i][r] = kk[i][r] >= 0 ? pupil[kk[i][r]] : 0.0;
        }
    }
    auto fetch = [&](int qq) {
        if constexpr (LINE) {
            const int lr = (qLine && qq < nQ) ? qLine[qq] : -1;
            const double* src = lr >= 0 ? line + (size_t)lr * nPS : phi + (size_t)qq * nPS;
            ls = lr >= 0 ? qLineS[qq] : 0.0;
#pragma unroll
            for (int i = 0; i < IR; ++i)
#pragma unroll
                for (int r = 0; r < RF / 2; ++r) {
                    ph[i][r] = (qq < nQ && kk[i][r] >= 0) ? src[kk[i][r]] : 0.0;
                    pd[i][r] = (lr >= 0 && kk[i][r] >= 0) ? src[dirOff + kk[i][r]] : 0.0;
                }
        } else {
#pragma unroll
            for (int i = 0; i < IR; ++i)
#pragma unroll
                for (int r = 0; r < RF / 2; ++r)
                    ph[i][r] = (qq < nQ && kk[i][r] >= 0) ? phi[(size_t)qq * nPS + kk[i][r]] : 0.0;
        }
    };
    // images not probed this round (device line search: qSkip -2) have no S and are passed
    // over without a fetch of their phases
    auto nextq = [&](int qq) {
        do qq += gridDim.x;
        while (qSkip && qq < nQ && qSkip[qq] == -2);
        return qq;
    };
    int q0 = blockIdx.x;
    if (qSkip && q0 < nQ && qSkip[q0] == -2) q0 = nextq(q0);
    fetch(q0);
    auto at = [](int k, int y) { return y < P ? y * LD + k : (y - P) * LD + (k == 0 ? P : N - k); };
    for (int q = q0; q < nQ; q = nextq(q)) {
        const int qn = nextq(q);                // the image fetched during this one
        const double sc = imgScale[qImg[q]];
        if (sc == 0.0) {        // an image the patch lacks: its OTF is zero (t is not touched)
            for (int u = threadIdx.x; u < nOS; u += NT) S[(size_t)q * nOS + u] = {0.0, 0.0};
            fetch(qn);
            continue;
        }
        // 1-2.
        {
            cx<double> v[IR][RF];
#pragma unroll
            for (int i = 0; i < IR; ++i) {
#pragma unroll
                for (int r = 0; r < RF; ++r) v[i][r] = {0.0, 0.0};
#pragma unroll
                for (int r = 0; r < RF / 2; ++r) {
                    double sn, cs;
                    if constexpr (LINE)          // the expression of k_line_phase: base + s * dir
                        sincos(ls != 0.0 ? ph[i][r] + ls * pd[i][r] : ph[i][r], &sn, &cs);
                    else
                        sincos(ph[i][r], &sn, &cs);
                    const double m = amp[i][r] * sc;
                    v[i][r + RF / 4] = {m * cs, m * sn};
                }
                dft_reg<double, RF, -1>(v[i]);
            }
            if constexpr (!LINE) fetch(qn);
#pragma unroll
            for (int i = 0; i < IR; ++i) {
                const int b = threadIdx.x + i * NT, line = b % P, j = b / P;
#pragma unroll
                for (int r = 0; r < RF; ++r) t[line * LD + j * RF + r] = v[i][r];
            }
        }
        __syncthreads();
        spass<double, N, RL, RF, -1, P, NT, 1>(t, LD, 1, tw);
        // 3. butterfly bb of half h (0: left, 1: right) is column bb % P + h P; butterfly j
        //    reads grid rows j + r N/RF, r in [RF/4, 3RF/4), i.e. buffer rows j + (r - RF/4) N/RF
        cx<double> x1[2][IH][RF / 2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int i = 0; i < IH; ++i) {
                const int bb = threadIdx.x + i * NT, c = bb % P + h * P, j = bb / P;
#pragma unroll
                for (int r = 0; r < RF / 2; ++r) x1[h][i][r] = t[(j + r * (N / RF)) * LD + c];
            }
        __syncthreads();
        cx<double> pw[2][I2][RL / 2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int i = 0; i < IH; ++i) {
                const int bb = threadIdx.x + i * NT, c = bb % P, j = bb / P;
                cx<double> v[RF];
#pragma unroll
                for (int r = 0; r < RF; ++r)
                    v[r] = (r >= RF / 4 && r < 3 * RF / 4) ? x1[h][i][r - RF / 4] : cx<double>{0.0, 0.0};
                dft_reg<double, RF, -1>(v);
#pragma unroll
                for (int r = 0; r < RF; ++r) t[c * LD + j * RF + r] = v[r];     // column c as buffer row c
            }
            __syncthreads();
            // second pass (NS = RF); outputs y = j + r RF and y + N/2 give one power pair
#pragma unroll
            for (int i = 0; i < I2; ++i) {
                const int b = threadIdx.x + i * NT, c = b % P, j = b / P;
                cx<double> v[RL];
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    cx<double> x = t[c * LD + j + r * RF];
                    if (r > 0) x = cmul(x, tw[r * j * (N / (RF * RL))]);
                    v[r] = x;
                }
                dft_reg<double, RL, -1>(v);
#pragma unroll
                for (int r = 0; r < RL / 2; ++r) {
                    const cx<double> a = v[r], d = v[r + RL / 2];
                    pw[h][i][r] = {a.x * a.x + a.y * a.y, d.x * d.x + d.y * d.y};
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int i = 0; i < I2; ++i) {
                const int b = threadIdx.x + i * NT, c = b % P + h * P, j = b / P;
#pragma unroll
                for (int r = 0; r < RL / 2; ++r) t[(j + r * RF) * LD + c] = pw[h][i][r];
            }
        __syncthreads();
        // 4.
        fft_lines<double, N, +1, P, NT, 1>(t, LD, 1, tw);
        // 5. first pass with the half-spectrum split on its inputs, then the second pass
        {
            constexpr int NB = P * (N / RF), BT = (NB + NT - 1) / NT;
            cx<double> v[BT][RF];
#pragma unroll
            for (int i = 0; i < BT; ++i) {
                const int b = threadIdx.x + i * NT;
                if (NB % NT == 0 || b < NB) {
                    const int k = b % P, j = b / P;
#pragma unroll
                    for (int r = 0; r < RF / 2; ++r) {
                        const cx<double>* row = t + (j + r * (N / RF)) * LD;
                        const cx<double> z = row[k], zm = row[k == 0 ? P : N - k];
                        if (k == 0) {
                            v[i][r] = {z.x, zm.x};
                            v[i][r + RF / 2] = {z.y, zm.y};
                        } else {
                            v[i][r] = {0.5 * (z.x + zm.x), 0.5 * (z.y - zm.y)};
                            v[i][r + RF / 2] = {0.5 * (z.y + zm.y), -0.5 * (z.x - zm.x)};
                        }
                    }
                    dft_reg<double, RF, +1>(v[i]);
                }
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < BT; ++i) {
                const int b = threadIdx.x + i * NT;
                if (NB % NT == 0 || b < NB) {
                    const int k = b % P, j = b / P;
#pragma unroll
                    for (int r = 0; r < RF; ++r) t[at(k, j * RF + r)] = v[i][r];
                }
            }
            __syncthreads();
        }
        spass_at<double, N, RL, RF, +1, P, NT>(t, at, tw);
        // LINE: the next image's base and direction are fetched only here, after the heavy
        // passes, so that they do not hold registers through them (223 registers in place of
        // 255: a k_pq workgroup of another lane then fits beside the two waves per SIMD;
        // 258 against 249 patches/s on 256 config-2 patches). Fetched here without LINE the
        // phases would cost the GEMM-per-probe path 1.5 %.
        if constexpr (LINE) fetch(qn);
        // 6.
        constexpr double inv = 1.0 / ((double)N * N);
        cx<double>* Sq = S + (size_t)q * nOS;
        for (int u = threadIdx.x; u < nOS; u += NT) {
            const int code = hread[u];
            const int idx = code >= 0 ? code : -1 - code, y = idx / HC, k = idx % HC;
            cx<double> v;
            if (k != 0 && k != P) {
                v = t[at(k, y)];
            } else {
                const cx<double> x = t[at(0, y)], xm = t[at(0, (N - y) % N)];
                v = k == 0 ? cx<double>{0.5 * (x.x + xm.x), 0.5 * (x.y - xm.y)}
                           : cx<double>{0.5 * (x.y + xm.y), -0.5 * (x.x - xm.x)};
            }
            Sq[u] = {v.x * inv, (code >= 0 ? v.y : -v.y) * inv};
        }
        __syncthreads();
    }
}

// FP64 frame transform D = FFT2(padded windowed frame) on the stored half of the OTF
// support, one image per workgroup, in k_otf64_fused's P x (N+1) LDS buffer (N <= 128):
// the steps of k_window (cut s x s at the shifted start, (x - avg) window + avg, pad to
// N x N with the windowed mean), then the forward transform of a real image:
//   rows y and y + N/2 packed as one complex row, forward row FFTs, the split into the
//   half spectra of the two real rows folded into the first pass of the column FFTs
//   (line layout `at` of k_otf64_fused: column 0 combined with column N/2), second pass,
//   and the stored points through hread (unnormalised, as FourierTransform's forward FFT),
//   rounded once to TD (float for the FP32 forms, whose D is otherwise the same)
template <int N, int NT, typename TD>
__global__ void __launch_bounds__(NT) k_init64_fused(cx<TD>* __restrict__ D, const float* __restrict__ raw,
                                                     const double* __restrict__ window, const int* __restrict__ slot,
                                                     const int* __restrict__ img, const int* __restrict__ sx,
                                                     const int* __restrict__ sy, const int* __restrict__ hread,
                                                     const cx<double>* __restrict__ twg, int nImg, int dataY, int dataX,
                                                     int startY, int startX, int s, int nOS) {
    constexpr int P = N / 2, LD = N + 1, HC = N / 2 + 1;
    constexpr int RF = N == 64 ? 8 : 16, RL = 8;
    extern __shared__ cx<double> lds64[];
    cx<double>* t = lds64;
    cx<double>* tw = lds64 + P * LD;
    double* red = reinterpret_cast<double*>(tw + N);          // NT doubles for the block sums
    for (int e = threadIdx.x; e < N; e += NT) tw[e] = twg[e];
    auto bsum = [&](double v) {                                 // fixed-order tree over NT threads
        red[threadIdx.x] = v;
        __syncthreads();
        for (int w = NT / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        const double r = red[0];
        __syncthreads();
        return r;
    };
    const int q = blockIdx.x;
    const float* blk = raw + ((size_t)slot[q] * nImg + img[q]) * dataY * dataX;
    const int y0 = startY + sy[q], x0 = startX + sx[q];
    double acc = 0;
    for (int e = threadIdx.x; e < s * s; e += NT) acc += (double)blk[(size_t)(y0 + e / s) * dataX + x0 + e % s];
    const double avg = bsum(acc) / (double)(s * s);
    acc = 0;
    for (int e = threadIdx.x; e < s * s; e += NT)
        acc += ((double)blk[(size_t)(y0 + e / s) * dataX + x0 + e % s] - avg) * window[e] + avg;
    const double mean = bsum(acc) / (double)(s * s);
    const int off = (N - s) / 2;
    auto val = [&](int y, int x) {
        const int yy = y - off, xx = x - off;
        return (yy >= 0 && yy < s && xx >= 0 && xx < s)
                   ? ((double)blk[(size_t)(y0 + yy) * dataX + x0 + xx] - avg) * window[yy * s + xx] + avg
                   : mean;
    };
    for (int e = threadIdx.x; e < P * N; e += NT) {
        const int y = e / N, x = e % N;
        t[y * LD + x] = {val(y, x), val(y + P, x)};
    }
    __syncthreads();
    fft_lines<double, N, -1, P, NT, 1>(t, LD, 1, tw);
    auto at = [](int k, int y) { return y < P ? y * LD + k : (y - P) * LD + (k == 0 ? P : N - k); };
    {
        constexpr int NB = P * (N / RF), BT = (NB + NT - 1) / NT;
        cx<double> v[BT][RF];
#pragma unroll
        for (int i = 0; i < BT; ++i) {
            const int b = threadIdx.x + i * NT;
            if (NB % NT == 0 || b < NB) {
                const int k = b % P, j = b / P;
#pragma unroll
                for (int r = 0; r < RF / 2; ++r) {
                    const cx<double>* row = t + (j + r * (N / RF)) * LD;
                    const cx<double> z = row[k], zm = row[k == 0 ? P : N - k];
                    if (k == 0) {
                        v[i][r] = {z.x, zm.x};
                        v[i][r + RF / 2] = {z.y, zm.y};
                    } else {
                        v[i][r] = {0.5 * (z.x + zm.x), 0.5 * (z.y - zm.y)};
                        v[i][r + RF / 2] = {0.5 * (z.y + zm.y), -0.5 * (z.x - zm.x)};
                    }
                }
                dft_reg<double, RF, -1>(v[i]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < BT; ++i) {
            const int b = threadIdx.x + i * NT;
            if (NB % NT == 0 || b < NB) {
                const int k = b % P, j = b / P;
#pragma unroll
                for (int r = 0; r < RF; ++r) t[at(k, j * RF + r)] = v[i][r];
            }
        }
        __syncthreads();
    }
    spass_at<double, N, RL, RF, -1, P, NT>(t, at, tw);
    cx<TD>* Dq = D + ((size_t)slot[q] * nImg + img[q]) * nOS;
    for (int u = threadIdx.x; u < nOS; u += NT) {
        const int code = hread[u];
        const int idx = code >= 0 ? code : -1 - code, y = idx / HC, k = idx % HC;
        cx<double> v;
        if (k != 0 && k != P) {
            v = t[at(k, y)];
        } else {
            const cx<double> x = t[at(0, y)], xm = t[at(0, (N - y) % N)];
            v = k == 0 ? cx<double>{0.5 * (x.x + xm.x), 0.5 * (x.y - xm.y)}
                       : cx<double>{0.5 * (x.y + xm.y), -0.5 * (x.x - xm.x)};
        }
        Dq[u] = {(TD)v.x, (TD)(code >= 0 ? v.y : -v.y)};
    }
}

// FP64 Vogel / adjoint weights in one persistent kernel per gradient image (N = 64,
// 128), the quantities of the unfused chain (k_build_g, IFFT2, k_save_re, k_field,
// IFFT2, k_mul_re, FFT2, k_extract_v) with every intermediate but one in
// the P x (N+1) complex-double LDS buffer of k_otf64_fused (line/slot layout `at` as
// there).  gl = Re IFFT2(G'_H)/N^2 goes through HBM (`greal`, one N x N slice per
// gradient image, written by this workgroup and read back by it once L2 has acknowledged
// the stores), since it is needed while the buffer holds the pupil field.
//   A1. G'_H = (G'(u) + G'(-u)*)/2, the Hermitian part of G' = (PQ D - PS S)/QS on the
//       support (de-centred natural layout), as lines over y: line k (0 < k < N/2) is
//       column k, line 0 is G'_H(., 0) + i G'_H(., N/2).  G' is evaluated once per
//       stored point and written where the host table gw puts it: G' at u and G'* at -u
//       for a pair, G'/2 and G'*/2 when -u lies outside the support, Re G' when u = -u
//   A2. inverse FFTs of the lines: C(n, k), line 0 holding C(n, 0) + i C(n, N/2) (real
//       columns, since G'_H is Hermitian)
//   A3. rows n and n + N/2 packed into one complex row, Z_n(k) = C(n, k) + i C(n + N/2, k),
//       extended to k > N/2 by C(n, N-k) = C(n, k)*
//   A4. inverse row FFTs: Re Z_n = N^2 gl(n, .), Im Z_n = N^2 gl(n + N/2, .)
//   B1. inverse row FFTs of the P rows of the de-centred pupil field PF' (non-zero in
//       rows and columns [0, N/4) and [3N/4, N); buffer row i is grid row i for i < N/4,
//       i + N/2 otherwise), first pass pruned to the two ends and formed from phi
//   B2. per column: inverse FFT (first pass pruned to the pupil rows) -> hj, times
//       gl / N^2, forward FFT, keeping the output rows of the pupil; the columns go in
//       two halves as in k_otf64_fused
//   B3. forward row FFTs; v_i = Im(conj(PF_i) t(pde_i)) [* pupil_i] at the pupil pixels
template <int N, int NT>
__global__ void __launch_bounds__(NT) k_grad64_fused(double* __restrict__ v, double* __restrict__ greal,
                                                     const cx<double>* __restrict__ S, const cx<double>* __restrict__ D,
                                                     const double* __restrict__ P, const double* __restrict__ Q,
                                                     const int* __restrict__ gw, QMap qm,
                                                     const int* __restrict__ qReq, const int* __restrict__ imgObj,
                                                     const int* __restrict__ gIdx, const double* __restrict__ phi,
                                                     const int* __restrict__ invPs, const double* __restrict__ pupil,
                                                     const cx<double>* __restrict__ twg, int nPS, int nOS, int nImg,
                                                     int nObj, int nG, int extraPupil) {
    constexpr int H = N / 2, LD = N + 1, Q4 = N / 4;
    constexpr int RF = N == 64 ? 8 : 16, RL = 8;
    static_assert(RF * RL == N, "two-pass transforms only (N = 64, 128)");
    constexpr int NB1 = N * (N / RF), NB2 = H * (N / RL);
    constexpr int IH = NB1 / 2 / NT, I2 = NB2 / NT;
    static_assert(NB1 / 2 % NT == 0 && NB2 % NT == 0, "thread count must divide the butterflies");
    constexpr double inv = 1.0 / ((double)N * N);
    extern __shared__ cx<double> lds64[];
    cx<double>* t = lds64;
    cx<double>* tw = lds64 + H * LD;
    cx<double>* side = tw + N;                                   // G'_H columns 0 and N/2
    for (int e = threadIdx.x; e < N; e += NT) tw[e] = twg[e];
    auto at = [](int k, int y) { return y < H ? y * LD + k : (y - H) * LD + (k == 0 ? H : N - k); };
    for (int qt = blockIdx.x; qt < nG; qt += gridDim.x) {
        const int q = gIdx[qt], img = qm.img[q], r = qReq[q], o = imgObj[img];
        const cx<double>* Dq = D + ((size_t)qm.slot[q] * nImg + img) * nOS;
        const cx<double>* Sq = S + ((size_t)r * nImg + img) * nOS;
        const double* Po = P + 2 * ((size_t)r * nObj + o) * nOS;
        const double* Qo = Q + ((size_t)r * nObj + o) * nOS;
        // A1. zero the lines and the side area, write G'_H of every stored point to its
        //     targets, then line 0 = G'_H(., 0) + i G'_H(., N/2)
        for (int e = threadIdx.x; e < H * LD; e += NT) t[e] = {0.0, 0.0};
        for (int e = threadIdx.x; e < 2 * N; e += NT) side[e] = {0.0, 0.0};
        __syncthreads();
        for (int u = threadIdx.x; u < nOS; u += NT) {
            const double pr = Po[2 * u], pi = Po[2 * u + 1], qq = Qo[u];
            const cx<double> d = Dq[u], sv = Sq[u];
            const double pqr = pr * qq, pqi = pi * qq, ps = pr * pr + pi * pi, qs = qq * qq;
            const double gr = ((pqr * d.x - pqi * d.y) - ps * sv.x) / qs;
            const double gi = ((pqr * d.y + pqi * d.x) - ps * sv.y) / qs;
#pragma unroll
            for (int w = 0; w < 2; ++w) {
                const int c = gw[w * nOS + u];
                if (c < 0) continue;
                const int f = c & 7;
                const double sc = (f & 2) ? 0.5 : 1.0;
                t[c >> 3] = {sc * gr, (f & 4) ? 0.0 : ((f & 1) ? -sc * gi : sc * gi)};
            }
        }
        __syncthreads();
        for (int y = threadIdx.x; y < N; y += NT) {
            const cx<double> a = side[y], b = side[N + y];
            t[at(0, y)] = {a.x - b.y, a.y + b.x};
        }
        __syncthreads();
        // A2.
        spass_at<double, N, RF, 1, +1, H, NT>(t, at, tw);
        spass_at<double, N, RL, RF, +1, H, NT>(t, at, tw);
        // A3. row n: slot k holds C(n, k), slot N-k holds C(n + N/2, k); slot 0 holds
        //     C(n, 0) + i C(n, N/2), slot N/2 holds C(n + N/2, 0) + i C(n + N/2, N/2)
        for (int e = threadIdx.x; e < H * H; e += NT) {
            const int n = e / H, k = e % H;
            cx<double>* row = t + n * LD;
            const cx<double> a = row[k], b = row[k == 0 ? H : N - k];
            if (k == 0) {
                row[0] = {a.x, b.x};
                row[H] = {a.y, b.y};
            } else {
                row[k] = {a.x - b.y, a.y + b.x};
                row[N - k] = {a.x + b.y, b.x - a.y};
            }
        }
        __syncthreads();
        // A4.
        fft_lines<double, N, +1, H, NT, 1>(t, LD, 1, tw);
        double* gq = greal + (size_t)qt * N * N;
        for (int e = threadIdx.x; e < H * N; e += NT) {
            const int n = e / N, x = e % N;
            const cx<double> z = t[n * LD + x];
            gq[(size_t)n * N + x] = z.x * inv;
            gq[(size_t)(n + H) * N + x] = z.y * inv;
        }
        // gl is read back only by this workgroup, from a slice that nothing else touches
        // during the launch (so no L1 line of it can be stale): its stores must have been
        // acknowledged by L2 before any wave passes the barrier.  A device-scope fence
        // would also write back the whole L2, which this does not need.
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        // B1. butterfly (row i, j) of the first pass reads columns j + m N/RF, m in the
        //     two ends [0, RF/4) and [3RF/4, RF)
        const double* ph = phi + (size_t)q * nPS;
        {
            constexpr int NB = H * (N / RF), BT = NB / NT;
            static_assert(NB % NT == 0, "");
            cx<double> w[BT][RF];
#pragma unroll
            for (int i = 0; i < BT; ++i) {
                const int b = threadIdx.x + i * NT, row = b % H, j = b / H;
                const int py = (row + Q4) % H;                      // pupil row of buffer row `row`
#pragma unroll
                for (int m = 0; m < RF; ++m) {
                    w[i][m] = {0.0, 0.0};
                    if (m < RF / 4 || m >= 3 * RF / 4) {
                        const int x = j + m * (N / RF), px = x < Q4 ? x + Q4 : x - 3 * Q4;
                        const int k = invPs[py * H + px];
                        if (k >= 0) {
                            double sn, cs;
                            sincos(ph[k], &sn, &cs);
                            w[i][m] = {pupil[k] * cs, pupil[k] * sn};
                        }
                    }
                }
                dft_reg<double, RF, +1>(w[i]);
            }
#pragma unroll
            for (int i = 0; i < BT; ++i) {
                const int b = threadIdx.x + i * NT, row = b % H, j = b / H;
#pragma unroll
                for (int m = 0; m < RF; ++m) t[row * LD + j * RF + m] = w[i][m];
            }
        }
        __syncthreads();
        spass<double, N, RL, RF, +1, H, NT, 1>(t, LD, 1, tw);
        // B2. column butterfly (c, j) of the first pass reads grid rows j + m N/RF for m in
        //     the two ends, i.e. buffer rows j + m' N/RF, m' = m for m < RF/4, m - RF/2 after
        cx<double> x1[2][IH][RF / 2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int i = 0; i < IH; ++i) {
                const int bb = threadIdx.x + i * NT, c = bb % H + h * H, j = bb / H;
#pragma unroll
                for (int m = 0; m < RF / 2; ++m) x1[h][i][m] = t[(j + m * (N / RF)) * LD + c];
            }
        __syncthreads();
        cx<double> keep[2][I2][RL / 2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            // inverse first pass: column c as buffer row c % H, element j RF + m
#pragma unroll
            for (int i = 0; i < IH; ++i) {
                const int bb = threadIdx.x + i * NT, c = bb % H, j = bb / H;
                cx<double> w[RF];
#pragma unroll
                for (int m = 0; m < RF; ++m)
                    w[m] = m < RF / 4 ? x1[h][i][m] : m >= 3 * RF / 4 ? x1[h][i][m - RF / 2] : cx<double>{0.0, 0.0};
                dft_reg<double, RF, +1>(w);
#pragma unroll
                for (int m = 0; m < RF; ++m) t[c * LD + j * RF + m] = w[m];
            }
            __syncthreads();
            // inverse second pass; output y = j + m RF times gl(y, column) / N^2.  A last
            // pass writes exactly the elements it read (j + m RF), so each butterfly
            // finishes in place without waiting for the others.
#pragma unroll
            for (int i = 0; i < I2; ++i) {
                const int b = threadIdx.x + i * NT, c = b % H, j = b / H;
                cx<double> w[RL];
                double f[RL];
#pragma unroll
                for (int m = 0; m < RL; ++m) f[m] = gq[(size_t)(j + m * RF) * N + c + h * H] * inv;
#pragma unroll
                for (int m = 0; m < RL; ++m) {
                    cx<double> x = t[c * LD + j + m * RF];
                    if (m > 0) {
                        cx<double> tv = tw[m * j * (N / (RF * RL))];
                        tv.y = -tv.y;
                        x = cmul(x, tv);
                    }
                    w[m] = x;
                }
                dft_reg<double, RL, +1>(w);
#pragma unroll
                for (int m = 0; m < RL; ++m) t[c * LD + j + m * RF] = {w[m].x * f[m], w[m].y * f[m]};
            }
            __syncthreads();
            // forward transform of the natural-order column
            spass<double, N, RF, 1, -1, H, NT, 1>(t, LD, 1, tw);
            // forward second pass, keeping outputs y = j + m RF in the two ends
#pragma unroll
            for (int i = 0; i < I2; ++i) {
                const int b = threadIdx.x + i * NT, c = b % H, j = b / H;
                cx<double> w[RL];
#pragma unroll
                for (int m = 0; m < RL; ++m) {
                    cx<double> x = t[c * LD + j + m * RF];
                    if (m > 0) x = cmul(x, tw[m * j * (N / (RF * RL))]);
                    w[m] = x;
                }
                dft_reg<double, RL, -1>(w);
#pragma unroll
                for (int m = 0; m < RL / 4; ++m) {
                    keep[h][i][m] = w[m];
                    keep[h][i][m + RL / 4] = w[m + 3 * RL / 4];
                }
            }
            __syncthreads();
        }
        // kept row y = j + m RF (ends) -> buffer row y (y < N/4) or y - N/2
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int i = 0; i < I2; ++i) {
                const int b = threadIdx.x + i * NT, c = b % H + h * H, j = b / H;
#pragma unroll
                for (int m = 0; m < RL / 4; ++m) {
                    t[(j + m * RF) * LD + c] = keep[h][i][m];
                    t[(j + (m + 3 * RL / 4) * RF - H) * LD + c] = keep[h][i][m + RL / 4];
                }
            }
        __syncthreads();
        // B3.
        fft_lines<double, N, -1, H, NT, 1>(t, LD, 1, tw);
        double* vq = v + (size_t)qt * nPS;
        for (int e = threadIdx.x; e < H * H; e += NT) {
            const int py = e / H, px = e % H;
            const int k = invPs[e];
            if (k < 0) continue;
            const int row = (py + Q4) % H, x = ((px - Q4) % N + N) % N;
            const cx<double> tv = t[row * LD + x];
            double sn, cs;
            sincos(ph[k], &sn, &cs);
            const double pu = pupil[k];
            const double im = (pu * cs) * tv.y - (pu * sn) * tv.x;   // Im(conj(PF) t)
            vq[k] = extraPupil ? im * pu : im;
        }
        __syncthreads();
    }
}

// Vogel / adjoint weights (subimage.cpp:348-390) for one image:
//   gl = IFFT(G')/N^2 with G' = (PQ D - PS S)/QS de-centred, hj = IFFT(PF')/N^2,
//   t = FFT(hj * Re gl), v_i = Im(conj(PF_i) t(pde_i)) [* pupil_i]
// Re gl is the transform of the Hermitian part G'_H = (G'(u) + G'(-u)*)/2, a real image:
// G'_H on the columns [0, N/2] (the table gwt, as for k_grad64_fused), inverse transforms
// of the N/2 lines columns 1 .. N/2-1 and G'_H(., 0) + i G'_H(., N/2) (real results), then
// rows n and n + N/2 packed as one complex row (its columns above N/2 from C(n, N-k) =
// C(n, k)*) and N/2 inverse row transforms: Re row n = N^2 gl(n, .), Im = N^2 gl(n + N/2, .)
template <int N, int NT>
__global__ void __launch_bounds__(NT) k_vogel_fused(float* __restrict__ v, const cx<float>* __restrict__ S,
                                                    const cx<float>* __restrict__ D, const double* __restrict__ P,
                                                    const double* __restrict__ Q, const float* __restrict__ phi,
                                                    const int* __restrict__ pdeY, const int* __restrict__ pdeX,
                                                    const float* __restrict__ pupil, const int* __restrict__ gwt,
                                                    QMap qm, const int* __restrict__ qReq,
                                                    const int* __restrict__ imgObj, const int* __restrict__ gIdx,
                                                    const cx<float>* __restrict__ tw,
                                                    int nPS, int nOS, int nImg, int nObj, int extraPupil) {
    extern __shared__ cx<float> lds_dyn[];
    cx<float>* t = lds_dyn;
    constexpr int PER = (N * N + NT - 1) / NT;
    const int q = gIdx[blockIdx.x];
    const int img = qm.img[q], r = qReq[q], o = imgObj[img];
    for (int e = threadIdx.x; e < Tile<N>::SIZE; e += NT) t[e] = {0.f, 0.f};
    __syncthreads();
    const cx<float>* Dq = D + ((size_t)qm.slot[q] * nImg + img) * nOS;
    const cx<float>* Sq = S + ((size_t)r * nImg + img) * nOS;
    const double* Po = P + 2 * ((size_t)r * nObj + o) * nOS;
    const double* Qo = Q + ((size_t)r * nObj + o) * nOS;
    for (int u = threadIdx.x; u < nOS; u += NT) {
        const double pr = Po[2 * u], pi = Po[2 * u + 1], qq = Qo[u];
        const double pqr = pr * qq, pqi = pi * qq, ps = pr * pr + pi * pi, qs = qq * qq;
        const cx<float> d = Dq[u], s = Sq[u];
        const double gr = (pqr * d.x - pqi * d.y) - ps * s.x;
        const double gi = (pqr * d.y + pqi * d.x) - ps * s.y;
        // de-centred position = natural index of the centred support point
        const float g_r = (float)(gr / qs), g_i = (float)(gi / qs);
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            const int c = gwt[w * nOS + u];
            if (c < 0) continue;
            const int f = c & 7;
            const float sc = (f & 2) ? 0.5f : 1.0f;
            t[c >> 3] = {sc * g_r, (f & 4) ? 0.f : ((f & 1) ? -sc * g_i : sc * g_i)};
        }
    }
    __syncthreads();
    for (int y = threadIdx.x; y < N; y += NT) {
        const cx<float> a = t[Tile<N>::at(y, 0)], b = t[Tile<N>::at(y, N / 2)];
        t[Tile<N>::at(y, 0)] = {a.x - b.y, a.y + b.x};
    }
    __syncthreads();
    fft_lines<float, N, +1, N / 2, NT, 1>(t, 1, Tile<N>::LD, tw);
    for (int e = threadIdx.x; e < (N / 2) * (N / 2); e += NT) {
        const int n = e / (N / 2), k = e % (N / 2);
        const cx<float> a = t[Tile<N>::at(n, k)], b = t[Tile<N>::at(n + N / 2, k)];
        if (k == 0) {            // a = C(n, 0) + i C(n, N/2), b the same of row n + N/2
            t[Tile<N>::at(n, 0)] = {a.x, b.x};
            t[Tile<N>::at(n, N / 2)] = {a.y, b.y};
        } else {
            t[Tile<N>::at(n, k)] = {a.x - b.y, a.y + b.x};
            t[Tile<N>::at(n, N - k)] = {a.x + b.y, b.x - a.y};
        }
    }
    __syncthreads();
    fft_lines<float, N, +1, N / 2, NT, 1>(t, Tile<N>::LD, 1, tw);
    float greal[PER];
    constexpr float inv = 1.0f / (N * N);
    static_assert(PER * NT == N * N, "point e = threadIdx.x + k NT lies in the first half of the rows for k < PER/2");
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int e = threadIdx.x + k * NT, y = e / N, x = e % N;
        greal[k] = (k < PER / 2 ? t[Tile<N>::at(y, x)].x : t[Tile<N>::at(y - N / 2, x)].y) * inv;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < Tile<N>::SIZE; e += NT) t[e] = {0.f, 0.f};
    __syncthreads();
    const float* ph = phi + (size_t)q * nPS;
    for (int i = threadIdx.x; i < nPS; i += NT) {
        float s, c;
        sincosf(ph[i], &s, &c);
        const float m = pupil[i];
        t[Tile<N>::at(pdeY[i], pdeX[i])] = {m * c, m * s};
    }
    __syncthreads();
    // the de-centred field is non-zero only in the rows and columns [0, N/4), [3N/4, N):
    // row transforms of those rows only, column transforms whose first pass reads only them
    constexpr int RF = N == 64 ? 8 : 16, RL = 8;
    auto endRow = [](int l, int e) { return (l < N / 4 ? l : l + N / 2) * Tile<N>::LD + e; };
    spass_at<float, N, RF, 1, +1, N / 2, NT>(t, endRow, tw);
    spass_at<float, N, RL, RF, +1, N / 2, NT>(t, endRow, tw);
    spass_first_pruned<float, N, RF, N, NT, 1, +1, 1>(t, 1, Tile<N>::LD);
    spass<float, N, RL, RF, +1, N, NT, 1>(t, 1, Tile<N>::LD, tw);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int e = threadIdx.x + k * NT;
        if (e < N * N) {
            const int a = Tile<N>::at(e / N, e % N);
            const float f = greal[k] * inv;
            const cx<float> h = t[a];
            t[a] = {h.x * f, h.y * f};
        }
    }
    __syncthreads();
    // forward: all rows, then only the columns of the pupil pixels
    fft_lines<float, N, -1, N, NT, 1>(t, Tile<N>::LD, 1, tw);
    auto endCol = [](int l, int e) { return e * Tile<N>::LD + (l < N / 4 ? l : l + N / 2); };
    spass_at<float, N, RF, 1, -1, N / 2, NT>(t, endCol, tw);
    spass_at<float, N, RL, RF, -1, N / 2, NT>(t, endCol, tw);
    float* vq = v + (size_t)blockIdx.x * nPS;
    for (int i = threadIdx.x; i < nPS; i += NT) {
        float s, c;
        sincosf(ph[i], &s, &c);
        const float pu = pupil[i];
        const cx<float> tv = t[Tile<N>::at(pdeY[i], pdeX[i])];
        const float im = (pu * c) * tv.y - (pu * s) * tv.x;
        vq[i] = extraPupil ? im * pu : im;
    }
}

// ------------------------------------------------------------- evaluator ----

// Stream capture (the device line search's graphs) and the calls HIP refuses while any
// stream captures -- the device-wide synchronisation before a buffer is replaced, frees --
// exclude each other across the contexts of one process (mfbd_solve_group runs one host
// thread per context)
// The streams of a context's lanes (kSyncLane + 1 of them, the evaluation lanes' first)
// are taken as one set from a per-device pool and given back when the context closes, so
// that the contexts a process opens one after the other run on the same streams.  On new
// streams, a context opened after one that had replayed captured graphs measured 11 %
// slower (FP32 480 -> 426 patches/s on 256 config-2 patches), as if two of its lanes shared
// a hardware queue; on the pooled streams it runs as the first one.  Contexts open at the
// same time (mfbd_solve_group) each get a set of their own.
struct StreamSets {
    std::mutex m;
    std::map<int, std::vector<std::array<hipStream_t, kSyncLane + 1>>> free;
};
StreamSets& stream_sets() {
    static StreamSets* p = new StreamSets;     // outlives the contexts (never destroyed)
    return *p;
}
std::mutex& capture_mutex() {
    static std::mutex m;
    return m;
}

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        if (count <= n) return;
        std::lock_guard<std::mutex> lk(capture_mutex());
        if (p) {
            HIPCHK(hipDeviceSynchronize());   // in-flight kernels may still read the old buffer
            HIPCHK(hipFree(p));
        }
        HIPCHK(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void free_() {
        if (p) {
            std::lock_guard<std::mutex> lk(capture_mutex());
            (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
    }
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { free_(); }
};

// Host staging arrays are ordinary vectors, so each upload completes before return.
template <typename T>
void upload(DBuf<T>& b, const T* h, size_t n, hipStream_t st) {
    if (n > b.n) HIPCHK(hipStreamSynchronize(st));
    b.alloc(std::max<size_t>(n, 1));
    if (n) {
        HIPCHK(hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }
}

template <typename T>
void upload(DBuf<T>& b, const std::vector<T>& h, hipStream_t st) { upload(b, h.data(), h.size(), st); }

// Pinned host staging: asynchronous copies without a per-copy synchronisation.
template <typename T>
struct HBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        if (count <= n) return;
        std::lock_guard<std::mutex> lk(capture_mutex());
        if (p) {
            HIPCHK(hipDeviceSynchronize());
            HIPCHK(hipHostFree(p));
        }
        HIPCHK(hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault));
        n = count;
    }
    void free_() {
        if (p) {
            std::lock_guard<std::mutex> lk(capture_mutex());
            (void)hipHostFree(p);
        }
        p = nullptr;
        n = 0;
    }
    HBuf() = default;
    HBuf(const HBuf&) = delete;
    HBuf& operator=(const HBuf&) = delete;
    ~HBuf() { free_(); }
};

inline unsigned nblk(size_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

template <typename T, int N>
class Pipeline final : public DeviceEvaluator {
public:
    using C = cx<T>;
    static constexpr int N2 = N * N;
    static constexpr int LINES = sizeof(T) == 8 ? 16 : 32;   // lines per FFT block (32 KB of LDS)
    // whole-tile LDS kernels: FP32 tiles up to 128 x 129 complex (132 KB)
    static constexpr bool FUSED = std::is_same_v<T, float> && N <= 128;
    static constexpr int FNT = N == 128 ? 1024 : 256;
    static constexpr size_t FLDS = (size_t)Tile<N>::SIZE * sizeof(cx<float>);
    static constexpr size_t OLDS = ((size_t)(N / 2) * Tile<N>::LD + (size_t)N * (N / 2 + 1) + N) * sizeof(cx<float>);
    static constexpr size_t WLDS = (size_t)wl::SIZE * sizeof(cx<float>);     // k_otf_wave (N = 128)
    // FP64 OTF in one kernel with an N/2 x (N+1) complex-double LDS buffer per image
    static constexpr bool FUSED64 = std::is_same_v<T, double> && N <= 128;
    static constexpr int F64NT = N == 128 ? 512 : 256;
    static constexpr size_t F64LDS = ((size_t)(N / 2) * (N + 1) + N) * sizeof(cx<double>);
    static constexpr size_t G64LDS = F64LDS + 2 * N * sizeof(cx<double>);     // plus two G'_H columns
    static constexpr size_t I64LDS = F64LDS + F64NT * sizeof(double);           // plus the block-sum scratch
    // frame transforms by k_init64_fused in every precision (N <= 128)
    static constexpr bool INIT64 = N <= 128;

    Pipeline(const Geometry& g, int maxSlots, int maxBatch, int device, bool bf16Phase = false)
        : g_(g), maxSlots_(maxSlots), maxBatch_(maxBatch), bf16Phase_(bf16Phase && std::is_same_v<T, float>) {
        device_ = device;
        HIPCHK(hipSetDevice(device));
        // one stream per lane, so the two evaluation lanes' kernels overlap on the device
        // (a latency-bound transform beside a bandwidth-bound reduction); the synchronous lane (set-up,
        // synchronous calls, restoration sums) is the pipeline's main stream.  With kernel
        // timing on, every lane uses the main stream, so each launch's duration is its own.
        timing_ = getenv("MFBD_KERNEL_TIMING") != nullptr;
        // MFBD_MAX_TILES caps the tile workspace (images per chunk of an evaluation or a
        // restoration), so that tests can run the multi-chunk paths on small problems
        if (const char* e = getenv("MFBD_MAX_TILES")) tileCap_ = std::max(1, atoi(e));
        // the evaluation lanes' streams first: the runtime spreads streams over its hardware
        // queues (4 by default) in creation order, and with the main stream created first
        // two evaluation lanes could share a queue and serialise (measured: the process's
        // first FP32 context 12 % slower than later ones)
        {
            std::array<hipStream_t, kSyncLane + 1> set{};
            bool have = false;
            {
                std::lock_guard<std::mutex> lk(stream_sets().m);
                auto& v = stream_sets().free[device];
                if (!v.empty()) {
                    set = v.back();
                    v.pop_back();
                    have = true;
                }
            }
            if (!have)
                for (auto& s : set) HIPCHK(hipStreamCreate(&s));
            for (int l = 0; l <= kSyncLane; ++l) lanes_[l].stream = set[l];
            streamSet_ = set;
        }
        if (timing_)
            for (int l = 0; l < kMaxEvalLanes; ++l) lanes_[l].stream = lanes_[kSyncLane].stream;
        st_ = lanes_[kSyncLane].stream;
        HIPCHK(hipEventCreate(&ev0_));
        HIPCHK(hipEventCreate(&ev1_));
        for (auto& l : lanes_) HIPCHK(hipEventCreateWithFlags(&l.done, hipEventDisableTiming));
        std::vector<C> tw(N);
        for (int k = 0; k < N; ++k) {
            long double a = -2.0L * 3.14159265358979323846264338327950288L * k / N;
            tw[k] = {(T)cosl(a), (T)sinl(a)};
        }
        upload(tw_, tw, st_);
        if constexpr (INIT64 && !std::is_same_v<T, double>) {     // k_init64_fused's FP64 twiddles
            std::vector<cx<double>> tw64(N);
            for (int k = 0; k < N; ++k) {
                long double a = -2.0L * 3.14159265358979323846264338327950288L * k / N;
                tw64[k] = {(double)cosl(a), (double)sinl(a)};
            }
            upload(tw64_, tw64, st_);
        }
        autotune_fft();
    }

    ~Pipeline() override {
        for (auto& l : lanes_) (void)hipStreamSynchronize(l.stream);
        ls_clear();
        for (auto& l : ls_)
            if (l.done) (void)hipEventDestroy(l.done);
        (void)hipEventDestroy(ev0_);
        (void)hipEventDestroy(ev1_);
        for (auto& l : lanes_) (void)hipEventDestroy(l.done);
        for (auto& e : evPool_) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
        for (auto& e : evPending_) { (void)hipEventDestroy(e.ev.first); (void)hipEventDestroy(e.ev.second); }
        {
            std::lock_guard<std::mutex> lk(stream_sets().m);
            stream_sets().free[device_].push_back(streamSet_);
        }
        // device and pinned buffers are released by their own destructors
    }

    int capacity() const override { return maxSlots_; }
    void bind_thread() override { HIPCHK(hipSetDevice(device_)); }
    double* alpha_staging() override { return hStage_.p; }
    int nOS() const override { return nOSfull_; }
    const std::vector<int>& support_map() const override { return fullMap_; }
    int nOSStored() const override { return nOS_; }
    int nPS() const override { return nPS_; }
    const std::vector<int>& otf_support() const override { return os_; }

    void set_basis(const double* pupil, const double* modes) override {
        use_lane(kSyncLane);
        const int P = g_.P, P2 = P * P;
        ps_.clear();
        std::vector<int> pio;
        for (int i = 0; i < P2; ++i)
            if (pupil[i] > 1e-9) {
                ps_.push_back(i);
                pio.push_back((i / P) * N + (i % P) + P / 2 + (P / 2) * N);
            }
        nPS_ = (int)ps_.size();
        std::vector<int> invPio(N2, -1), invPde(N2, -1), pde(nPS_);
        std::vector<T> pup(nPS_);
        for (int k = 0; k < nPS_; ++k) {
            invPio[pio[k]] = k;
            const int y = ps_[k] / P, x = ps_[k] % P;
            const int yd = ((y - P / 2) % N + N) % N, xd = ((x - P / 2) % N + N) % N;
            pde[k] = yd * N + xd;
            invPde[pde[k]] = k;
            pup[k] = (T)pupil[ps_[k]];
        }
        // OTF support: autocorrelation of the pupil above 1e-9 (Pupil::generateSupport)
        os_ = otf_support_of(pupil);
        nOSfull_ = (int)os_.size();
        // Hermitian half of the support: S(-u) = S(u)*, D(-u) = D(u)* (real frames), so
        // P, Q, ftSum and the metric terms of u and -u agree.  Keep one representative
        // of each pair (the smaller centred index) with weight 2, unpaired points (-u
        // outside the support, or u = -u) with weight 1.
        auto natOf = [&](int c) { return ((c / N + N / 2) % N) * N + (c % N + N / 2) % N; };
        auto mirOf = [&](int c) { return ((N - c / N) % N) * N + (N - c % N) % N; };
        std::vector<int> where(N2, -1);
        for (int u = 0; u < nOSfull_; ++u) where[os_[u]] = u;
        std::vector<int> rep;
        std::vector<double> wt;
        std::vector<int> mirNat;
        fullMap_.assign(nOSfull_, 0);
        for (int u = 0; u < nOSfull_; ++u) {
            const int c = os_[u], m = mirOf(c);
            const bool paired = where[m] >= 0 && m != c;
            // representatives sit in the natural columns [0, N/2] (the half the fused OTF
            // kernel computes); on columns 0 and N/2 the smaller natural row
            const int ny = natOf(c) / N, nx = natOf(c) % N;
            const bool isRep = (nx >= 1 && nx < N / 2) || ((nx == 0 || nx == N / 2) && ny <= (N - ny) % N);
            if (paired && !isRep) continue;                  // the mirror of a representative
            fullMap_[u] = (int)rep.size();
            if (paired) fullMap_[where[m]] = -1 - (int)rep.size();
            rep.push_back(c);
            wt.push_back(paired ? 2.0 : 1.0);
            mirNat.push_back(paired ? natOf(m) : -1);
        }
        nOS_ = (int)rep.size();
        {   // each grid point's place on the stored support, for the restoration (k_rs_build)
            std::vector<int> code(N2, -1);
            for (int u = 0; u < nOSfull_; ++u)
                code[os_[u]] = fullMap_[u] >= 0 ? 2 * fullMap_[u] : 2 * (-1 - fullMap_[u]) + 1;
            upload(rsCode_, code, st_);
        }
        for (int c : rep)       // k_cols_extract transforms only the natural columns [0, N/2]
            if (natOf(c) % N > N / 2) throw std::runtime_error("OTF support is not centro-symmetric");
        {   // where each stored point sits in the fused kernel's (N) x (N/2+1) half array
            std::vector<int> hread(nOS_);
            for (int h = 0; h < nOS_; ++h) {
                const int nat = natOf(rep[h]), ny = nat / N, nx = nat % N;
                hread[h] = nx <= N / 2 ? ny * (N / 2 + 1) + nx : -1 - (((N - ny) % N) * (N / 2 + 1) + (N - nx));
            }
            upload(hread_, hread, st_);
            if constexpr (FUSED && N == 128) {  // k_otf_wave: where each stored point's value sits in LDS
                std::vector<int> hsrc(nOS_, 0);
                otfWave_ = true;
                for (int h = 0; h < nOS_; ++h) {
                    const int c = hread[h], y = c / (N / 2 + 1), k = c % (N / 2 + 1);
                    if (c < 0) otfWave_ = false;                         // not a plain gather
                    else if (k > 0 && k < N / 2)
                        hsrc[h] = y < N / 2 ? y * wl::LD + k : (y - N / 2) * wl::LD + (N - k);
                    else if (y > N / 2) otfWave_ = false;                // columns 0, N/2: rows <= N/2 only
                    else if (y < N / 2) hsrc[h] = y * wl::LD + k;        // T0(y), T64(y)
                    else hsrc[h] = wl::SIDE + (k ? 1 : 0);               // T0(64), T64(64)
                }
                if (nPS_ > 7 * 512) otfWave_ = false;  // seven support points per lane of the row waves
                if (const char* e = std::getenv("MFBD_OTF_WAVE")) otfWave_ = otfWave_ && std::atoi(e) != 0;
                upload(hsrc_, hsrc, st_);
            }
            // where k_grad64_fused writes G'_H of stored point h (two targets, -1 for none):
            // target * 8 + flags (1: conjugate, 2: half, 4: real part), in its LDS line
            // layout, columns 0 and N/2 in a side area after the buffer and twiddles
            std::vector<int> gw(2 * (size_t)nOS_, -1);
            auto target = [&](int nat) {
                const int y = nat / N, x = nat % N, base = (N / 2) * (N + 1) + N;
                if (x > N / 2) return -1;
                if (x == 0) return base + y;
                if (x == N / 2) return base + N + y;
                return y < N / 2 ? y * (N + 1) + x : (y - N / 2) * (N + 1) + (N - x);
            };
            auto enc = [](int t, int f) { return t < 0 ? -1 : t * 8 + f; };
            for (int h = 0; h < nOS_; ++h) {
                const int pn = natOf(rep[h]), mn = ((N - pn / N) % N) * N + (N - pn % N) % N;
                if (mirNat[h] >= 0) {                       // paired: G'_H = G' at p, G'* at -p
                    gw[h] = enc(target(pn), 0);
                    gw[nOS_ + h] = enc(target(mn), 1);
                } else if (mn == pn) {                      // its own mirror: Re G'
                    gw[h] = enc(target(pn), 4);
                } else {                                    // -p outside the support
                    gw[h] = enc(target(pn), 2);
                    gw[nOS_ + h] = enc(target(mn), 3);
                }
            }
            upload(gw_, gw, st_);
            std::vector<int> invPs(P2, -1);
            for (int k = 0; k < nPS_; ++k) invPs[ps_[k]] = k;
            upload(invPs_, invPs, st_);
            upload(psd_, ps_, st_);
            if constexpr (FUSED && N == 128) {   // k_otf_wave: the pupil on its P x P grid, 0 off the support
                std::vector<float> pg(P2, 0.f);
                for (int k = 0; k < nPS_; ++k) pg[ps_[k]] = (float)pupil[ps_[k]];
                upload(pupilG_, pg, st_);
            }
        }
        std::vector<int> osNat(nOS_), invOs(N2, -1);
        for (int h = 0; h < nOS_; ++h) {
            osNat[h] = natOf(rep[h]);
            invOs[osNat[h]] = h;
            if (mirNat[h] >= 0) invOs[mirNat[h]] = -2 - h;
        }
        upload(wt_, wt, st_);
        upload(invPio_, invPio, st_);
        {
            std::vector<unsigned char> rowNZ(N, 0);
            for (int k = 0; k < nPS_; ++k) rowNZ[pio[k] / N] = 1;
            upload(rowNZ_, rowNZ, st_);
        }
        upload(invPde_, invPde, st_);
        {
            std::vector<int> yx(4 * (size_t)nPS_);
            for (int k = 0; k < nPS_; ++k) {
                yx[k] = pio[k] / N;
                yx[nPS_ + k] = pio[k] % N;
                yx[2 * nPS_ + k] = pde[k] / N;
                yx[3 * nPS_ + k] = pde[k] % N;
            }
            upload(pyx_, yx, st_);
        }
        upload(pde_, pde, st_);
        upload(pupil_, pup, st_);
        upload(osNat_, osNat, st_);
        upload(invOs_, invOs, st_);
        {   // where k_vogel_fused writes G'_H of stored point u in its N x (N+1) tile (natural
            // layout, columns [0, N/2] only): two targets (-1 for none), target * 8 + flags
            // (1: conjugate, 2: half, 4: real part), as the table of k_grad64_fused
            std::vector<int> gwt(2 * (size_t)nOS_, -1);
            auto target = [&](int nat) {
                const int y = nat / N, x = nat % N;
                return x > N / 2 ? -1 : y * (N + 1) + x;
            };
            auto enc = [](int t, int f) { return t < 0 ? -1 : t * 8 + f; };
            for (int u = 0; u < nOS_; ++u) {
                const int pn = osNat[u], mn = ((N - pn / N) % N) * N + (N - pn % N) % N;
                if (mirNat[u] >= 0) {
                    gwt[u] = enc(target(pn), 0);
                    gwt[nOS_ + u] = enc(target(mn), 1);
                } else if (mn == pn) {
                    gwt[u] = enc(target(pn), 4);
                } else {
                    gwt[u] = enc(target(pn), 2);
                    gwt[nOS_ + u] = enc(target(mn), 3);
                }
            }
            upload(gwt_, gwt, st_);
        }
        if constexpr (FUSED) {
            HIPCHK(hipFuncSetAttribute((const void*)k_otf_fused<N, FNT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)OLDS));
            int nCU = 0, dev = 0, perCU = 0;
            HIPCHK(hipGetDevice(&dev));
            HIPCHK(hipDeviceGetAttribute(&nCU, hipDeviceAttributeMultiprocessorCount, dev));
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_otf_fused<N, FNT>, FNT, OLDS));
            if constexpr (N == 128) {
                HIPCHK(hipFuncSetAttribute((const void*)k_otf_wave<FNT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)WLDS));
                if (otfWave_) {
                    int perW = 0;
                    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perW, k_otf_wave<FNT>, FNT, WLDS));
                    perCU = perW;
                }
            }
            otfGrid_ = std::max(nCU * std::max(perCU, 1), 1);   // persistent workgroups: all that fit at once
            HIPCHK(hipFuncSetAttribute((const void*)k_vogel_fused<N, FNT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)FLDS));
        }
        if constexpr (FUSED64) {
            HIPCHK(hipFuncSetAttribute((const void*)k_otf64_fused<N, F64NT, false>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)F64LDS));
            HIPCHK(hipFuncSetAttribute((const void*)k_otf64_fused<N, F64NT, true>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)F64LDS));
            HIPCHK(hipFuncSetAttribute((const void*)k_grad64_fused<N, F64NT>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)G64LDS));
            int nCU = 0, dev = 0;
            HIPCHK(hipGetDevice(&dev));
            HIPCHK(hipDeviceGetAttribute(&nCU, hipDeviceAttributeMultiprocessorCount, dev));
            otfGrid_ = std::max(nCU, 1);         // persistent workgroups: one per CU (the LDS fits one)
        }
        if constexpr (INIT64)
            HIPCHK(hipFuncSetAttribute((const void*)k_init64_fused<N, F64NT, T>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)I64LDS));
        HIPCHK(hipFuncSetAttribute((const void*)k_noise, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)((128 * 128 + 6 * 128) * sizeof(double))));
        area_ = 0;
        for (int k = 0; k < nPS_; ++k) area_ += pupil[ps_[k]];
        const int K = g_.K;
        std::vector<T> Bc((size_t)g_.nObj * K * nPS_);
        for (int o = 0; o < g_.nObj; ++o)
            for (int m = 0; m < K; ++m)
                for (int k = 0; k < nPS_; ++k)
                    Bc[((size_t)o * K + m) * nPS_ + k] = (T)modes[((size_t)o * K + m) * P2 + ps_[k]];
        upload(B_, Bc, st_);
        if (bf16Phase_) {       // transposed BF16 copy [obj][nPS][Kp] for k_phase_bf16
            const int Kp = (K + 31) & ~31;
            std::vector<__bf16> bt((size_t)g_.nObj * nPS_ * Kp, (__bf16)0.0f);
            for (int o = 0; o < g_.nObj; ++o)
                for (int m = 0; m < K; ++m)
                    for (int k = 0; k < nPS_; ++k)
                        bt[((size_t)o * nPS_ + k) * Kp + m] = (__bf16)(float)modes[((size_t)o * K + m) * P2 + ps_[k]];
            upload(Bt_, bt, st_);
        }
        HIPCHK(hipStreamSynchronize(st_));
    }

    void set_structure(const Structure& s, const std::vector<double>& otfNorm) override {
        use_lane(kSyncLane);
        st = s;
        otfNorm_ = otfNorm;
        std::vector<int> chFirst(g_.nCh, 0), chCount(g_.nCh, 0);
        for (int c = 0; c < g_.nCh; ++c) {
            chFirst[c] = -1;
            for (int j = 0; j < g_.nImg; ++j)
                if (s.imgCh[j] == c) {
                    if (chFirst[c] < 0) chFirst[c] = j;
                    chCount[c]++;
                }
            if (chFirst[c] < 0) chFirst[c] = 0;
        }
        upload(chObj_, s.chObj, st_);
        upload(chFirst_, chFirst, st_);
        upload(chCount_, chCount, st_);
        upload(imgObj_, s.imgObj, st_);
        upload(imgCh_, s.imgCh, st_);
        upload(w_, s.weight, st_);
        // pupil amplitude scale per (slot, image): the channel's otfNormalization, 0 for an
        // image a patch does not have (its OTF, and with it its share of P, Q and the
        // gradient, is then exactly zero)
        hScale_.assign((size_t)maxSlots_ * g_.nImg, T(0));
        imgOn_.assign((size_t)maxSlots_ * g_.nImg, 1);
        for (int sl = 0; sl < maxSlots_; ++sl)
            for (int j = 0; j < g_.nImg; ++j) hScale_[(size_t)sl * g_.nImg + j] = (T)otfNorm[s.imgCh[j]];
        upload(imgScale_, hScale_, st_);
        hStage_.alloc((size_t)maxSlots_ * g_.nImg * g_.K);
        std::fill(hStage_.p, hStage_.p + hStage_.n, 0.0);
        const size_t nS = maxSlots_;
        D_.alloc(nS * g_.nImg * nOS_);
        ftSum_.alloc(nS * g_.nObj * nOS_);
        gamma_.alloc(nS * g_.nObj);
        phiCh_.alloc(nS * g_.nCh * nPS_);
        rawF_.alloc(nS * g_.nImg * (size_t)g_.dataY * g_.dataX);
        HIPCHK(hipMemsetAsync(gamma_.p, 0, nS * g_.nObj * sizeof(double), st_));
        HIPCHK(hipStreamSynchronize(st_));
    }

    void load(int first, int n, const float* data, const double* phiCh) override {
        use_lane(kSyncLane);
        const size_t per = (size_t)g_.nImg * g_.dataY * g_.dataX;
        HIPCHK(hipMemcpyAsync(rawF_.p + (size_t)first * per, data, n * per * sizeof(float), hipMemcpyHostToDevice, st_));
        for (int k = 0; k < n; ++k) set_phi_channel(first + k, phiCh + (size_t)k * g_.nCh * g_.P * g_.P);
        HIPCHK(hipStreamSynchronize(st_));
    }

    void set_phi_channel(int slot, const double* phiCh) override {
        use_lane(kSyncLane);
        std::vector<T> h((size_t)g_.nCh * nPS_);
        for (int c = 0; c < g_.nCh; ++c)
            for (int k = 0; k < nPS_; ++k) h[(size_t)c * nPS_ + k] = (T)phiCh[(size_t)c * g_.P * g_.P + ps_[k]];
        HIPCHK(hipMemcpyAsync(phiCh_.p + (size_t)slot * g_.nCh * nPS_, h.data(), h.size() * sizeof(T),
                              hipMemcpyHostToDevice, st_));
        HIPCHK(hipStreamSynchronize(st_));
    }

    void frame_stats(const std::vector<ImgInit>& imgs, const std::vector<double>& window,
                     const std::vector<double>& noiseWindow, double* noise, double* stddev) override {
        use_lane(kSyncLane);
        const int nQ = (int)imgs.size();
        if (!nQ) return;
        const int s = g_.s;
        if (s > 128) throw std::runtime_error("frame_stats: patch side above 128");
        std::vector<double> wins(window);
        wins.insert(wins.end(), noiseWindow.begin(), noiseWindow.end());
        upload(win_, wins, st_);      // the window first: k_window reads win_ as the window
        winSet_ = true;
        std::vector<int> ints(4 * (size_t)nQ);
        for (int q = 0; q < nQ; ++q) {
            ints[q] = imgs[q].slot;
            ints[nQ + q] = imgs[q].img;
            ints[2 * nQ + q] = imgs[q].shiftX;
            ints[3 * nQ + q] = imgs[q].shiftY;
        }
        upload(sc_->ints, ints, st_);
        sc_->statOut.alloc(2 * (size_t)nQ);
        const size_t lds = ((size_t)s * s + 6 * (size_t)s) * sizeof(double);
        launch("k_noise", nQ, [&] {
            hipLaunchKernelGGL(k_noise, dim3(nQ), dim3(256), lds, st_, sc_->statOut.p, sc_->statOut.p + nQ, rawF_.p, win_.p,
                               win_.p + (size_t)s * s, sc_->ints.p, sc_->ints.p + nQ, sc_->ints.p + 2 * nQ, sc_->ints.p + 3 * nQ,
                               g_.nImg, g_.dataY, g_.dataX, g_.startY, g_.startX, s);
        });
        HIPCHK(hipMemcpyAsync(noise, sc_->statOut.p, nQ * sizeof(double), hipMemcpyDeviceToHost, st_));
        HIPCHK(hipMemcpyAsync(stddev, sc_->statOut.p + nQ, nQ * sizeof(double), hipMemcpyDeviceToHost, st_));
        HIPCHK(hipStreamSynchronize(st_));
        collect(kSyncLane);
    }

    // gamma[k * nObj + o] of slot slots[k], all in one copy
    void set_images_enabled(int slot, const unsigned char* on) override {
        use_lane(kSyncLane);
        const size_t b = (size_t)slot * g_.nImg;
        for (int j = 0; j < g_.nImg; ++j) {
            imgOn_[b + j] = on[j] ? 1 : 0;
            hScale_[b + j] = on[j] ? (T)otfNorm_[st.imgCh[j]] : T(0);
        }
        HIPCHK(hipMemcpyAsync(imgScale_.p + b, hScale_.data() + b, g_.nImg * sizeof(T), hipMemcpyHostToDevice, st_));
        HIPCHK(hipStreamSynchronize(st_));
    }

    void set_gamma(int n, const int* slots, const double* gamma) override {
        use_lane(kSyncLane);
        if (n <= 0) return;
        hGamma_.resize((size_t)maxSlots_ * g_.nObj, 0.0);
        for (int k = 0; k < n; ++k)
            std::copy(gamma + (size_t)k * g_.nObj, gamma + (size_t)(k + 1) * g_.nObj,
                      hGamma_.begin() + (size_t)slots[k] * g_.nObj);
        HIPCHK(hipMemcpyAsync(gamma_.p, hGamma_.data(), hGamma_.size() * sizeof(double), hipMemcpyHostToDevice, st_));
        HIPCHK(hipStreamSynchronize(st_));
    }

    // The synchronous lane waits for completion; on an evaluation lane the work is only enqueued (after that
    // lane's finished evaluation, before its next one) from the lane's own pinned staging.
    void init_images(const std::vector<ImgInit>& imgs, const std::vector<double>& window, int lane) override {
        use_lane(lane);
        if (imgs.empty()) return;
        if (!winSet_) {
            upload(win_, window, st_);
            winSet_ = true;
        }
        const int nQ = (int)imgs.size();
        const int chunk = maxTiles();
        // ftSum for every touched (slot, object)
        std::vector<int> pairs;
        std::map<std::pair<int, int>, int> seen;
        for (auto& im : imgs) {
            auto key = std::make_pair(im.slot, st.imgObj[im.img]);
            if (!seen.count(key)) {
                seen[key] = 1;
                pairs.push_back(key.first);
                pairs.push_back(key.second);
            }
        }
        std::vector<int> ints(4 * (size_t)nQ);
        for (int q = 0; q < nQ; ++q) {
            ints[q] = imgs[q].slot;
            ints[nQ + q] = imgs[q].img;
            ints[2 * nQ + q] = imgs[q].shiftX;
            ints[3 * nQ + q] = imgs[q].shiftY;
        }
        ints.insert(ints.end(), pairs.begin(), pairs.end());
        stage_async(sc_->initInts, sc_->hInit, ints);
        const int* base = sc_->initInts.p;
        if constexpr (INIT64) {
            const cx<double>* tw64 = nullptr;
            if constexpr (std::is_same_v<T, double>) tw64 = tw_.p;
            else tw64 = tw64_.p;
            launch("k_init64_fused", nQ, [&] {
                hipLaunchKernelGGL((k_init64_fused<N, F64NT, T>), dim3(nQ), dim3(F64NT), I64LDS, st_, D_.p, rawF_.p,
                                   win_.p, base, base + nQ, base + 2 * nQ, base + 3 * nQ, hread_.p, tw64, g_.nImg,
                                   g_.dataY, g_.dataX, g_.startY, g_.startX, g_.s, nOS_);
            });
        } else {
            sc_->tiles.alloc((size_t)std::min(chunk, nQ) * N2);
            for (int b = 0; b < nQ; b += chunk) {
                const int n = std::min(chunk, nQ - b);
                const int* ds = base + b;
                const int* di = base + nQ + b;
                const int* dx = base + 2 * nQ + b;
                const int* dy = base + 3 * nQ + b;
                launch("k_window", n, [&] {
                    hipLaunchKernelGGL(k_window<T>, dim3(n), dim3(256), 0, st_, sc_->tiles.p, rawF_.p, win_.p, ds, di, dx,
                                       dy, g_.nImg, g_.dataY, g_.dataX, g_.startY, g_.startX, g_.s, N);
                });
                fft2(sc_->tiles.p, n, -1);
                launch("k_store_d", n, [&] {
                    hipLaunchKernelGGL(k_store_d<T>, dim3(nblk((size_t)n * nOS_)), dim3(256), 0, st_, D_.p, sc_->tiles.p,
                                       osNat_.p, ds, di, n, nOS_, g_.nImg, N2);
                });
            }
        }
        const int nPairs = (int)pairs.size() / 2;
        launch("k_ftsum", nPairs, [&] {
            hipLaunchKernelGGL(k_ftsum<T>, dim3(nblk((size_t)nPairs * nOS_)), dim3(256), 0, st_, ftSum_.p, D_.p,
                               base + 4 * nQ, imgObj_.p, nPairs, g_.nImg, g_.nObj, nOS_);
        });
        if (lane == kSyncLane) {
            HIPCHK(hipStreamSynchronize(st_));
            collect(kSyncLane);
        }
    }

    void evaluate(std::vector<EvalReq>& reqs) override {
        evaluate_begin(kSyncLane, reqs);
        evaluate_end(kSyncLane, reqs);
    }

    // ---- device line search: per lane a fixed list of patches and one captured round ----
    struct LsLane {
        int nL = 0;
        std::vector<int> slots;
        DBuf<LsState> st;
        DBuf<double> L, qS, part, fl, dirA;
        DBuf<int> tab, startIdx, dirTab;   // tab: slots [nL] | scale rows [nQ] | qLine [nQ] | active [nL] | lq [nQ] | lrow [nQ]
        DBuf<C> S;
        DBuf<T> phi;
        HBuf<LsState> hSt;
        HBuf<double> hFl, hDir;
        HBuf<int> hIdx, hDirTab;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        hipEvent_t done = nullptr;
        bool useGraph = false;
        std::vector<int> probes;           // each line's probes at the last read-back
    };
    LsLane ls_[kSyncLane + 1];

    void ls_clear() override {
        for (auto& l : ls_) {
            if (l.exec) (void)hipGraphExecDestroy(l.exec);
            if (l.graph) (void)hipGraphDestroy(l.graph);
            l.exec = nullptr;
            l.graph = nullptr;
            l.nL = 0;
        }
    }

    bool ls_begin(int lane, const std::vector<int>& slots, bool graph) override {
        if constexpr (!(FUSED || FUSED64)) {
            return false;
        } else {
            use_lane(lane);
            LsLane& l = ls_[lane];
            if (l.exec) (void)hipGraphExecDestroy(l.exec);
            if (l.graph) (void)hipGraphDestroy(l.graph);
            l.exec = nullptr;
            l.graph = nullptr;
            const int nL = (int)slots.size(), nImg = g_.nImg, nQ = nL * nImg;
            l.nL = nL;
            l.slots = slots;
            l.useGraph = graph && !timing_;     // kernel events time each launch: direct launches
            l.probes.assign(nL, 0);
            if (!nL) return true;
            if (!l.done) HIPCHK(hipEventCreateWithFlags(&l.done, hipEventDisableTiming));
            lineBuf_.alloc((size_t)2 * maxSlots_ * nImg * nPS_);
            const int nChunk = (nOS_ + 255) / 256;
            l.st.alloc(nL);
            l.L.alloc(nL);
            l.qS.alloc(nQ);
            l.part.alloc((size_t)nL * g_.nObj * nChunk);
            l.S.alloc((size_t)nQ * nOS_);
            l.phi.alloc((size_t)nQ * nPS_);
            std::vector<int> t((size_t)2 * nL + 4 * (size_t)nQ);
            for (int r = 0; r < nL; ++r) {
                t[r] = slots[r];
                for (int j = 0; j < nImg; ++j) {
                    const int q = r * nImg + j;
                    t[nL + q] = slots[r] * nImg + j;               // pupil scale row
                    t[nL + nQ + q] = -2;                           // qLine: no probe yet
                    t[2 * nL + 2 * nQ + q] = q;                    // lq
                    t[2 * nL + 3 * nQ + q] = slots[r] * nImg + j;  // lrow
                }
                t[nL + 2 * nQ + r] = 0;                            // active
            }
            upload(l.tab, t, st_);
            std::vector<LsState> idle(nL);
            for (auto& x : idle) { x = LsState{}; x.pc = -1; }
            l.hSt.alloc(nL);
            std::copy(idle.begin(), idle.end(), l.hSt.p);
            HIPCHK(hipMemcpyAsync(l.st.p, l.hSt.p, nL * sizeof(LsState), hipMemcpyHostToDevice, st_));
            HIPCHK(hipStreamSynchronize(st_));
            if (l.useGraph) {
                std::lock_guard<std::mutex> lk(capture_mutex());
                HIPCHK(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
                ls_round(l, false);
                HIPCHK(hipStreamEndCapture(st_, &l.graph));
                HIPCHK(hipGraphInstantiate(&l.exec, l.graph, nullptr, nullptr, 0));
            }
            return true;
        }
    }

    // One round of a lane's line searches: the probes' transforms from the lines' phases,
    // P / Q, then each probed patch's metric and its line's next probe (k_ls_metric).  Raw
    // launches when captured; with `named` each through launch() (its statistics, events
    // when timing).
    void ls_round(LsLane& l, bool named) {
        const int nL = l.nL, nImg = g_.nImg, nQ = nL * nImg, nChunk = (nOS_ + 255) / 256;
        const int* tSlot = l.tab.p;
        const int* tScale = l.tab.p + nL;
        int* tLine = l.tab.p + nL + nQ;
        int* tActive = l.tab.p + nL + 2 * nQ;
        const size_t dirOff = (size_t)maxSlots_ * nImg * nPS_;
        auto go = [&](const char* name, auto&& f) {
            if (named)
                launch(name, 0, f);
            else
                f();
        };
        if constexpr (FUSED) {
            if (otfWave_) {
                if constexpr (N == 128)
                    go("k_otf_wave", [&] {
                        hipLaunchKernelGGL((k_otf_wave<FNT>), dim3(std::min(nQ, otfGrid_)), dim3(FNT), WLDS, st_, l.S.p,
                                           l.phi.p, pupilG_.p, tScale, imgScale_.p, hsrc_.p, psd_.p, tw_.p, nPS_,
                                           nOS_, nQ, lineBuf_.p, tLine, l.qS.p, dirOff);
                    });
            } else
            go("k_otf_fused", [&] {
                hipLaunchKernelGGL((k_otf_fused<N, FNT>), dim3(std::min(nQ, otfGrid_)), dim3(FNT), OLDS, st_, l.S.p,
                                   l.phi.p, invPs_.p, pupil_.p, tScale, imgScale_.p, hread_.p, tw_.p, nPS_, nOS_, nQ,
                                   lineBuf_.p, tLine, l.qS.p, dirOff);
            });
        } else if constexpr (FUSED64) {
            go("k_otf64_fused", [&] {
                hipLaunchKernelGGL((k_otf64_fused<N, F64NT, true>), dim3(std::min(nQ, otfGrid_)), dim3(F64NT), F64LDS,
                                   st_, l.S.p, l.phi.p, invPs_.p, pupil_.p, tScale, imgScale_.p, hread_.p, tw_.p, nPS_,
                                   nOS_, nQ, tLine, lineBuf_.p, tLine, l.qS.p, dirOff);
            });
        }
        go("k_pq", [&] {
            hipLaunchKernelGGL(k_pq<T>, dim3(nChunk, g_.nObj, nL), dim3(256), 0, st_, (double*)nullptr,
                               (double*)nullptr, l.part.p, l.S.p,
                               D_.p, ftSum_.p, gamma_.p, tSlot, chObj_.p, chFirst_.p, chCount_.p, wt_.p, imgScale_.p,
                               g_.nCh, g_.nObj, nImg, nOS_, 1, tActive);
        });
        go("k_ls_metric", [&] {
            hipLaunchKernelGGL(k_ls_metric, dim3(nL), dim3(64), 0, st_, l.st.p, l.L.p, l.part.p, w_.p, g_.nObj, nChunk,
                               1.0 / ((double)g_.s * g_.s), tLine, l.qS.p, tActive, tSlot, nImg);
        });
    }

    // new lines: their states and the directions' phases B dAlpha (no channel phase) into
    // the slots' line buffers, on the lane's stream before its next rounds
    void ls_start(int lane, const std::vector<LsStart>& starts) override {
        if (starts.empty()) return;
        use_lane(lane);
        LsLane& l = ls_[lane];
        const int n = (int)starts.size(), nImg = g_.nImg, K = g_.K, rows = n * nImg;
        const size_t row = (size_t)nImg * K;
        l.hIdx.alloc(n);
        l.hFl.alloc(2 * (size_t)n);
        l.hDir.alloc((size_t)n * row);
        l.hDirTab.alloc(4 * (size_t)rows);
        const int LS = maxSlots_ * nImg;
        for (int i = 0; i < n; ++i) {
            const LsStart& a = starts[i];
            if (a.idx < 0 || a.idx >= l.nL || !a.dAlpha) throw std::runtime_error("ls_start: bad line");
            l.hIdx.p[i] = a.idx;
            l.probes[a.idx] = 0;
            l.hFl.p[2 * i] = a.fa;
            l.hFl.p[2 * i + 1] = a.lambda;
            std::memcpy(l.hDir.p + (size_t)i * row, a.dAlpha, sizeof(double) * row);
            for (int j = 0; j < nImg; ++j) {
                const int g = i * nImg + j;
                l.hDirTab.p[g] = g;                                        // alpha row
                l.hDirTab.p[rows + g] = -1;                                // no channel phase
                l.hDirTab.p[2 * rows + g] = st.imgObj[j];
                l.hDirTab.p[3 * rows + g] = -1 - (LS + l.slots[a.idx] * nImg + j);   // line direction row
            }
        }
        l.startIdx.alloc(n);
        l.fl.alloc(2 * (size_t)n);
        l.dirA.alloc((size_t)n * row);
        l.dirTab.alloc(4 * (size_t)rows);
        HIPCHK(hipMemcpyAsync(l.startIdx.p, l.hIdx.p, n * sizeof(int), hipMemcpyHostToDevice, st_));
        HIPCHK(hipMemcpyAsync(l.fl.p, l.hFl.p, 2 * n * sizeof(double), hipMemcpyHostToDevice, st_));
        HIPCHK(hipMemcpyAsync(l.dirA.p, l.hDir.p, (size_t)n * row * sizeof(double), hipMemcpyHostToDevice, st_));
        HIPCHK(hipMemcpyAsync(l.dirTab.p, l.hDirTab.p, 4 * (size_t)rows * sizeof(int), hipMemcpyHostToDevice, st_));
        launch("k_ls_init", n, [&] {
            const int nQ = l.nL * nImg;
            hipLaunchKernelGGL(k_ls_init, dim3((n + 63) / 64), dim3(64), 0, st_, l.st.p, l.startIdx.p, l.fl.p, n,
                               l.tab.p + l.nL + nQ, l.qS.p, l.tab.p + l.nL + 2 * nQ, l.tab.p, nImg);
        });
        const int* gA = l.dirTab.p;
        if (bf16Phase_) {
            if constexpr (std::is_same_v<T, float>) {
                launch("k_phase_bf16", rows, [&] {
                    hipLaunchKernelGGL(k_phase_bf16, dim3((nPS_ + 63) / 64, (rows + 63) / 64), dim3(256), 0, st_,
                                       l.phi.p, reinterpret_cast<const bf16x8*>(Bt_.p), phiCh_.p, l.dirA.p, gA, gA + rows,
                                       gA + 2 * rows, rows, K, (K + 31) & ~31, nPS_, lineBuf_.p, gA + 3 * rows,
                                       (const int*)nullptr);
                });
            }
        } else {
            launch("k_phase_mfma", rows, [&] {
                hipLaunchKernelGGL(k_phase_mfma<T>, dim3((nPS_ + 63) / 64, (rows + 63) / 64), dim3(256), 0, st_,
                                   l.phi.p, B_.p, phiCh_.p, l.dirA.p, gA, gA + rows, gA + 2 * rows, rows, K, nPS_,
                                   lineBuf_.p, gA + 3 * rows, (const int*)nullptr);
            });
        }
    }

    void ls_run(int lane, int rounds) override {
        use_lane(lane);
        LsLane& l = ls_[lane];
        if (!l.nL || (l.useGraph && !l.exec)) throw std::runtime_error("ls_run: no device line search on this lane");
        if (l.useGraph) {
            launch("k_ls_round", 0, [&] {
                for (int k = 0; k < rounds; ++k) HIPCHK(hipGraphLaunch(l.exec, st_));
            });
        } else {
            for (int k = 0; k < rounds; ++k) ls_round(l, true);
        }
        l.hSt.alloc(l.nL);
        HIPCHK(hipMemcpyAsync(l.hSt.p, l.st.p, l.nL * sizeof(LsState), hipMemcpyDeviceToHost, st_));
        HIPCHK(hipEventRecord(l.done, st_));
    }

    void ls_end(int lane, std::vector<LsResult>& out) override {
        LsLane& l = ls_[lane];
        HIPCHK(hipEventSynchronize(l.done));
        out.resize(l.nL);
        long long probes = 0;
        for (int r = 0; r < l.nL; ++r) {
            const LsState& x = l.hSt.p[r];
            out[r] = {x.pc == -1, x.stepOut, x.fOut, x.nProbe};
            probes += x.nProbe - l.probes[r];
            l.probes[r] = x.nProbe;
        }
        collect(lane);
        // the rounds' work in images evaluated (skipped images not counted); "_line_images"
        // counts the images whose transform read the line (base and direction) for phi
        const long long img = probes * g_.nImg;
        for (const char* k : {"k_otf_fused", "k_otf_wave", "k_line_phase", "k_otf64_fused", "k_pq", "k_ls_round"}) {
            auto it = stats_.find(k);
            if (it != stats_.end() && (std::strcmp(k, "k_ls_round") == 0 || !l.useGraph)) it->second.units += img;
        }
        if (timing_ && !l.useGraph && (FUSED || FUSED64)) {
            auto& li = stats_["_line_images"];
            li.name = "_line_images";
            li.units += img;
        }
    }

    // Enqueue the evaluation of `reqs` on an evaluation lane (kSyncLane: the synchronous calls) and
    // return without waiting; evaluate_end waits for that lane and fills f and g.  The
    // requests' alpha (and the lane's staging) must stay untouched until then.  Requests
    // beyond one chunk of the tile workspace are evaluated synchronously first.
    void evaluate_begin(int lane, std::vector<EvalReq>& reqs) override {
        use_lane(lane);
        const int nR = (int)reqs.size();
        sc_->nReq = 0;
        if (!nR) return;
        const int K = g_.K, nImg = g_.nImg;
        const int perChunk = std::max(1, std::min(maxBatch_, maxTiles() / nImg));
        int b = 0;
        for (; nR - b > perChunk; b += perChunk) {
            eval_chunk_launch(&reqs[b], perChunk, K, nImg);
            HIPCHK(hipEventSynchronize(sc_->done));
            eval_chunk_finish(&reqs[b], perChunk, K, nImg);
        }
        eval_chunk_launch(&reqs[b], nR - b, K, nImg);
        sc_->nReq = nR - b;
        laneFirst_[lane] = b;
    }

    void evaluate_end(int lane, std::vector<EvalReq>& reqs) override {
        use_lane(lane);
        if (!sc_->nReq) return;
        HIPCHK(hipEventSynchronize(sc_->done));
        eval_chunk_finish(&reqs[laneFirst_[lane]], sc_->nReq, g_.K, g_.nImg);
        sc_->nReq = 0;
        collect(lane);
    }

    // The restored objects of n slots at their alphas on the synchronous lane, in chunks of
    // patches: the forward chain and the sums P0 = sum conj(D) S, Q0 = sum |S|^2 (no gamma),
    // then per (patch, object) the restoration proper on the device in FP64: conj(P0) / Q0,
    // the Scharmer filter (its box smoothings, weights and connected region), the inverse
    // transform and the crop, each chunk's floats downloaded to pinned memory.
    void restore_objects_begin(int n, const int* slots, const double* const* alpha, const double* nf,
                               const double* c2, bool filter, bool full) override {
        use_lane(kSyncLane);
        const int nO = g_.nObj, perChunk = std::max(1, std::min(maxBatch_, maxTiles() / g_.nImg));
        const size_t outPer = (size_t)nO * (full ? N2 : (size_t)g_.s * g_.s);
        if (!twD_.p) {
            std::vector<cx<double>> tw(N);
            for (int k = 0; k < N; ++k) {
                long double a = -2.0L * 3.14159265358979323846264338327950288L * k / N;
                tw[k] = {(double)cosl(a), (double)sinl(a)};
            }
            upload(twD_, tw, st_);
        }
        std::vector<EvalReq> rq(n);
        for (int k = 0; k < n; ++k) {
            rq[k].slot = slots[k];
            rq[k].alpha = alpha[k];
        }
        hRs_.alloc((size_t)n * outPer);
        std::vector<double> c2v(c2, c2 + nO);
        upload(rsC2_, c2v, st_);
        for (int b = 0; b < n; b += perChunk) {
            const int m = std::min(perChunk, n - b), R = m * nO;
            // forward() rewrites the lane's pinned table and alpha staging, which the previous
            // chunk's queued uploads still read: let them finish first
            if (b) HIPCHK(hipStreamSynchronize(st_));
            forward(&rq[b], m, /*useGamma=*/0);
            const size_t RN = (size_t)R * N2;
            rsAo_.alloc(RN);
            rsT_.alloc(RN);
            rsDn_.alloc(RN);
            rsA_.alloc(RN);
            rsB_.alloc(RN);
            rsOut_.alloc((size_t)m * outPer);
            std::vector<double> nfv(nf + (size_t)b * nO, nf + (size_t)(b + m) * nO);
            upload(rsNf_, nfv, st_);
            launch("k_rs_build", R, [&] {
                hipLaunchKernelGGL(k_rs_build, dim3(nblk(RN)), dim3(256), 0, st_, rsAo_.p, rsDn_.p, rsA_.p, sc_->Pp.p,
                                   sc_->Qp.p, rsCode_.p, R, nOS_, (int)N2);
            });
            if (filter) {
                auto box = [&](double* o, const double* i, const double* nfp, int w) {
                    launch("k_rs_box", R, [&] {
                        hipLaunchKernelGGL(k_rs_box, dim3(nblk(RN)), dim3(256), 0, st_, o, i, nfp, R, N, w);
                    });
                };
                box(rsB_.p, rsA_.p, rsNf_.p, 1);       // v = nf / smooth(pt, 1)
                launch("k_rs_weights", R, [&] {
                    hipLaunchKernelGGL(k_rs_weights, dim3(nblk(RN)), dim3(256), 0, st_, rsA_.p, rsB_.p, R, N);
                });
                launch("k_rs_region", R, [&] {
                    hipLaunchKernelGGL(k_rs_region<N>, dim3(R), dim3(1024), 0, st_, rsB_.p, rsA_.p);
                });
                box(rsA_.p, rsB_.p, nullptr, 4);       // rs = smooth(weights in the region, 4)
            }
            launch("k_rs_apply", R, [&] {
                hipLaunchKernelGGL(k_rs_apply, dim3(nblk(RN)), dim3(256), 0, st_, rsT_.p, rsAo_.p,
                                   filter ? rsA_.p : (const double*)nullptr, rsC2_.p, R, N, nO);
            });
            launch("k_rs_ifft", R, [&] {
                constexpr int L = 16;
                hipLaunchKernelGGL((k_fft_rows<double, N, +1, L>), dim3(R * (N / L)), dim3(L * N / 16), 0, st_, rsT_.p,
                                   twD_.p);
                hipLaunchKernelGGL((k_fft_cols<double, N, +1, L>), dim3(R * (N / L)), dim3(L * N / 16), 0, st_, rsT_.p,
                                   twD_.p);
            });
            const size_t nOut = (size_t)m * outPer;
            launch("k_rs_crop", R, [&] {
                hipLaunchKernelGGL(k_rs_crop, dim3(nblk(nOut)), dim3(256), 0, st_, rsOut_.p, rsT_.p, R, N, g_.s,
                                   full ? 1 : 0);
            });
            HIPCHK(hipMemcpyAsync(hRs_.p + (size_t)b * outPer, rsOut_.p, nOut * sizeof(float), hipMemcpyDeviceToHost,
                                  st_));
        }
        HIPCHK(hipEventRecord(sc_->done, st_));
        restoreN_ = n;
        restoreOut_ = n * outPer;
    }

    void restore_objects_end(float* out) override {
        use_lane(kSyncLane);
        HIPCHK(hipEventSynchronize(sc_->done));
        std::memcpy(out, hRs_.p, restoreOut_ * sizeof(float));
        restoreN_ = 0;
        restoreOut_ = 0;
        collect(kSyncLane);
    }

    void restore_extra(int n, const int* slots, const double* const* alpha,
                       const std::vector<std::vector<ImgInit>>& imgs, const std::vector<float>& objFull, int want,
                       float* psf, float* cobj, float* res) override {
        use_lane(kSyncLane);
        if (n < 1 || !want) return;
        const int nO = g_.nObj, nImg = g_.nImg;
        const size_t s2 = (size_t)g_.s * g_.s;
        const bool wPsf = want & (kOutPsf | kOutPsfAvg), wConv = want & (kOutCobj | kOutRes), wWin = want & kOutRes;
        if (wConv && objFull.size() < (size_t)n * nO * N2) throw std::runtime_error("restore_extra: objFull too small");
        const int perChunk = std::max(1, std::min(maxBatch_, maxTiles() / nImg));
        const size_t psfPer = (size_t)((want & kOutPsfAvg) ? nO : nImg) * s2, imgPer = (size_t)nImg * s2;
        for (int b = 0; b < n; b += perChunk) {
            const int m = std::min(perChunk, n - b);
            std::vector<EvalReq> rq(m);
            int nI = 0;
            for (int r = 0; r < m; ++r) {
                rq[r].slot = slots[b + r];
                rq[r].alpha = alpha[b + r];
                nI += (int)imgs[b + r].size();
            }
            // the chunk's float outputs, zero where a patch lacks an image
            if (wPsf) xF_[0].alloc(std::max<size_t>((size_t)m * psfPer, 1));
            if (want & kOutCobj) xF_[1].alloc(std::max<size_t>((size_t)m * imgPer, 1));
            if (want & kOutRes) xF_[2].alloc(std::max<size_t>((size_t)m * imgPer, 1));
            if (wPsf) HIPCHK(hipMemsetAsync(xF_[0].p, 0, (size_t)m * psfPer * sizeof(float), st_));
            if (want & kOutCobj) HIPCHK(hipMemsetAsync(xF_[1].p, 0, (size_t)m * imgPer * sizeof(float), st_));
            if (want & kOutRes) HIPCHK(hipMemsetAsync(xF_[2].p, 0, (size_t)m * imgPer * sizeof(float), st_));
            if (nI) {
                forward(rq.data(), m, /*useGamma=*/0);        // S of request r: [r][nImg][nOS_]
                // per image q: slot, image, shifts, object tile (request * nObj + object), S row;
                // then each request's first image row (m + 1)
                std::vector<int> ints(6 * (size_t)nI + m + 1);
                int q = 0;
                for (int r = 0; r < m; ++r) {
                    ints[6 * (size_t)nI + r] = q;
                    for (const ImgInit& im : imgs[b + r]) {
                        ints[q] = im.slot;
                        ints[nI + q] = im.img;
                        ints[2 * nI + q] = im.shiftX;
                        ints[3 * nI + q] = im.shiftY;
                        ints[4 * nI + q] = r * nO + st.imgObj[im.img];
                        ints[5 * nI + q] = r * nImg + im.img;
                        ++q;
                    }
                }
                ints[6 * (size_t)nI + m] = q;
                upload(xInts_, ints, st_);
                const int* dSlot = xInts_.p;
                const int* dImg = xInts_.p + nI;
                const int* dTile = xInts_.p + 4 * nI;
                const int* dRow = xInts_.p + 5 * nI;
                const int* dFirst = xInts_.p + 6 * nI;
                const size_t nT = (size_t)nI * N2;
                xT_.alloc(nT);
                auto crop = [&](DBuf<double>& out, int sh, double scale) {
                    out.alloc((size_t)nI * s2);
                    launch("k_crop_re", nI, [&] {
                        hipLaunchKernelGGL(k_crop_re<T>, dim3(nblk((size_t)nI * s2)), dim3(256), 0, st_, out.p, xT_.p,
                                           nI, N, g_.s, sh, scale);
                    });
                };
                auto scatter = [&] {
                    launch("k_otf_scatter", nI, [&] {
                        hipLaunchKernelGGL(k_otf_scatter<T>, dim3(nblk(nT)), dim3(256), 0, st_, xT_.p, sc_->S.p,
                                           invOs_.p, dRow, nI, nOS_, N2);
                    });
                };
                if (wPsf) {            // PSF: normalised inverse transform of the OTF, re-centred
                    scatter();
                    fft2(xT_.p, nI, +1);
                    crop(xD_[0], N / 2, 1.0 / N2);
                }
                if (wConv) {           // convolved objects
                    std::vector<C> h((size_t)m * nO * N2);
                    const float* src = objFull.data() + (size_t)b * nO * N2;
                    for (size_t k = 0; k < h.size(); ++k) h[k] = {(T)src[k], T(0)};
                    upload(xO_, h, st_);
                    fft2(xO_.p, m * nO, -1);
                    scatter();
                    launch("k_mul_obj", nI, [&] {
                        hipLaunchKernelGGL(k_mul_obj<T>, dim3(nblk(nT)), dim3(256), 0, st_, xT_.p, xO_.p, dTile, nI, N2,
                                           (T)(1.0 / N2));
                    });
                    fft2(xT_.p, nI, +1);
                    crop(xD_[1], 0, 1.0);
                }
                if (wWin) {            // windowed, padded frames (the inverse of imgFT)
                    if (!winSet_) throw std::runtime_error("restore_extra: frames not initialised");
                    launch("k_window", nI, [&] {
                        hipLaunchKernelGGL(k_window<T>, dim3(nI), dim3(256), 0, st_, xT_.p, rawF_.p, win_.p, dSlot, dImg,
                                           xInts_.p + 2 * nI, xInts_.p + 3 * nI, nImg, g_.dataY, g_.dataX, g_.startY,
                                           g_.startX, g_.s, N);
                    });
                    crop(xD_[2], 0, 1.0);
                }
                // the float outputs in the caller's layout (SubImage::getPSF / addPSF,
                // convolveImage, convolvedResidual): per image, or the PSFs averaged per object
                const int ns2 = (int)s2;
                launch("k_out_finish", nI, [&] {
                    if (want & kOutPsf)
                        hipLaunchKernelGGL(k_out_rows, dim3(nblk((size_t)nI * s2)), dim3(256), 0, st_, xF_[0].p,
                                           xD_[0].p, (const double*)nullptr, dTile, dImg, nI, nO, nImg, ns2);
                    if (want & kOutPsfAvg)
                        hipLaunchKernelGGL(k_psf_avg, dim3(nblk((size_t)m * nO * s2)), dim3(256), 0, st_, xF_[0].p,
                                           xD_[0].p, dFirst, dImg, imgObj_.p, m, nO, ns2);
                    if (want & kOutCobj)
                        hipLaunchKernelGGL(k_out_rows, dim3(nblk((size_t)nI * s2)), dim3(256), 0, st_, xF_[1].p,
                                           xD_[1].p, (const double*)nullptr, dTile, dImg, nI, nO, nImg, ns2);
                    if (want & kOutRes)
                        hipLaunchKernelGGL(k_out_rows, dim3(nblk((size_t)nI * s2)), dim3(256), 0, st_, xF_[2].p,
                                           xD_[1].p, xD_[2].p, dTile, dImg, nI, nO, nImg, ns2);
                });
            }
            // through pinned staging into the caller's arrays
            auto fetch = [&](DBuf<float>& d, float* dst, size_t per) {
                hF_.alloc((size_t)m * per);
                HIPCHK(hipMemcpyAsync(hF_.p, d.p, (size_t)m * per * sizeof(float), hipMemcpyDeviceToHost, st_));
                HIPCHK(hipStreamSynchronize(st_));
                std::memcpy(dst + (size_t)b * per, hF_.p, (size_t)m * per * sizeof(float));
            };
            if (wPsf) fetch(xF_[0], psf, psfPer);
            if (want & kOutCobj) fetch(xF_[1], cobj, imgPer);
            if (want & kOutRes) fetch(xF_[2], res, imgPer);
        }
        collect(kSyncLane);
    }

    bool restore_ready() override {
        const hipError_t e = hipEventQuery(lanes_[kSyncLane].done);
        if (e == hipErrorNotReady) return false;
        HIPCHK(e);
        return true;
    }

    void debug_get(const std::string& what, int slot, std::vector<double>& out) override {
        if (what == "fft_lines") {      // the transform passes' lines per workgroup (rows, columns)
            out = {(double)lr_, (double)lc_};
            return;
        }
        if (what == "otf_lines") {      // ... of the four OTF passes (0 before the first OTF)
            out = {(double)stageL_[0], (double)stageL_[1], (double)stageL_[2], (double)stageL_[3]};
            return;
        }
        use_lane(kSyncLane);
        auto getc = [&](const C* p, size_t n) {
            std::vector<C> h(n);
            HIPCHK(hipMemcpy(h.data(), p, n * sizeof(C), hipMemcpyDeviceToHost));
            out.resize(2 * n);
            for (size_t k = 0; k < n; ++k) { out[2 * k] = h[k].x; out[2 * k + 1] = h[k].y; }
        };
        auto getd = [&](const double* p, size_t n) {
            out.resize(n);
            HIPCHK(hipMemcpy(out.data(), p, n * sizeof(double), hipMemcpyDeviceToHost));
        };
        // support arrays are stored on the Hermitian half; expand rows of nOS_ to the full
        // support (complex: conjugate for mirrors)
        auto expand = [&](int rows, bool cplx) {
            std::vector<double> h;
            h.swap(out);
            const int w = cplx ? 2 : 1;
            out.assign((size_t)rows * nOSfull_ * w, 0.0);
            for (int k = 0; k < rows; ++k)
                for (int u = 0; u < nOSfull_; ++u) {
                    const int code = fullMap_[u], hh = code >= 0 ? code : -1 - code;
                    const size_t src = ((size_t)k * nOS_ + hh) * w, dst = ((size_t)k * nOSfull_ + u) * w;
                    out[dst] = h[src];
                    if (cplx) out[dst + 1] = code >= 0 ? h[src + 1] : -h[src + 1];
                }
        };
        HIPCHK(hipStreamSynchronize(st_));
        const size_t nI = (size_t)g_.nImg * nOS_, nO = (size_t)g_.nObj * nOS_;
        if (what == "D") { getc(D_.p + slot * nI, nI); expand(g_.nImg, true); }
        else if (what == "S") { getc(sc_->S.p + slot * nI, nI); expand(g_.nImg, true); }   // slot = request index
        else if (what == "P") { getd(sc_->Pp.p + slot * 2 * nO, 2 * nO); expand(g_.nObj, true); }
        else if (what == "Q") { getd(sc_->Qp.p + slot * nO, nO); expand(g_.nObj, false); }
        else if (what == "ftsum") { getd(ftSum_.p + slot * nO, nO); expand(g_.nObj, false); }
        else if (what == "gamma") getd(gamma_.p + slot * g_.nObj, g_.nObj);
        else if (what == "phi") {
            const size_t n = (size_t)g_.nImg * nPS_;
            std::vector<T> h(n);
            HIPCHK(hipMemcpy(h.data(), sc_->phi.p + slot * n, n * sizeof(T), hipMemcpyDeviceToHost));
            out.assign(h.begin(), h.end());
        } else out.clear();
    }

    std::vector<KernelStat> kernel_stats() override {
        std::vector<KernelStat> v;
        for (auto& e : stats_) v.push_back(e.second);
        return v;
    }
    void reset_stats() override { stats_.clear(); }

private:
    int maxTiles() const {
        // tile workspace (per lane, plus a real array of half its size) bounded to 4 GiB
        const size_t cap = (size_t)4 << 30;
        const size_t per = (size_t)N2 * sizeof(C);
        const size_t t = std::max<size_t>(16, std::min<size_t>(cap / per, (size_t)maxBatch_ * g_.nImg));
        return (int)(tileCap_ > 0 ? std::min<size_t>(t, (size_t)tileCap_) : t);
    }

    // Per-kernel timing (MFBD_KERNEL_TIMING set): an event pair around each launch on
    // the pipeline's stream, read back after the stream synchronises.
    template <typename F>
    void launch(const char* name, F&& f) { launch(name, 0, f); }

    template <typename F>
    void launch(const char* name, long long units, F&& f) {
        if (!timing_) {
            f();
            HIPCHK(hipGetLastError());
            return;
        }
        if (evPool_.empty()) {
            hipEvent_t a, b;
            HIPCHK(hipEventCreate(&a));
            HIPCHK(hipEventCreate(&b));
            evPool_.push_back({a, b});
        }
        auto pr = evPool_.back();
        evPool_.pop_back();
        HIPCHK(hipEventRecord(pr.first, st_));
        f();
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(pr.second, st_));
        evPending_.push_back({name, pr, units, lane_});
    }

    // fold the timings of finished launches into the statistics: every pending launch
    // (lane < 0) or those of one lane; waits for each launch's end event
    void collect(int lane = -1) {
        if (evPending_.empty()) return;
        std::vector<PendingEv> keep;
        for (auto& e : evPending_) {
            if (lane >= 0 && e.lane != lane) {
                keep.push_back(e);
                continue;
            }
            HIPCHK(hipEventSynchronize(e.ev.second));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e.ev.first, e.ev.second));
            auto& st = stats_[e.name];
            st.name = e.name;
            st.ms += ms;
            st.launches++;
            st.units += e.units;
            evPool_.push_back(e.ev);
        }
        evPending_.swap(keep);
    }

    // One row or column pass of L lines per workgroup over nTiles tiles.  Every line's
    // butterflies are the same whatever L, so the choice changes only the speed.
    template <int DIR, int L>
    void fft_pass(C* tiles, int nTiles, bool rows) {
        const unsigned gr = nTiles * (N / L);
        if (rows)
            hipLaunchKernelGGL((k_fft_rows<T, N, DIR, L>), dim3(gr), dim3(L * N / 16), 0, st_, tiles, tw_.p);
        else
            hipLaunchKernelGGL((k_fft_cols<T, N, DIR, L>), dim3(gr), dim3(L * N / 16), 0, st_, tiles, tw_.p);
    }

    template <int DIR>
    void fft_pass_any(C* tiles, int nTiles, bool rows, int l) {
        if constexpr (N == 256) {
            if (l == LINES / 4) return fft_pass<DIR, LINES / 4>(tiles, nTiles, rows);
            if (l == LINES / 2) return fft_pass<DIR, LINES / 2>(tiles, nTiles, rows);
        }
        fft_pass<DIR, LINES>(tiles, nTiles, rows);
    }

    // The four passes of the multi-pass OTF (see k_field_rows) at L lines per workgroup
    template <int L>
    void otf_stage(int stage, C* tiles, T* power, int nQ, const T* phi, const int* qImg, const T* scale, C* Sout) {
        const dim3 bs(L * N / 16);
        if (stage == 0)
            hipLaunchKernelGGL((k_field_rows<T, N, L>), dim3(nQ * (N / L)), bs, 0, st_, tiles, phi, invPio_.p,
                               pupil_.p, qImg, scale, rowNZ_.p, nPS_, tw_.p);
        else if (stage == 1)
            hipLaunchKernelGGL((k_cols_power<T, N, L>), dim3(nQ * (N / L)), bs, 0, st_, power, tiles, rowNZ_.p, qImg,
                               scale, tw_.p);
        else if (stage == 2)
            hipLaunchKernelGGL((k_rows_inv_real<T, N, L>), dim3(nQ * (N / (2 * L))), bs, 0, st_, tiles, power, qImg,
                               scale, tw_.p);
        else {
            const int chunks = (N / 2) / L + 1;
            hipLaunchKernelGGL((k_cols_extract<T, N, L>), dim3(nQ * chunks), bs, 0, st_, Sout, tiles, invOs_.p, nOS_,
                               (T)(1.0 / N2), qImg, scale, chunks, tw_.p);
        }
    }

    void otf_stage_l(int stage, int l, C* tiles, T* power, int nQ, const T* phi, const int* qImg, const T* scale,
                     C* Sout) {
        if constexpr (N == 256) {
            if (l == LINES / 4) return otf_stage<LINES / 4>(stage, tiles, power, nQ, phi, qImg, scale, Sout);
            if (l == LINES / 2) return otf_stage<LINES / 2>(stage, tiles, power, nQ, phi, qImg, scale, Sout);
        }
        otf_stage<LINES>(stage, tiles, power, nQ, phi, qImg, scale, Sout);
    }

    // Tile shape of each of the four OTF passes at N = 256, picked by timing every
    // candidate on 32 scratch images at the first OTF of the context (the geometry is known
    // by then), once per precision and device; with MFBD_FFT_LINES set, its rows / columns
    // shapes are used.  The shape does not change any result.
    void tune_otf_stages() {
        for (int st = 0; st < 4; ++st) stageL_[st] = (st == 0 || st == 2) ? lr_ : lc_;
        if constexpr (N == 256) {
            if (getenv("MFBD_FFT_LINES")) return;
            static std::mutex mu;
            static std::map<std::pair<int, int>, std::vector<int>> chosen;
            std::lock_guard<std::mutex> lock(mu);
            const auto key = std::make_pair((int)sizeof(T), device_);
            auto it = chosen.find(key);
            if (it == chosen.end()) {
                const int nT = 32;
                DBuf<C> tiles, S;
                DBuf<T> power, phi, one;
                DBuf<int> qz;
                tiles.alloc((size_t)nT * N2);
                power.alloc((size_t)nT * N2);
                phi.alloc((size_t)nT * nPS_);
                S.alloc((size_t)nT * nOS_);
                qz.alloc(nT);
                upload(one, std::vector<T>(1, T(1)), st_);
                HIPCHK(hipMemsetAsync(tiles.p, 0, (size_t)nT * N2 * sizeof(C), st_));
                HIPCHK(hipMemsetAsync(power.p, 0, (size_t)nT * N2 * sizeof(T), st_));
                HIPCHK(hipMemsetAsync(phi.p, 0, (size_t)nT * nPS_ * sizeof(T), st_));
                HIPCHK(hipMemsetAsync(qz.p, 0, nT * sizeof(int), st_));
                const int cand[3] = {LINES / 4, LINES / 2, LINES};
                std::vector<int> best(4, LINES);
                for (int st = 0; st < 4; ++st) {
                    float bestMs = 1e30f;
                    for (int l : cand) {
                        otf_stage_l(st, l, tiles.p, power.p, nT, phi.p, qz.p, one.p, S.p);      // warm-up
                        float ms = 1e30f;
                        for (int rep = 0; rep < 3; ++rep) {
                            HIPCHK(hipEventRecord(ev0_, st_));
                            otf_stage_l(st, l, tiles.p, power.p, nT, phi.p, qz.p, one.p, S.p);
                            HIPCHK(hipEventRecord(ev1_, st_));
                            HIPCHK(hipEventSynchronize(ev1_));
                            float t = 0;
                            HIPCHK(hipEventElapsedTime(&t, ev0_, ev1_));
                            ms = std::min(ms, t);
                        }
                        HIPCHK(hipGetLastError());
                        if (ms < bestMs) {
                            bestMs = ms;
                            best[st] = l;
                        }
                    }
                }
                HIPCHK(hipStreamSynchronize(st_));
                it = chosen.emplace(key, best).first;
            }
            for (int st = 0; st < 4; ++st) stageL_[st] = it->second[st];
        }
    }

    void fft2(C* tiles, int nTiles, int dir) {
        if (dir < 0) {
            launch("fft_rows", nTiles, [&] { fft_pass_any<-1>(tiles, nTiles, true, lr_); });
            launch("fft_cols", nTiles, [&] { fft_pass_any<-1>(tiles, nTiles, false, lc_); });
        } else {
            launch("fft_rows", nTiles, [&] { fft_pass_any<+1>(tiles, nTiles, true, lr_); });
            launch("fft_cols", nTiles, [&] { fft_pass_any<+1>(tiles, nTiles, false, lc_); });
        }
    }

    // Tile shape of the N = 256 transform passes (lines per workgroup: LINES/4, LINES/2 or
    // LINES, separately for rows and columns), picked once per precision and device by
    // timing each candidate on 32 scratch tiles; MFBD_FFT_LINES="rows,cols" fixes it.
    void autotune_fft() {
        lr_ = lc_ = LINES;
        if constexpr (N == 256) {
            const int cand[3] = {LINES / 4, LINES / 2, LINES};
            auto valid = [&](int l) { return l == cand[0] || l == cand[1] || l == cand[2]; };
            if (const char* e = getenv("MFBD_FFT_LINES")) {
                int a = 0, b = 0;
                if (sscanf(e, "%d,%d", &a, &b) != 2 || !valid(a) || !valid(b))
                    throw std::runtime_error("MFBD_FFT_LINES must be two of " + std::to_string(cand[0]) + ", " +
                                             std::to_string(cand[1]) + ", " + std::to_string(cand[2]));
                lr_ = a;
                lc_ = b;
                return;
            }
            static std::mutex mu;
            static std::map<std::pair<int, int>, std::pair<int, int>> chosen;   // (bytes of T, device) -> (rows, cols)
            std::lock_guard<std::mutex> lock(mu);
            const auto key = std::make_pair((int)sizeof(T), device_);
            auto it = chosen.find(key);
            if (it == chosen.end()) {
                const int nT = 32;
                DBuf<C> scratch;
                scratch.alloc((size_t)nT * N2);
                HIPCHK(hipMemsetAsync(scratch.p, 0, (size_t)nT * N2 * sizeof(C), st_));
                std::pair<int, int> best{LINES, LINES};
                for (int pass = 0; pass < 2; ++pass) {
                    float bestMs = 1e30f;
                    for (int l : cand) {
                        fft_pass_any<-1>(scratch.p, nT, pass == 0, l);      // warm-up
                        float ms = 1e30f;
                        for (int rep = 0; rep < 3; ++rep) {
                            HIPCHK(hipEventRecord(ev0_, st_));
                            fft_pass_any<-1>(scratch.p, nT, pass == 0, l);
                            HIPCHK(hipEventRecord(ev1_, st_));
                            HIPCHK(hipEventSynchronize(ev1_));
                            float t = 0;
                            HIPCHK(hipEventElapsedTime(&t, ev0_, ev1_));
                            ms = std::min(ms, t);
                        }
                        HIPCHK(hipGetLastError());
                        if (ms < bestMs) {
                            bestMs = ms;
                            (pass == 0 ? best.first : best.second) = l;
                        }
                    }
                }
                HIPCHK(hipStreamSynchronize(st_));
                it = chosen.emplace(key, best).first;
            }
            lr_ = it->second.first;
            lc_ = it->second.second;
        }
    }

    // OTF chain for nQ virtual images whose phases are in `phi`: S = reorder(IFFT(|FFT(a)|^2))/N^2
    // line-search probes read by the fused FP32 transform: line buffers, per image the line
    // row (-1: phases from phi) and the probe step, the direction's offset
    struct LineIn {
        const T* line;
        const int* qLine;
        const double* qS;
        size_t dirOff;
    };

    void otf_chain(const T* phi, const int* qImg, int nQ, C* Sout, LineIn lin = {}) {
        if constexpr (FUSED && N == 128) {
            if (otfWave_) {
                launch("k_otf_wave", nQ, [&] {
                    hipLaunchKernelGGL((k_otf_wave<FNT>), dim3(std::min(nQ, otfGrid_)), dim3(FNT), WLDS, st_, Sout, phi,
                                       pupilG_.p, qImg, imgScale_.p, hsrc_.p, psd_.p, tw_.p, nPS_, nOS_, nQ, lin.line,
                                       lin.qLine, lin.qS, lin.dirOff);
                });
                return;
            }
        }
        if constexpr (FUSED) {
            launch("k_otf_fused", nQ, [&] {
                hipLaunchKernelGGL((k_otf_fused<N, FNT>), dim3(std::min(nQ, otfGrid_)), dim3(FNT), OLDS, st_, Sout, phi,
                                   invPs_.p, pupil_.p, qImg, imgScale_.p, hread_.p, tw_.p, nPS_, nOS_, nQ, lin.line,
                                   lin.qLine, lin.qS, lin.dirOff);
            });
            return;
        }
        if constexpr (FUSED64) {
            launch("k_otf64_fused", nQ, [&] {
                if (lin.line)                  // line-search probes from their lines' phases
                    hipLaunchKernelGGL((k_otf64_fused<N, F64NT, true>), dim3(std::min(nQ, otfGrid_)), dim3(F64NT),
                                       F64LDS, st_, Sout, phi, invPs_.p, pupil_.p, qImg, imgScale_.p, hread_.p, tw_.p,
                                       nPS_, nOS_, nQ, nullptr, lin.line, lin.qLine, lin.qS, lin.dirOff);
                else
                    hipLaunchKernelGGL((k_otf64_fused<N, F64NT, false>), dim3(std::min(nQ, otfGrid_)), dim3(F64NT),
                                       F64LDS, st_, Sout, phi, invPs_.p, pupil_.p, qImg, imgScale_.p, hread_.p, tw_.p,
                                       nPS_, nOS_, nQ);
            });
            return;
        }
        if (!stageL_[0]) {
            const hipStream_t keep = st_;
            st_ = lanes_[kSyncLane].stream;
            tune_otf_stages();
            st_ = keep;
        }
        sc_->tiles.alloc((size_t)nQ * N2);
        sc_->power.alloc((size_t)nQ * N2);
        static const char* names[4] = {"k_field_rows", "k_cols_power", "k_rows_inv_real", "k_cols_extract"};
        for (int stage = 0; stage < 4; ++stage)
            launch(names[stage], nQ, [&] {
                otf_stage_l(stage, stageL_[stage], sc_->tiles.p, sc_->power.p, nQ, phi, qImg, imgScale_.p, Sout);
            });
    }

    // phases, OTFs, P/Q and the metric for n requests (virtual images q = r*nImg + j).
    // Alpha rows come from the slot-indexed pinned staging when every request is
    // staged there (the solver's case), else they are packed per request.
    // needPQ false: only the metric partials are kept (no request reads P and Q afterwards)
    void forward(EvalReq* rq, int n, int useGamma, const std::vector<int>* xi = nullptr,
                 const std::vector<double>* xd = nullptr, bool needPQ = true) {
        const int K = g_.K, nImg = g_.nImg, nQ = n * nImg;
        const size_t row = (size_t)nImg * K;
        bool staged = true;
        int lo = maxSlots_, hi = -1, nStart = 0, nLine = 0, nSave = 0;
        for (int r = 0; r < n; ++r) {
            staged &= rq[r].alpha == hStage_.p + (size_t)rq[r].slot * row;
            lo = std::min(lo, rq[r].slot);
            hi = std::max(hi, rq[r].slot);
            nStart += rq[r].line == 1;
            nLine += rq[r].line != 0;
            nSave += rq[r].line == 0 && rq[r].saveBase;
            if (rq[r].line == 1 && !rq[r].dAlpha) throw std::runtime_error("forward: line start without a direction");
        }
        const bool rowsMode = nLine || nSave;      // the GEMM through row tables (lines in the slots' buffers)
        if (rowsMode) {      // a line lives in its slot's buffers: one request per slot
            std::vector<char> seen(maxSlots_, 0);
            for (int r = 0; r < n; ++r) {
                if (seen[rq[r].slot]) throw std::runtime_error("forward: line requests need distinct slots");
                seen[rq[r].slot] = 1;
            }
            lineBuf_.alloc((size_t)2 * maxSlots_ * nImg * nPS_);
        }
        // one pinned block and one copy per evaluation: the caller's doubles xd, with line
        // rows each image's probe step (nQ), then the request tables (7 nQ + n ints), the
        // caller's ints xi, the GEMM's row tables (5 per GEMM row), and in rows mode the
        // line rows' (2) and each image's line row (nQ)
        const size_t nd = xd ? xd->size() : 0, nx = xi ? xi->size() : 0;
        const int nLR = nLine * nImg, nGemm = nQ - nLR + nStart * nImg;
        const size_t nd2 = nd + (nLine ? (size_t)nQ : 0);
        const size_t nInts = 7 * (size_t)nQ + n + nx + 5 * (size_t)nGemm + (rowsMode ? 2 * (size_t)nLR + nQ : 0);
        const size_t nBlk = nd2 + (nInts + 1) / 2;
        sc_->hTab.alloc(nBlk);
        if (nd) std::memcpy(sc_->hTab.p, xd->data(), nd * sizeof(double));
        int* ints = reinterpret_cast<int*>(sc_->hTab.p + nd2);
        for (int r = 0; r < n; ++r) {
            for (int j = 0; j < nImg; ++j) {
                const int q = r * nImg + j;
                ints[q] = rq[r].slot;                                          // slot
                ints[nQ + q] = j;                                              // image
                ints[2 * nQ + q] = staged ? (rq[r].slot - lo) * nImg + j : q;  // alpha row
                ints[3 * nQ + q] = r;                                          // request
                ints[4 * nQ + n + q] = rq[r].slot * g_.nCh + st.imgCh[j];     // phiCh row
                ints[5 * nQ + n + q] = st.imgObj[j];                           // object
                ints[6 * nQ + n + q] = rq[r].slot * nImg + j;                  // pupil scale row
            }
            ints[4 * nQ + r] = rq[r].slot;
        }
        if (nx) std::memcpy(ints + 7 * (size_t)nQ + n, xi->data(), nx * sizeof(int));
        const size_t aRows = staged ? (size_t)(hi - lo + 1) * nImg : (size_t)nQ;   // alpha rows before the directions
        const int LS = maxSlots_ * nImg;
        {
            // GEMM rows: phases of the requests without a line to the work array (row q), and
            // of a gradient request that starts a conjugate-gradient iteration also to its
            // slot's line base; for each line start the direction B dAlpha (no channel phase)
            // to the slot's line direction.  The rows go to the GEMM grouped by object (in
            // order within an object), so a wave's 16 rows rarely straddle two objects (each
            // object in a wave costs a pass over K); every row's sum is the same either way.
            // Line rows: the probes' phases from the line.
            int* gA = ints + 7 * (size_t)nQ + n + nx;
            int* gP = gA + nGemm;
            int* gO = gP + nGemm;
            int* gOut = gO + nGemm;
            int* gOut2 = gOut + nGemm;
            int* lq = gOut2 + nGemm;
            int* lrow = lq + nLR;
            int* qLine = lrow + nLR;
            double* qS = sc_->hTab.p + nd;
            struct GRow { int a, p, o, out, out2; };
            std::vector<GRow> rows;
            rows.reserve(nGemm);
            int d = 0, l = 0;
            for (int r = 0; r < n; ++r)
                for (int j = 0; j < nImg; ++j) {
                    const int q = r * nImg + j, srow = rq[r].slot * nImg + j;
                    if (rq[r].line == 0)
                        rows.push_back({ints[2 * nQ + q], ints[4 * nQ + n + q], st.imgObj[j], q,
                                        rq[r].saveBase ? srow : -1});
                    else if (rq[r].line == 1)
                        rows.push_back({(int)aRows + d * nImg + j, -1, st.imgObj[j], -1 - (LS + srow), -1});
                    if (rowsMode) {
                        qLine[q] = rq[r].line ? srow : -1;
                        if (nLine) qS[q] = rq[r].line ? rq[r].lineS : 0.0;
                        if (rq[r].line) {
                            lq[l] = q;
                            lrow[l] = srow;
                            ++l;
                        }
                    }
                    if (rq[r].line == 1 && j == nImg - 1) ++d;
                }
            if ((int)rows.size() != nGemm || l != nLR) throw std::runtime_error("forward: line tables");
            std::stable_sort(rows.begin(), rows.end(), [](const GRow& x, const GRow& y) { return x.o < y.o; });
            for (int g = 0; g < nGemm; ++g) {
                gA[g] = rows[g].a;
                gP[g] = rows[g].p;
                gO[g] = rows[g].o;
                gOut[g] = rows[g].out;
                gOut2[g] = rows[g].out2;
            }
            if (nStart) {
                sc_->hDir.alloc(std::max<size_t>((size_t)nStart * row, 1));
                d = 0;
                for (int r = 0; r < n; ++r)
                    if (rq[r].line == 1) std::memcpy(sc_->hDir.p + (size_t)d++ * row, rq[r].dAlpha, sizeof(double) * row);
            }
        }
        const size_t aAll = aRows + (size_t)nStart * nImg;
        sc_->alpha.alloc(std::max(aAll * K, sc_->alpha.n));
        if (staged) {
            HIPCHK(hipMemcpyAsync(sc_->alpha.p, hStage_.p + (size_t)lo * row, aRows * K * sizeof(double),
                                  hipMemcpyHostToDevice, st_));
        } else {
            sc_->hAlpha.alloc((size_t)nQ * K);
            for (int r = 0; r < n; ++r) std::memcpy(sc_->hAlpha.p + (size_t)r * row, rq[r].alpha, sizeof(double) * row);
            HIPCHK(hipMemcpyAsync(sc_->alpha.p, sc_->hAlpha.p, (size_t)nQ * K * sizeof(double), hipMemcpyHostToDevice, st_));
        }
        if (nStart)
            HIPCHK(hipMemcpyAsync(sc_->alpha.p + aRows * K, sc_->hDir.p, (size_t)nStart * row * sizeof(double),
                                  hipMemcpyHostToDevice, st_));
        sc_->tab.alloc(nBlk);
        HIPCHK(hipMemcpyAsync(sc_->tab.p, sc_->hTab.p, nBlk * sizeof(double), hipMemcpyHostToDevice, st_));
        const int* ti = reinterpret_cast<const int*>(sc_->tab.p + nd2);
        sc_->qm = {ti, ti + nQ, ti + 2 * nQ};
        sc_->qReq = ti + 3 * nQ;
        sc_->reqSlot = ti + 4 * nQ;
        const int* qScale = ti + 6 * nQ + n;
        sc_->tx = ti + 7 * nQ + n;
        sc_->td = sc_->tab.p;
        sc_->phi.alloc((size_t)nQ * nPS_);
        sc_->S.alloc((size_t)nQ * nOS_);
        // the mode GEMM over its row tables (grouped by object)
        const int* gA = sc_->tx + nx;
        const int* gP = gA + nGemm;
        const int* gO = gP + nGemm;
        const int* gOut = gO + nGemm;
        const int* gOut2 = gOut + nGemm;
        const int* lq = rowsMode ? gOut2 + nGemm : nullptr;
        const int* qLineD = rowsMode ? lq + 2 * nLR : nullptr;
        const int rows = nGemm;
        T* lineBuf = rowsMode ? lineBuf_.p : nullptr;
        if (rows > 0) {
            if (bf16Phase_) {
                if constexpr (std::is_same_v<T, float>) {
                    launch("k_phase_bf16", rows, [&] {
                        hipLaunchKernelGGL(k_phase_bf16, dim3((nPS_ + 63) / 64, (rows + 63) / 64), dim3(256), 0, st_,
                                           sc_->phi.p, reinterpret_cast<const bf16x8*>(Bt_.p), phiCh_.p, sc_->alpha.p,
                                           gA, gP, gO, rows, K, (K + 31) & ~31, nPS_, lineBuf, gOut, gOut2);
                    });
                }
            } else {
                launch("k_phase_mfma", rows, [&] {
                    hipLaunchKernelGGL(k_phase_mfma<T>, dim3((nPS_ + 63) / 64, (rows + 63) / 64), dim3(256), 0, st_,
                                       sc_->phi.p, B_.p, phiCh_.p, sc_->alpha.p, gA, gP, gO, rows, K, nPS_, lineBuf,
                                       gOut, gOut2);
                });
            }
        }
        const double* qSD = nLine ? sc_->tab.p + nd : nullptr;
        LineIn lin{};
        if (nLine) {
            if constexpr (FUSED || FUSED64) {  // the fused transforms read the line in place of phi
                lin = {lineBuf_.p, qLineD, qSD, (size_t)maxSlots_ * nImg * nPS_};
            } else {
                launch("k_line_phase", nLR, [&] {
                    hipLaunchKernelGGL(k_line_phase<T>, dim3(nblk((size_t)nLR * nPS_)), dim3(256), 0, st_, sc_->phi.p,
                                       lineBuf_.p, lq, lq + nLR, qSD, nLR, nPS_, (size_t)maxSlots_ * nImg * nPS_);
                });
            }
        }
        otf_chain(sc_->phi.p, qScale, nQ, sc_->S.p, lin);
        const int nChunk = (nOS_ + 255) / 256;
        sc_->Pp.alloc((size_t)n * g_.nObj * nOS_ * 2);
        sc_->Qp.alloc((size_t)n * g_.nObj * nOS_);
        sc_->part.alloc((size_t)n * g_.nObj * nChunk);
        launch("k_pq", nQ, [&] {
            hipLaunchKernelGGL(k_pq<T>, dim3(nChunk, g_.nObj, n), dim3(256), 0, st_, needPQ ? sc_->Pp.p : nullptr,
                               needPQ ? sc_->Qp.p : nullptr, sc_->part.p, sc_->S.p, D_.p,
                               ftSum_.p, gamma_.p, sc_->reqSlot, chObj_.p, chFirst_.p, chCount_.p, wt_.p, imgScale_.p,
                               g_.nCh, g_.nObj, nImg, nOS_, useGamma);
        });
    }

    // host vector -> pinned staging -> device, without waiting (the lane's staging is not
    // reused before its evaluation has finished)
    template <typename U>
    void stage_async(DBuf<U>& d, HBuf<U>& h, const std::vector<U>& v) {
        h.alloc(std::max<size_t>(v.size(), 1));
        std::copy(v.begin(), v.end(), h.p);
        d.alloc(std::max<size_t>(v.size(), 1));
        if (!v.empty()) HIPCHK(hipMemcpyAsync(d.p, h.p, v.size() * sizeof(U), hipMemcpyHostToDevice, st_));
    }

    void use_lane(int lane) {
        lane_ = lane;
        sc_ = &lanes_[lane];
        st_ = sc_->stream;
    }

    // Gradient tables of the requests that want a Vogel / adjoint gradient: their virtual
    // images (gIdx, request order), enabled-mode counts (gEn) and output scales (gs)
    void grad_tables(const EvalReq* rq, int n, int nImg, std::vector<int>& gIdx, std::vector<int>& gEn,
                     std::vector<double>& gs, int& nEn, int& extraPupil) {
        nEn = 0;
        extraPupil = 1;
        for (int r = 0; r < n; ++r) {
            if (!rq[r].wantGrad || rq[r].gradMethod == 1) continue;
            nEn = std::max(nEn, rq[r].nEnabled);
            for (int j = 0; j < nImg; ++j) {
                if (!imgOn_[(size_t)rq[r].slot * nImg + j]) continue;   // its gradient row is zero
                const int o = st.imgObj[j];
                const double w = st.weight[o];
                gIdx.push_back(r * nImg + j);
                gEn.push_back(rq[r].nEnabled);
                if (rq[r].gradMethod == 2) {
                    // gradientVogel: -2 area/N^2 * w, times w again in Solver::gradient
                    gs.push_back((w == 0.0) ? 0.0 : -2.0 * area_ / N2 * w * w);
                } else {
                    // exact dL/dalpha = -(4 c^2 / s^2) w sum Im(conj(a) (G * a)), t = (G * PF)/N^2
                    const double c = otfNorm_[st.imgCh[j]];
                    gs.push_back(-4.0 * c * c * N2 / ((double)g_.s * g_.s) * w);
                    extraPupil = 0;
                }
            }
        }
    }

    // One chunk of requests: one table upload, the forward chain, the metric, the
    // gradients, and one download of [L (n) | Vogel / adjoint gradient rows (nG x K)]
    void eval_chunk_launch(EvalReq* rq, int n, int K, int nImg) {
        std::vector<int> gIdx, gEn;
        std::vector<double> gs;
        int nEn = 0, extraPupil = 1;
        grad_tables(rq, n, nImg, gIdx, gEn, gs, nEn, extraPupil);
        const int nG = (int)gIdx.size();
        gIdx.insert(gIdx.end(), gEn.begin(), gEn.end());       // [0, nG): images, [nG, 2nG): counts
        bool anyFd = false;
        for (int r = 0; r < n; ++r) anyFd |= rq[r].wantGrad && rq[r].gradMethod == 1;
        forward(rq, n, 1, &gIdx, &gs, nG > 0 || anyFd);        // P and Q for the gradients only
        const int nChunk = (nOS_ + 255) / 256;
        const size_t nOut = (size_t)n + (size_t)nG * K;
        sc_->out.alloc(nOut);
        launch("k_metric", [&] {
            hipLaunchKernelGGL(k_metric, dim3(nblk(n, 64)), dim3(64), 0, st_, sc_->out.p, sc_->part.p, w_.p, n, g_.nObj,
                               nChunk, 1.0 / ((double)g_.s * g_.s));
        });
        if (nG) vogel_gradients(n, K, nImg, nG, nEn, extraPupil);
        if (anyFd) fd_gradient(rq, n, K, nImg);
        sc_->hOut.alloc(nOut);
        HIPCHK(hipMemcpyAsync(sc_->hOut.p, sc_->out.p, nOut * sizeof(double), hipMemcpyDeviceToHost, st_));
        HIPCHK(hipEventRecord(sc_->done, st_));
    }

    void eval_chunk_finish(EvalReq* rq, int n, int K, int nImg) {
        size_t row = 0;                                          // Vogel / adjoint gradient rows, request order
        for (int r = 0; r < n; ++r) {
            rq[r].f = sc_->hOut.p[r];
            if (!rq[r].wantGrad || rq[r].gradMethod == 1) continue;   // finite differences: filled already
            const unsigned char* on = &imgOn_[(size_t)rq[r].slot * nImg];
            for (int j = 0; j < nImg; ++j) {
                if (!on[j]) {               // an image the patch lacks has no gradient row
                    if (rq[r].g) std::fill(rq[r].g + (size_t)j * K, rq[r].g + (size_t)(j + 1) * K, 0.0);
                    continue;
                }
                if (rq[r].g) std::memcpy(rq[r].g + (size_t)j * K, sc_->hOut.p + n + row * K, sizeof(double) * K);
                row++;
            }
        }
    }

    // Vogel / adjoint weights and their mode projection for the nG gradient images whose
    // tables forward() uploaded (sc_->tx: gIdx then counts, sc_->td: scales)
    void vogel_gradients(int n, int K, int nImg, int nG, int nEn, int extraPupil) {
        const int* gIdxD = sc_->tx;
        {
            sc_->v.alloc((size_t)nG * nPS_);
            if constexpr (FUSED) {
                launch("k_vogel_fused", nG, [&] {
                    hipLaunchKernelGGL((k_vogel_fused<N, FNT>), dim3(nG), dim3(FNT), FLDS, st_, sc_->v.p, sc_->S.p, D_.p, sc_->Pp.p,
                                       sc_->Qp.p, sc_->phi.p, pyx_.p + 2 * nPS_, pyx_.p + 3 * nPS_, pupil_.p, gwt_.p,
                                       sc_->qm, sc_->qReq, imgObj_.p, gIdxD, tw_.p, nPS_, nOS_, nImg, g_.nObj,
                                       extraPupil);
                });
            } else if constexpr (FUSED64) {
                sc_->greal.alloc((size_t)nG * N2);
                launch("k_grad64_fused", nG, [&] {
                    hipLaunchKernelGGL((k_grad64_fused<N, F64NT>), dim3(std::min(nG, otfGrid_)), dim3(F64NT), G64LDS,
                                       st_, sc_->v.p, sc_->greal.p, sc_->S.p, D_.p, sc_->Pp.p, sc_->Qp.p, gw_.p,
                                       sc_->qm, sc_->qReq, imgObj_.p, gIdxD, sc_->phi.p, invPs_.p, pupil_.p,
                                       tw_.p, nPS_, nOS_, nImg, g_.nObj, nG, extraPupil);
                });
            } else {
                sc_->tiles.alloc((size_t)nG * N2);
                sc_->greal.alloc((size_t)nG * N2);
                launch("k_build_g", nG, [&] {
                    hipLaunchKernelGGL(k_build_g<T>, dim3(nblk((size_t)nG * N2)), dim3(256), 0, st_, sc_->tiles.p, sc_->S.p,
                                       D_.p, sc_->Pp.p, sc_->Qp.p, invOs_.p, sc_->qm, sc_->qReq, imgObj_.p, gIdxD, nG, N2, nOS_, nImg,
                                       g_.nObj);
                });
                fft2(sc_->tiles.p, nG, +1);
                launch("k_save_re", nG, [&] {
                    hipLaunchKernelGGL(k_save_re<T>, dim3(nblk((size_t)nG * N2)), dim3(256), 0, st_, sc_->greal.p,
                                       sc_->tiles.p, (size_t)nG * N2, (T)(1.0 / N2));
                });
                launch("k_field", nG, [&] {
                    hipLaunchKernelGGL(k_field<T>, dim3(nblk((size_t)nG * N2)), dim3(256), 0, st_, sc_->tiles.p, sc_->phi.p,
                                       invPde_.p, pupil_.p, sc_->qm.img, (const T*)nullptr, gIdxD, nG, N2, nPS_);
                });
                fft2(sc_->tiles.p, nG, +1);
                launch("k_mul_re", nG, [&] {
                    hipLaunchKernelGGL(k_mul_re<T>, dim3(nblk((size_t)nG * N2)), dim3(256), 0, st_, sc_->tiles.p,
                                       sc_->greal.p, (size_t)nG * N2, (T)(1.0 / N2));
                });
                fft2(sc_->tiles.p, nG, -1);
                launch("k_extract_v", nG, [&] {
                    hipLaunchKernelGGL(k_extract_v<T>, dim3(nblk((size_t)nG * nPS_)), dim3(256), 0, st_, sc_->v.p,
                                       sc_->tiles.p, sc_->phi.p, pde_.p, pupil_.p, gIdxD, nG, nPS_, N2, extraPupil);
                });
            }
            {
                const int Kt = std::min(K, (nEn + 15) & ~15);      // mode columns actually computed
                sc_->gpart.alloc((size_t)GR_SLICES * nG * std::max(Kt, 1));
                launch("k_grad_mfma", nG, [&] {
                    hipLaunchKernelGGL(k_grad_mfma<T>, dim3((nG + 15) / 16, GR_SLICES), dim3(256), 0, st_, sc_->gpart.p,
                                       sc_->v.p, B_.p, sc_->qm, imgObj_.p, gIdxD, nG, K, Kt, nPS_, nEn);
                });
                launch("k_grad_sum", [&] {
                    hipLaunchKernelGGL(k_grad_sum, dim3(nblk((size_t)nG * K)), dim3(256), 0, st_, sc_->out.p + n,
                                       sc_->gpart.p, sc_->td, gIdxD + nG, nG, K, Kt);
                });
            }
        }
    }

    void fd_gradient(EvalReq* rq, int n, int K, int nImg) {
        // probes (request r, image j, mode m < nEnabled): phi + step * B_m
        std::vector<int> base, mode, obj, pimg;
        std::vector<T> step;
        std::vector<double> wgt;
        std::vector<std::pair<int, int>> dst;   // (alpha row, mode)
        for (int r = 0; r < n; ++r) {
            if (!rq[r].wantGrad || rq[r].gradMethod != 1) continue;
            for (int j = 0; j < nImg; ++j) {
                const int o = st.imgObj[j];
                if (!imgOn_[(size_t)rq[r].slot * nImg + j]) continue;   // an image the patch lacks
                for (int m = 0; m < rq[r].nEnabled; ++m) {
                    base.push_back(r * nImg + j);
                    mode.push_back(m);
                    obj.push_back(o);
                    step.push_back((T)(st.graddiffStep * st.wavelength[o]));
                    pimg.push_back(rq[r].slot * nImg + j);
                    wgt.push_back(st.weight[o]);
                    dst.push_back({r * nImg + j, m});
                }
            }
        }
        const int nP = (int)base.size();
        std::vector<double> res(nP);
        const int chunk = maxTiles();
        // probe tables live after the request tables in a second integer buffer
        for (int b = 0; b < nP; b += chunk) {
            const int m = std::min(chunk, nP - b);
            std::vector<int> ints;
            ints.insert(ints.end(), base.begin() + b, base.begin() + b + m);
            ints.insert(ints.end(), mode.begin() + b, mode.begin() + b + m);
            ints.insert(ints.end(), obj.begin() + b, obj.begin() + b + m);
            ints.insert(ints.end(), pimg.begin() + b, pimg.begin() + b + m);
            upload(sc_->fdInts, ints, st_);
            upload(sc_->fdStep, step.data() + b, m, st_);
            sc_->phiFd.alloc((size_t)m * nPS_);
            sc_->Sfd.alloc((size_t)m * nOS_);
            sc_->fdOut.alloc(m);
            launch("k_phase_pert", [&] {
                hipLaunchKernelGGL(k_phase_pert<T>, dim3(nblk((size_t)m * nPS_)), dim3(256), 0, st_, sc_->phiFd.p, sc_->phi.p,
                                   B_.p, sc_->fdInts.p, sc_->fdInts.p + m, sc_->fdInts.p + 2 * m, sc_->fdStep.p, m, K, nPS_);
            });
            otf_chain(sc_->phiFd.p, sc_->fdInts.p + 3 * m, m, sc_->Sfd.p);
            launch("k_metric_change", [&] {
                hipLaunchKernelGGL(k_metric_change<T>, dim3(m), dim3(256), 0, st_, sc_->fdOut.p, sc_->Sfd.p, sc_->S.p, D_.p, sc_->Pp.p,
                                   sc_->Qp.p, sc_->fdInts.p, sc_->qm, sc_->qReq, imgObj_.p, wt_.p, nOS_, nImg, g_.nObj, 1.0 / N2);
            });
            HIPCHK(hipMemcpyAsync(res.data() + b, sc_->fdOut.p, m * sizeof(double), hipMemcpyDeviceToHost, st_));
        }
        HIPCHK(hipStreamSynchronize(st_));
        // g = w * (dl / step * w) for the probed modes, 0 for the others, straight into the
        // requests' gradient arrays (their patches are waiting for this evaluation)
        for (int r = 0; r < n; ++r)
            if (rq[r].wantGrad && rq[r].gradMethod == 1 && rq[r].g) std::fill(rq[r].g, rq[r].g + (size_t)nImg * K, 0.0);
        for (int k = 0; k < nP; ++k) {
            const double stp = (double)step[k];
            const double w = wgt[k];
            const int r = dst[k].first / nImg, j = dst[k].first % nImg;
            if (rq[r].g) rq[r].g[(size_t)j * K + dst[k].second] = (w == 0.0) ? 0.0 : w * (res[k] / stp * w);
        }
    }

    std::vector<int> otf_support_of(const double* pupil) {
        // autocorrelation of the pupil placed in the lower-left P x P block, centred,
        // computed on the host in double (same threshold as the reference)
        const int P = g_.P;
        std::vector<double> ac((size_t)N2, 0.0);
        for (int dy = -(P - 1); dy <= P - 1; ++dy)
            for (int dx = -(P - 1); dx <= P - 1; ++dx) {
                double acc = 0;
                for (int y = std::max(0, -dy); y < std::min(P, P - dy); ++y)
                    for (int x = std::max(0, -dx); x < std::min(P, P - dx); ++x)
                        acc += pupil[(y + dy) * P + x + dx] * pupil[y * P + x];
                ac[(size_t)(dy + N / 2) * N + dx + N / 2] = acc;
            }
        std::vector<int> os;
        for (int i = 0; i < N2; ++i)
            if (std::fabs(ac[i]) > 1e-9) os.push_back(i);
        return os;
    }

public:
    Structure st;

private:
    Geometry g_;
    int maxSlots_, maxBatch_;
    int otfGrid_ = 1;                        // workgroups of the persistent FP64 OTF kernel
    std::array<hipStream_t, kSyncLane + 1> streamSet_{};     // from stream_sets(), given back at close
    bool otfWave_ = false;                   // FP32, N = 128: k_otf_wave in place of k_otf_fused
    int restoreN_ = 0;                       // slots of the restoration in flight on the synchronous lane
    size_t restoreOut_ = 0;                  // its output floats
    hipStream_t st_;
    hipEvent_t ev0_, ev1_;
    bool timing_ = false;
    int tileCap_ = 0;            // MFBD_MAX_TILES (0: no cap beyond the 4 GiB bound)
    std::map<std::string, KernelStat> stats_;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> evPool_;
    struct PendingEv {
        const char* name;
        std::pair<hipEvent_t, hipEvent_t> ev;
        long long units;
        int lane;
    };
    std::vector<PendingEv> evPending_;
    std::vector<int> ps_, os_;
    int nPS_ = 0, nOS_ = 0, nOSfull_ = 0;   // nOS_: stored (half) support, nOSfull_: full support
    std::vector<int> fullMap_;               // full support point -> half index h, or -1-h for a mirror
    double area_ = 0;
    std::vector<double> otfNorm_;
    // resident: basis, supports, per-patch frames, D, ftSum, gamma, channel phases
    DBuf<C> tw_, D_;
    DBuf<cx<double>> tw64_;                  // FP64 twiddles of k_init64_fused in the FP32 forms
    DBuf<C> xT_, xO_;                  // optional restoration outputs: image and object tiles
    DBuf<double> xD_[3];               // PSFs, convolved objects, windowed frames [images][s][s]
    DBuf<float> xF_[3];                // the float outputs: PSFs (or their averages), cobj, res
    HBuf<float> hF_;
    DBuf<int> xInts_;
    DBuf<T> B_, phiCh_, pupil_, imgScale_;
    DBuf<T> lineBuf_;                 // [2][maxSlots][nImg][nPS]: each slot's line base phases, then directions
    // the restoration: grid codes, FP64 twiddles, per (patch, object) grids and outputs
    DBuf<int> rsCode_;
    DBuf<cx<double>> twD_, rsAo_, rsT_;
    DBuf<double> rsDn_, rsA_, rsB_, rsNf_, rsC2_;
    DBuf<float> rsOut_;
    HBuf<float> hRs_;
    int device_ = 0;
    int lr_ = LINES, lc_ = LINES;           // lines per workgroup of the transform passes
    int stageL_[4] = {0, 0, 0, 0};          // ... of the four OTF passes at N = 256 (0: not yet tuned)
    std::vector<T> hScale_;                 // host copy of imgScale_ [maxSlots][nImg]
    std::vector<unsigned char> imgOn_;      // images each slot's patch has [maxSlots][nImg]
    DBuf<__bf16> Bt_;                 // BF16 basis, transposed, for the mixed-precision form
    bool bf16Phase_ = false;
    HBuf<double> hStage_;
    DBuf<double> ftSum_, gamma_, win_, w_, wt_;
    std::vector<double> hGamma_;             // host copy of gamma_ (slot, object)
    DBuf<int> gw_;
    DBuf<unsigned char> rowNZ_;             // natural rows of the OTF grid the pupil field touches
    DBuf<int> invPio_, invPde_, osNat_, invOs_, pde_, chObj_, chFirst_, chCount_, imgObj_, imgCh_, pyx_, gwt_, hread_,
        invPs_, hsrc_, psd_;
    DBuf<float> rawF_, pupilG_;
    bool winSet_ = false;
    // per-round work buffers: the evaluation lanes take contiguous shares of a batch
    // (evaluate_begin / evaluate_end), kSyncLane serves the synchronous calls
    struct Scratch {
        DBuf<C> tiles, S, Sfd;
        DBuf<T> phi, phiFd, v, greal, fdStep, power;
        DBuf<double> Pp, Qp, part, alpha, fdOut, statOut, gpart, tab, out;
        DBuf<int> ints, fdInts, initInts;
        HBuf<double> hAlpha, hL, hG, hTab, hOut;   // hTab: the evaluation's tables, hOut: its results
        HBuf<double> hDir;                         // line starts' directions in alpha
        HBuf<int> hInit;
        QMap qm{};                       // device tables of the current evaluation (in tab)
        const int* qReq = nullptr;
        const int* reqSlot = nullptr;
        const int* tx = nullptr;         // the caller's ints after the request tables
        const double* td = nullptr;      // the caller's doubles before them
        hipEvent_t done = nullptr;       // recorded after the lane's result copies
        hipStream_t stream = nullptr;
        int nReq = 0;
    };
    Scratch lanes_[kSyncLane + 1];
    Scratch* sc_ = &lanes_[kSyncLane];
    int lane_ = 2;
    int laneFirst_[kSyncLane + 1] = {};
};

std::unique_ptr<DeviceEvaluator> make_device_evaluator(const Geometry& g, int precision, int maxSlots, int maxBatch,
                                                       int device, std::string& err) {
    try {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device) {
            err = "no HIP device available";
            return nullptr;
        }
        if (precision == 0) {
            if (g.N == 64) return std::make_unique<Pipeline<double, 64>>(g, maxSlots, maxBatch, device);
            if (g.N == 128) return std::make_unique<Pipeline<double, 128>>(g, maxSlots, maxBatch, device);
            if (g.N == 256) return std::make_unique<Pipeline<double, 256>>(g, maxSlots, maxBatch, device);
        } else {
            const bool bf = precision == 2;      // MFBD_MIXED: BF16 mode GEMM, FP32 FFTs
            if (g.N == 64) return std::make_unique<Pipeline<float, 64>>(g, maxSlots, maxBatch, device, bf);
            if (g.N == 128) return std::make_unique<Pipeline<float, 128>>(g, maxSlots, maxBatch, device, bf);
            if (g.N == 256) return std::make_unique<Pipeline<float, 256>>(g, maxSlots, maxBatch, device, bf);
        }
        err = "unsupported OTF grid (N must be 64, 128 or 256)";
    } catch (const std::exception& e) {
        err = e.what();
    }
    return nullptr;
}

}  // namespace mfbd
