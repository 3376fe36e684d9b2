// Host-side FFT used outside the hot loop: pupil-support autocorrelation,
// the per-frame noise estimate on the (possibly non power-of-two) patch, and the
// final inverse transform of the restored object.  FFTW conventions: forward
// e^{-i}, backward e^{+i}, unnormalised.
#pragma once

#include <cmath>
#include <complex>
#include <vector>

namespace mfbd {

using cd = std::complex<double>;

class HostFFT {
public:
    explicit HostFFT(int n = 0) { plan(n); }

    void plan(int n) {
        n_ = n;
        if (n <= 0) return;
        tw_.resize(n);
        for (int k = 0; k < n; ++k) {
            double a = -2.0 * M_PI * double(k) / double(n);
            tw_[k] = cd(std::cos(a), std::sin(a));
        }
        factors_.clear();
        int m = n;
        for (int f : {4, 2, 3, 5, 7, 11, 13}) {
            while (m % f == 0) { factors_.push_back(f); m /= f; }
        }
        if (m > 1) factors_.push_back(m);   // remaining prime: direct butterfly
        scratch_.resize(n);
    }

    int size() const { return n_; }

    // in-place 1-D transform of x[0], x[stride], ..., sign -1 forward / +1 backward
    void run(cd* x, int stride, int sign) {
        if (n_ <= 1) return;
        if ((n_ & (n_ - 1)) == 0) {      // iterative radix-2, no allocation
            cd* a = scratch_.data();
            for (int k = 0; k < n_; ++k) a[k] = x[(long)k * stride];
            for (int i = 1, j = 0; i < n_; ++i) {
                int bit = n_ >> 1;
                for (; j & bit; bit >>= 1) j ^= bit;
                j ^= bit;
                if (i < j) std::swap(a[i], a[j]);
            }
            for (int len = 2; len <= n_; len <<= 1) {
                const int half = len >> 1, step = n_ / len;
                for (int i = 0; i < n_; i += len)
                    for (int j = 0; j < half; ++j) {
                        cd w = tw_[j * step];
                        if (sign > 0) w = std::conj(w);
                        const cd u = a[i + j], v = a[i + j + half] * w;
                        a[i + j] = u + v;
                        a[i + j + half] = u - v;
                    }
            }
            for (int k = 0; k < n_; ++k) x[(long)k * stride] = a[k];
            return;
        }
        for (int k = 0; k < n_; ++k) scratch_[k] = x[(long)k * stride];
        std::vector<cd> out(n_);
        rec(scratch_.data(), 1, out.data(), n_, 0, sign);
        for (int k = 0; k < n_; ++k) x[(long)k * stride] = out[k];
    }

private:
    // decimation-in-time mixed radix: out[0..n) = DFT of in[0], in[s], ..., in[(n-1)s]
    void rec(const cd* in, int s, cd* out, int n, int fi, int sign) {
        if (n == 1) { out[0] = in[0]; return; }
        int p = factors_[fi];
        int m = n / p;
        for (int q = 0; q < p; ++q) rec(in + (long)q * s, s * p, out + (long)q * m, m, fi + 1, sign);
        std::vector<cd> t(p);
        const int step = n_ / n;   // twiddle stride in the length-n_ table
        for (int k = 0; k < m; ++k) {
            for (int q = 0; q < p; ++q) {
                cd w = tw_[((long)q * k * step) % n_];
                if (sign > 0) w = std::conj(w);
                t[q] = out[(long)q * m + k] * w;
            }
            for (int r = 0; r < p; ++r) {
                cd acc = 0;
                for (int q = 0; q < p; ++q) {
                    cd w = tw_[((long)q * r * m * step) % n_];
                    if (sign > 0) w = std::conj(w);
                    acc += t[q] * w;
                }
                out[(long)r * m + k] = acc;
            }
        }
    }

    int n_ = 0;
    std::vector<cd> tw_;
    std::vector<int> factors_;
    std::vector<cd> scratch_;
};

// 2-D in-place, unnormalised
inline void fft2d(cd* a, int ny, int nx, int sign) {
    HostFFT fx(nx), fy(ny);
    for (int y = 0; y < ny; ++y) fx.run(a + (long)y * nx, 1, sign);
    for (int x = 0; x < nx; ++x) fy.run(a + x, nx, sign);
}

// fftshift for even sides / the reference's reorder (quadrant swap) for any side
template <typename T>
inline void reorder_into(const T* in, int iny, int inx, T* out) {
    const int mY = iny / 2, mX = inx / 2, oY = iny % 2, oX = inx % 2;
    for (int y = 0; y < iny; ++y)
        for (int x = 0; x < inx; ++x) {
            // input (y,x) lands at ((y - (mY+oY)) mod ny, (x - (mX+oX)) mod nx)
            int yy = (y + iny - (mY + oY)) % iny, xx = (x + inx - (mX + oX)) % inx;
            out[(long)yy * inx + xx] = in[(long)y * inx + x];
        }
}

}  // namespace mfbd
