// cpu_fft.h -- single-threaded FP32 dense 2-D real-to-complex FFT on the host: the CLI's
// --cpu-reference row.  The reference times FFTW3's fftwf_plan_dft_r2c_2d there
// (src/cpu_reference.cpp:8-64, driver src/main.cu:324-342); FFTW3 is not part of this
// build, so this is a transform of its own, labelled as such wherever it is printed.
#pragma once

#include <complex>
#include <memory>
#include <vector>

namespace spfft_cpu {

using cf = std::complex<float>;

// Forward DFT of one length n (exponent -), in place.  Mixed radix (decimation in time)
// over the factors 4, 2, 3, 5, 7 and the odd primes up to 31; a larger prime factor makes
// it Bluestein's algorithm over the smallest 5-smooth length m >= 2n - 1.  Twiddles and
// chirps are computed in FP64 with exact integer angle reduction and rounded once to FP32;
// the arithmetic is FP32.
class Dft {
public:
    explicit Dft(int n);
    int size() const { return n_; }
    // x[0..n) in place; scratch: scratch_size() entries.
    void run(cf* x, cf* scratch) const;
    size_t scratch_size() const;

private:
    void rec(const cf* in, size_t is, cf* out, int n, int level) const;
    int n_;
    std::vector<int> fac_;  // radices, outermost first (mixed radix)
    std::vector<cf> tw_;    // W_n^e, e < n (mixed radix)
    // Bluestein: chirp c[k] = e^{-i pi k^2 / n}, B^ = DFT_m(conj chirp, wrapped) / m
    int m_ = 0;
    std::vector<cf> chirp_, bhat_;
    std::unique_ptr<Dft> inner_;
};

// out[u * (C/2 + 1) + v] = sum over r < R, c < C of in[r * C + c] e^{-2 pi i (u r / R + v c / C)}
// for u < R, v <= C/2: the row-major R x (C/2 + 1) half spectrum of fftwf_plan_dft_r2c_2d.
void rfft2(int R, int C, const float* in, cf* out);

}  // namespace spfft_cpu
