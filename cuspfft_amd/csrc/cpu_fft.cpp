// cpu_fft.cpp -- the host FP32 2-D real-to-complex FFT of the CLI's --cpu-reference row
// (declarations and scope in cpu_fft.h).
#include "cpu_fft.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace spfft_cpu {

namespace {

// e^{-2 pi i e / n} in FP64 (e reduced mod n first), rounded once
cf unit_root(long long e, long long n) {
    e %= n;
    if (e < 0) e += n;
    const double a = -2.0 * M_PI * (double)e / (double)n;
    return cf((float)std::cos(a), (float)std::sin(a));
}

bool smooth5(int m) {
    for (int d : {2, 3, 5})
        while (m % d == 0) m /= d;
    return m == 1;
}

}  // namespace

Dft::Dft(int n) : n_(n) {
    if (n < 1) throw std::invalid_argument("DFT length must be positive");
    int r = n;
    while (r % 4 == 0) {
        fac_.push_back(4);
        r /= 4;
    }
    for (int p : {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31})
        while (r % p == 0) {
            fac_.push_back(p);
            r /= p;
        }
    if (r == 1) {
        tw_.resize(n);
        for (int e = 0; e < n; e++) tw_[e] = unit_root(e, n);
        return;
    }
    // Bluestein: X[k] = c[k] sum_j (x[j] c[j]) conj(c[k - j]), a cyclic convolution of
    // length m >= 2n - 1 with b[j] = conj(c[j]) at j and m - j
    fac_.clear();
    m_ = 2 * n - 1;
    while (!smooth5(m_)) m_++;
    inner_ = std::make_unique<Dft>(m_);
    chirp_.resize(n);
    for (long long k = 0; k < n; k++) {
        // e^{-i pi k^2 / n} = W_{2n}^{k^2 mod 2n}
        chirp_[k] = unit_root(k * k % (2LL * n), 2LL * n);
    }
    std::vector<cf> b(m_, cf(0.f, 0.f)), scr(inner_->scratch_size());
    b[0] = std::conj(chirp_[0]);
    for (int k = 1; k < n; k++) b[k] = b[m_ - k] = std::conj(chirp_[k]);
    inner_->run(b.data(), scr.data());
    const float inv = 1.0f / (float)m_;
    for (cf& v : b) v *= inv;
    bhat_ = std::move(b);
}

size_t Dft::scratch_size() const { return m_ ? 2 * (size_t)m_ + inner_->scratch_size() : (size_t)n_; }

// out[0..n) = DFT_n of in[0], in[is], in[2 is], ...: the sub-DFTs of the r decimated
// sequences into the r blocks of out, then the radix-r butterflies with twiddles
// W_n^{q k} across the blocks (in place: the butterfly over k reads and writes the
// entries k + m q, q < r).
void Dft::rec(const cf* in, size_t is, cf* out, int n, int level) const {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    const int r = fac_[level], m = n / r;
    for (int q = 0; q < r; q++) rec(in + q * is, is * r, out + (size_t)q * m, m, level + 1);
    const int step = n_ / n;  // W_n^e = W_N^{e N / n}
    cf t[31], y[31];
    for (int k = 0; k < m; k++) {
        t[0] = out[k];
        for (int q = 1; q < r; q++) t[q] = out[(size_t)q * m + k] * tw_[(size_t)q * k * step];
        switch (r) {
            case 2:
                y[0] = t[0] + t[1];
                y[1] = t[0] - t[1];
                break;
            case 4: {
                const cf a = t[0] + t[2], b = t[0] - t[2], c = t[1] + t[3];
                const cf d = t[1] - t[3];
                const cf dn(d.imag(), -d.real());  // -i d
                y[0] = a + c;
                y[2] = a - c;
                y[1] = b + dn;
                y[3] = b - dn;
                break;
            }
            case 3: {
                const float h = 0.8660254037844386f;  // sqrt(3) / 2
                const cf s = t[1] + t[2], d = t[1] - t[2];
                const cf base = t[0] - 0.5f * s;
                const cf rot(h * d.imag(), -h * d.real());  // -i (sqrt(3)/2) d
                y[0] = t[0] + s;
                y[1] = base + rot;
                y[2] = base - rot;
                break;
            }
            default: {
                // direct: y_s = sum_q t_q W_r^{q s}
                const int rs = n_ / r;
                for (int s = 0; s < r; s++) {
                    cf acc = t[0];
                    for (int q = 1; q < r; q++) acc += t[q] * tw_[(size_t)((q * s) % r) * rs];
                    y[s] = acc;
                }
            }
        }
        for (int s = 0; s < r; s++) out[k + (size_t)m * s] = y[s];
    }
}

void Dft::run(cf* x, cf* scratch) const {
    if (!m_) {
        rec(x, 1, scratch, n_, 0);
        std::copy(scratch, scratch + n_, x);
        return;
    }
    cf* a = scratch;
    cf* inner_scr = scratch + 2 * (size_t)m_;
    for (int k = 0; k < n_; k++) a[k] = x[k] * chirp_[k];
    std::fill(a + n_, a + m_, cf(0.f, 0.f));
    inner_->run(a, inner_scr);
    // inverse by conjugation: IDFT(y) = conj(DFT(conj(y))) / m (the 1/m is in B^)
    for (int k = 0; k < m_; k++) a[k] = std::conj(a[k] * bhat_[k]);
    inner_->run(a, inner_scr);
    for (int k = 0; k < n_; k++) x[k] = std::conj(a[k]) * chirp_[k];
}

void rfft2(int R, int C, const float* in, cf* out) {
    if (R < 1 || C < 1) throw std::invalid_argument("rfft2: empty matrix");
    const size_t V = (size_t)C / 2 + 1;
    Dft row(C), col(R);
    std::vector<cf> z(C), scr(std::max(row.scratch_size(), col.scratch_size()));
    // rows, two at a time: z = x_a + i x_b, X_a[k] = (Z[k] + conj Z[-k]) / 2,
    // X_b[k] = (Z[k] - conj Z[-k]) / (2i)
    for (int j = 0; j < R; j += 2) {
        const bool pair = j + 1 < R;
        const float* xa = in + (size_t)j * C;
        const float* xb = pair ? xa + C : nullptr;
        for (int c = 0; c < C; c++) z[c] = cf(xa[c], pair ? xb[c] : 0.f);
        row.run(z.data(), scr.data());
        cf* oa = out + (size_t)j * V;
        cf* ob = pair ? oa + V : nullptr;
        for (size_t k = 0; k < V; k++) {
            const cf zk = z[k], zc = std::conj(z[(C - k) % C]);
            oa[k] = 0.5f * (zk + zc);
            if (pair) {
                const cf d = zk - zc;
                ob[k] = cf(0.5f * d.imag(), -0.5f * d.real());  // d / (2i)
            }
        }
    }
    // columns, a block of them at a time (gathered into contiguous buffers)
    constexpr size_t kBlock = 16;
    std::vector<cf> cols(kBlock * (size_t)R);
    for (size_t k0 = 0; k0 < V; k0 += kBlock) {
        const size_t nb = std::min(kBlock, V - k0);
        for (int u = 0; u < R; u++)
            for (size_t b = 0; b < nb; b++) cols[b * R + u] = out[(size_t)u * V + k0 + b];
        for (size_t b = 0; b < nb; b++) col.run(cols.data() + b * R, scr.data());
        for (int u = 0; u < R; u++)
            for (size_t b = 0; b < nb; b++) out[(size_t)u * V + k0 + b] = cols[b * R + u];
    }
}

}  // namespace spfft_cpu
