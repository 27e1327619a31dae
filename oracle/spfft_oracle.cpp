// spfft_oracle.cpp -- CPU oracle for the binary-sparse 2-D R2C FFT.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (cuspfft_amd/, include/)
// links, loads or calls this file.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg use it, and only as the checker.
//
// What it restates (reference = Darkviper7/cuSpFFT, paths relative to its root):
//   orc_next_3_smooth / orc_next_7_smooth / orc_factor_3_smooth
//       src/sparse_fft_bluestein.cu:1975-2011
//   orc_compact_csc          include/sparse_fft_internal.cuh:102-120, 200-222
//                            (sorted unique columns, prefix-summed col_ptr,
//                             stable counting sort of the input COO order)
//   orc_can_pack_u16         include/sparse_fft_internal.cuh:122-124
//   orc_bluestein_chirp      include/sparse_fft_internal.cuh:168-177
//   orc_bluestein_b          include/sparse_fft_internal.cuh:179-191
//   orc_sparse_rfft2_f64     the quantity the reference's variants compute,
//                            X[u][v] = sum_{(r,c)} exp(-2 pi i (u r/R + v c/C)),
//                            v <= C/2 (FFTW r2c_2d layout, src/cpu_reference.cpp:8-64),
//                            evaluated in FP64: column pass with the exact
//                            integer phase (u*r) mod R (sparse_fft_internal.cuh:51-53,
//                            sparse_fft_bluestein.cu:366-397), then an FP64 length-C
//                            DFT per output row.
//   orc_sparse_rfft2_f32     FP32 restatement of the plain algorithm of
//                            sparse_fft_csc_bluestein_mixed_radix
//                            (src/sparse_fft_bluestein.cu:2256-2392): compact CSC,
//                            pass 1 summed in FP32, chirp pre-multiply, FP32
//                            Stockham FFT at M = next_3_smooth(2C-1) with the
//                            factor order of factor_3_smooth, pointwise product
//                            with B^ = FFT_M(b) computed by the same FP32 FFT,
//                            FP32 inverse FFT, 1/M scale, chirp post-multiply
//                            (bluestein_finalize_kernel, :469-486).  Twiddles come
//                            from FP64 tables rounded to FP32 (no fast intrinsics).
//
// How each function is pinned (tests/test_oracle.py): hand-built CSC fixtures and
// a brute-force stable sort; closed-form smooth lengths; numpy.fft (pocketfft FP64)
// on dense inputs; direct O(n^2) DFT on tiny inputs; invariants X[0][0] = nnz,
// Parseval, symmetry of symmetric inputs, X[u][0] = DFT_R(row degrees).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef std::complex<double> cd;
typedef std::complex<float> cf;
const double kPi = 3.14159265358979323846264338327950288;

// exp(sign * 2 pi i * k / n) with exact integer reduction of k.
inline cd unit_root(int64_t k, int64_t n, int sign) {
    int64_t kk = k % n;
    if (kk < 0) kk += n;
    const double a = 2.0 * kPi * (double)kk / (double)n;
    return cd(std::cos(a), sign * std::sin(a));
}

// ---------------------------------------------------------------------------
// FP64 FFT: iterative radix-2 for powers of two, Bluestein for everything else.
// ---------------------------------------------------------------------------
void fft_pow2(cd* a, int n, int sign) {
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; k++) {
                const cd w = unit_root((int64_t)k, len, sign);
                const cd u = a[i + k];
                const cd v = a[i + k + half] * w;
                a[i + k] = u + v;
                a[i + k + half] = u - v;
            }
        }
    }
}

struct Fft64Plan {
    int n = 0, L = 0;
    std::vector<cd> chirp;   // exp(sign * -i pi k^2 / n), k < n
    std::vector<cd> bhat;    // FFT_L of the wrapped conjugate chirp
    int sign = -1;
    void init(int n_, int sign_) {
        n = n_;
        sign = sign_;
        if (n <= 1 || (n & (n - 1)) == 0) return;
        L = 1;
        while (L < 2 * n - 1) L <<= 1;
        chirp.resize(n);
        for (int k = 0; k < n; k++) {
            // exponent pi k^2 / n = 2 pi (k^2 mod 2n) / (2n), exact in integers
            const int64_t k2 = ((int64_t)k * k) % (2LL * n);
            chirp[k] = unit_root(k2, 2LL * n, sign);
        }
        bhat.assign(L, cd(0, 0));
        bhat[0] = std::conj(chirp[0]);
        for (int k = 1; k < n; k++) {
            bhat[k] = std::conj(chirp[k]);
            bhat[L - k] = std::conj(chirp[k]);
        }
        fft_pow2(bhat.data(), L, -1);
    }
    // in-place length-n DFT with exponent sign `sign`
    void run(cd* x, std::vector<cd>& work) const {
        if (n <= 1) return;
        if (L == 0) { fft_pow2(x, n, sign); return; }
        work.assign(L, cd(0, 0));
        for (int k = 0; k < n; k++) work[k] = x[k] * chirp[k];
        fft_pow2(work.data(), L, -1);
        for (int k = 0; k < L; k++) work[k] *= bhat[k];
        fft_pow2(work.data(), L, +1);
        const double s = 1.0 / L;
        for (int k = 0; k < n; k++) x[k] = work[k] * s * chirp[k];
    }
};

// ---------------------------------------------------------------------------
// FP32 Stockham FFT with explicit radix list (restating the reference's
// run_fft_mixed_23 stage sequence, sparse_fft_bluestein.cu:2215-2254).
// Each stage: gather R inputs at stride n/R, twiddle, R-point DFT, scatter.
// ---------------------------------------------------------------------------
struct Fft32Plan {
    int n = 0;
    std::vector<int> radices;
    std::vector<cf> tw_fwd, tw_inv;              // W_n^i, i < n (FP64 rounded)
    std::vector<std::vector<cf>> dft_fwd, dft_inv;  // per-radix R x R matrices
    void init(int n_, const std::vector<int>& rad) {
        n = n_;
        radices = rad;
        tw_fwd.resize(n);
        tw_inv.resize(n);
        for (int i = 0; i < n; i++) {
            cd w = unit_root(i, n, -1);
            tw_fwd[i] = cf((float)w.real(), (float)w.imag());
            tw_inv[i] = std::conj(tw_fwd[i]);
        }
        dft_fwd.clear();
        dft_inv.clear();
        for (int R : radices) {
            std::vector<cf> f(R * R), g(R * R);
            for (int a = 0; a < R; a++)
                for (int b = 0; b < R; b++) {
                    cd w = unit_root((int64_t)a * b, R, -1);
                    f[a * R + b] = cf((float)w.real(), (float)w.imag());
                    g[a * R + b] = std::conj(f[a * R + b]);
                }
            dft_fwd.push_back(f);
            dft_inv.push_back(g);
        }
    }
    // returns pointer to the buffer holding the result (a or b)
    cf* run(cf* a, cf* b, bool inverse) const {
        const std::vector<cf>& tw = inverse ? tw_inv : tw_fwd;
        int Ns = 1;
        cf v[16], y[16];
        for (size_t s = 0; s < radices.size(); s++) {
            const int R = radices[s];
            const std::vector<cf>& F = inverse ? dft_inv[s] : dft_fwd[s];
            const int nR = n / R;
            const int step = n / (Ns * R);
            for (int j = 0; j < nR; j++) {
                const int k = j % Ns;
                for (int r = 0; r < R; r++) {
                    v[r] = a[j + r * nR];
                    if (r && k) v[r] *= tw[(int64_t)r * k * step % n];
                }
                for (int p = 0; p < R; p++) {
                    cf acc(0.f, 0.f);
                    for (int q = 0; q < R; q++) acc += F[p * R + q] * v[q];
                    y[p] = acc;
                }
                const int d = (j / Ns) * Ns * R + k;
                for (int r = 0; r < R; r++) b[d + r * Ns] = y[r];
            }
            std::swap(a, b);
            Ns *= R;
        }
        return a;
    }
};

int next_smooth(int n, const int* primes, int np) {
    if (n < 1) n = 1;
    for (int m = n;; m++) {
        int x = m;
        for (int i = 0; i < np; i++)
            while (x % primes[i] == 0) x /= primes[i];
        if (x == 1) return m;
    }
}

}  // namespace

extern "C" {

// src/sparse_fft_bluestein.cu:1975-1983
int orc_next_3_smooth(int n) {
    static const int p[] = {2, 3};
    return next_smooth(n, p, 2);
}

// src/sparse_fft_bluestein.cu:1988-1995
int orc_next_7_smooth(int n) {
    static const int p[] = {2, 3, 5, 7};
    return next_smooth(n, p, 4);
}

// src/sparse_fft_bluestein.cu:2000-2011: pairs of 3 as radix 9 first, a
// leftover 3, then all 2s.  Returns the number of factors or -1 if n is not
// 3-smooth (the reference throws).
int orc_factor_3_smooth(int n, int* out, int cap) {
    int cnt3 = 0, k = 0;
    while (n % 3 == 0) { cnt3++; n /= 3; }
    for (int i = 0; i < cnt3 / 2; i++) { if (k < cap) out[k] = 9; k++; }
    if (cnt3 % 2) { if (k < cap) out[k] = 3; k++; }
    while (n % 2 == 0) { if (k < cap) out[k] = 2; k++; n /= 2; }
    if (n != 1) return -1;
    return k;
}

// include/sparse_fft_internal.cuh:122-124
int orc_can_pack_u16(int rows, int cols) { return rows <= 65536 && cols <= 65536; }

// include/sparse_fft_internal.cuh:102-120 and 200-222.  sparse_cols must hold
// `cols` ints, col_ptr `cols + 1`, row_idx `nnz`.  Returns n_sparse_cols.
int orc_compact_csc(int rows, int cols, int nnz, const int* row, const int* col,
                    int* sparse_cols, int* col_ptr, int* row_idx) {
    (void)rows;
    std::vector<int> sc(col, col + nnz);
    std::sort(sc.begin(), sc.end());
    sc.erase(std::unique(sc.begin(), sc.end()), sc.end());
    const int n_sp = (int)sc.size();
    std::vector<int> slot(cols > 0 ? cols : 1, -1);
    for (int i = 0; i < n_sp; i++) { slot[sc[i]] = i; sparse_cols[i] = sc[i]; }
    for (int i = 0; i <= n_sp; i++) col_ptr[i] = 0;
    for (int i = 0; i < nnz; i++) col_ptr[slot[col[i]] + 1]++;
    for (int i = 0; i < n_sp; i++) col_ptr[i + 1] += col_ptr[i];
    std::vector<int> next(col_ptr, col_ptr + n_sp + 1);
    for (int i = 0; i < nnz; i++) row_idx[next[slot[col[i]]]++] = row[i];
    return n_sp;
}

// include/sparse_fft_internal.cuh:168-177 (FP64 angle -pi k^2 / n, rounded)
void orc_bluestein_chirp(int n, float* out) {
    for (int k = 0; k < n; k++) {
        const double kk = (double)k * (double)k;
        const double angle = -kPi * kk / (double)n;
        out[2 * k] = (float)std::cos(angle);
        out[2 * k + 1] = (float)std::sin(angle);
    }
}

// include/sparse_fft_internal.cuh:179-191
void orc_bluestein_b(int n, int M, float* out) {
    for (int i = 0; i < 2 * M; i++) out[i] = 0.f;
    out[0] = 1.f;
    for (int k = 1; k < n; k++) {
        const double kk = (double)k * (double)k;
        const double angle = kPi * kk / (double)n;
        const float re = (float)std::cos(angle), im = (float)std::sin(angle);
        out[2 * k] = re;
        out[2 * k + 1] = im;
        out[2 * (M - k)] = re;
        out[2 * (M - k) + 1] = im;
    }
}

// In-place FP64 DFT of length n (interleaved re/im), exponent sign `sign`.
void orc_fft64(int n, double* data, int sign) {
    Fft64Plan p;
    p.init(n, sign < 0 ? -1 : 1);
    std::vector<cd> x(n), work;
    for (int i = 0; i < n; i++) x[i] = cd(data[2 * i], data[2 * i + 1]);
    p.run(x.data(), work);
    for (int i = 0; i < n; i++) { data[2 * i] = x[i].real(); data[2 * i + 1] = x[i].imag(); }
}

// FP64 2-D half spectrum for the output rows listed in u_list.
// out: n_u x (cols/2+1) complex double, interleaved.
int orc_sparse_rfft2_f64(int rows, int cols, int nnz, const int* row, const int* col,
                         const int* u_list, int n_u, double* out, int nthreads) {
    if (rows <= 0 || cols <= 0) return -1;
    std::vector<int> sc(cols), cp(cols + 1), ri(nnz > 0 ? nnz : 1);
    const int n_sp = orc_compact_csc(rows, cols, nnz, row, col, sc.data(), cp.data(), ri.data());
    std::vector<cd> wr(rows);
    for (int k = 0; k < rows; k++) wr[k] = unit_root(k, rows, -1);
    Fft64Plan plan;
    plan.init(cols, -1);
    const int V = cols / 2 + 1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        std::vector<cd> x(cols), work;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (int t = 0; t < n_u; t++) {
            const int u = u_list[t];
            std::fill(x.begin(), x.end(), cd(0, 0));
            for (int s = 0; s < n_sp; s++) {
                cd acc(0, 0);
                for (int p = cp[s]; p < cp[s + 1]; p++)
                    acc += wr[(int64_t)u * ri[p] % rows];
                x[sc[s]] = acc;
            }
            plan.run(x.data(), work);
            double* o = out + (size_t)t * V * 2;
            for (int v = 0; v < V; v++) { o[2 * v] = x[v].real(); o[2 * v + 1] = x[v].imag(); }
        }
    }
    return 0;
}

// FP32 restatement of the reference's plain algorithm (mixed-radix variant when
// smooth == 3, 7-smooth length when smooth == 7) for output rows u_list.
// out: n_u x (cols/2+1) complex float, interleaved.
int orc_sparse_rfft2_f32(int rows, int cols, int nnz, const int* row, const int* col,
                         int smooth, const int* u_list, int n_u, float* out, int nthreads) {
    if (rows <= 0 || cols <= 0) return -1;
    const int M = smooth == 7 ? orc_next_7_smooth(2 * cols - 1) : orc_next_3_smooth(2 * cols - 1);
    // radix list: factor_3_smooth order for 3-smooth M; for 7-smooth M the
    // 7s, 5s, then the 3-smooth remainder in factor_3_smooth order.
    std::vector<int> rad;
    {
        int m = M;
        while (m % 7 == 0) { rad.push_back(7); m /= 7; }
        while (m % 5 == 0) { rad.push_back(5); m /= 5; }
        int buf[64];
        const int k = orc_factor_3_smooth(m, buf, 64);
        if (k < 0) return -2;
        rad.insert(rad.end(), buf, buf + k);
    }
    std::vector<int> sc(cols), cp(cols + 1), ri(nnz > 0 ? nnz : 1);
    const int n_sp = orc_compact_csc(rows, cols, nnz, row, col, sc.data(), cp.data(), ri.data());
    std::vector<cf> wr(rows);
    for (int k = 0; k < rows; k++) {
        cd w = unit_root(k, rows, -1);
        wr[k] = cf((float)w.real(), (float)w.imag());
    }
    std::vector<float> hc(2 * cols), hb(2 * M);
    orc_bluestein_chirp(cols, hc.data());
    orc_bluestein_b(cols, M, hb.data());
    Fft32Plan plan;
    plan.init(M, rad);
    std::vector<cf> bhat(M), tmp(M);
    for (int i = 0; i < M; i++) bhat[i] = cf(hb[2 * i], hb[2 * i + 1]);
    {
        cf* r = plan.run(bhat.data(), tmp.data(), false);
        if (r != bhat.data()) std::copy(r, r + M, bhat.begin());
    }
    const int V = cols / 2 + 1;
    const float inv_m = 1.f / (float)M;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        std::vector<cf> a(M), b(M);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (int t = 0; t < n_u; t++) {
            const int u = u_list[t];
            std::fill(a.begin(), a.end(), cf(0.f, 0.f));
            for (int s = 0; s < n_sp; s++) {
                float re = 0.f, im = 0.f;
                for (int p = cp[s]; p < cp[s + 1]; p++) {
                    const cf w = wr[(uint32_t)u * (uint32_t)ri[p] % (uint32_t)rows];
                    re += w.real();
                    im += w.imag();
                }
                const int c = sc[s];
                a[c] = cf(re, im) * cf(hc[2 * c], hc[2 * c + 1]);
            }
            cf* f = plan.run(a.data(), b.data(), false);
            cf* g = (f == a.data()) ? b.data() : a.data();
            for (int i = 0; i < M; i++) f[i] *= bhat[i];
            cf* y = plan.run(f, g, true);
            float* o = out + (size_t)t * V * 2;
            for (int v = 0; v < V; v++) {
                const cf z = y[v] * inv_m * cf(hc[2 * v], hc[2 * v + 1]);
                o[2 * v] = z.real();
                o[2 * v + 1] = z.imag();
            }
        }
    }
    return 0;
}

}  // extern "C"
