// Host-side check of every register codelet RegFft<L> (the same template the
// gfx950 kernels instantiate) against a direct FP64 DFT.  Built with hipcc and
// run on the CPU; no device code is launched.
#include <cmath>
#include <complex>
#include <cstdio>
#include <random>
#include <vector>

#include "codelets.h"

template <int L>
static double check_one(std::mt19937& g) {
    std::uniform_real_distribution<float> d(-1.f, 1.f);
    double worst = 0;
    for (int inv = 0; inv < 2; inv++) {
        float2 v[L];
        std::vector<std::complex<double>> x(L);
        for (int i = 0; i < L; i++) {
            v[i] = make_float2(d(g), d(g));
            x[i] = {v[i].x, v[i].y};
        }
        spfft::RegFft<L>::run(v, inv != 0);
        double num = 0, den = 0;
        for (int p = 0; p < L; p++) {
            std::complex<double> acc = 0;
            for (int q = 0; q < L; q++) {
                const double a = 2 * M_PI * (double)((long)p * q % L) / L;
                acc += x[q] * std::complex<double>(std::cos(a), inv ? std::sin(a) : -std::sin(a));
            }
            num += std::norm(acc - std::complex<double>(v[p].x, v[p].y));
            den += std::norm(acc);
        }
        worst = std::max(worst, std::sqrt(num / den));
    }
    return worst;
}

template <int... Ls>
static int run_all(std::mt19937& g) {
    int bad = 0;
    ((void)[&] {
        const double e = check_one<Ls>(g);
        std::printf("L=%d rel_l2=%.3e\n", Ls, e);
        if (!(e < 1e-6)) bad++;
    }(), ...);
    return bad;
}

int main() {
    std::mt19937 g(1);
    const int bad = run_all<2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 15, 16, 18, 20, 21, 24, 25, 27, 28,
                            30, 32, 35, 36, 40, 42, 45, 48, 49, 50, 54, 56, 60, 63, 64>(g);
    std::printf("%s\n", bad ? "FAIL" : "OK");
    return bad ? 1 : 0;
}
