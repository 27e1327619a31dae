// codelets.h -- compile-time FFT codelets held entirely in registers.
//
// RegFft<L>::run(v, inv) transforms v[0..L) in place (forward: exponent -,
// inverse: exponent +, no scaling) for any L <= 64 whose prime factors above 7
// occur once (radices 8, 4, 2, 7, 5, 3, then an odd prime up to 61).  The codelet is
// the register form of the Stockham autosort stages used everywhere in this
// library (see spfft_device.h): the radix list of L is fixed at compile time,
// every index is a compile-time constant, and every twiddle W_N^i is a
// compile-time constant computed by the constexpr trigonometry below (a
// quadrant reduction on the exact fraction i/N followed by a Taylor series in
// double precision, rounded once to float).  No memory is read for twiddles.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>

#include "spfft_device.h"

namespace spfft {
namespace cx {

constexpr double kPi = 3.14159265358979323846264338327950288;

// Taylor series, |t| <= pi/2: 16 terms leave an error far below 1e-17.
constexpr double sin_series(double t) {
    double t2 = t * t, term = t, sum = t;
    for (int n = 1; n < 16; n++) {
        term *= -t2 / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double cos_series(double t) {
    double t2 = t * t, term = 1.0, sum = 1.0;
    for (int n = 1; n < 16; n++) {
        term *= -t2 / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}
// cos / sin of 2*pi*i/n with i reduced exactly: quadrant q = floor(4 i / n).
constexpr double cos2pi(long i, long n) {
    i %= n;
    if (i < 0) i += n;
    const long e = 4 * i, q = e / n, rem = e - q * n;
    const double t = (kPi / 2.0) * (double)rem / (double)n;
    const double c = cos_series(t), s = sin_series(t);
    return q == 0 ? c : q == 1 ? -s : q == 2 ? -c : s;
}
constexpr double sin2pi(long i, long n) {
    i %= n;
    if (i < 0) i += n;
    const long e = 4 * i, q = e / n, rem = e - q * n;
    const double t = (kPi / 2.0) * (double)rem / (double)n;
    const double c = cos_series(t), s = sin_series(t);
    return q == 0 ? s : q == 1 ? c : q == 2 ? -s : -c;
}

// Radix list of L <= 64 (powers of two first, then 7, 5, 3 and odd primes up to 61).
struct RadixList {
    int n;
    int r[8];
};
constexpr RadixList radix_list(int L) {
    RadixList out{0, {0, 0, 0, 0, 0, 0, 0, 0}};
    int c2 = 0;
    while (L % 2 == 0) { L /= 2; c2++; }
    // powers of two: 8s, then a 4 or 2; 16 = 4*4 and 32 = 8*4 keep stages balanced
    if (c2 == 4) { out.r[out.n++] = 4; out.r[out.n++] = 4; c2 = 0; }
    while (c2 >= 3) { out.r[out.n++] = 8; c2 -= 3; }
    if (c2 == 2) out.r[out.n++] = 4;
    if (c2 == 1) out.r[out.n++] = 2;
    while (L % 7 == 0) { out.r[out.n++] = 7; L /= 7; }
    while (L % 5 == 0) { out.r[out.n++] = 5; L /= 5; }
    while (L % 3 == 0) { out.r[out.n++] = 3; L /= 3; }
    // larger odd primes (the Rader lengths): one stage each, a direct butterfly
    for (int p = 11; p <= 61 && L > 1; p += 2) {
        bool prime = true;
        for (int d = 3; d * d <= p; d += 2) prime = prime && p % d != 0;
        while (prime && L % p == 0) { out.r[out.n++] = p; L /= p; }
    }
    if (L != 1) out.n = -1;
    return out;
}

template <class F, int... Is>
SPFFT_HD void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
SPFFT_HD void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

}  // namespace cx

// v * W_N^{+-i} with a compile-time twiddle (forward: exp(-2 pi i I/N)).
template <int N, int I>
SPFFT_HD float2 twc(float2 v, bool inv) {
    constexpr float c = (float)cx::cos2pi(I, N);
    constexpr float s = (float)cx::sin2pi(I, N);
    // forward multiplier (c - i s); inverse (c + i s)
    return inv ? make_float2(fmaf(v.x, c, -v.y * s), fmaf(v.x, s, v.y * c))
               : make_float2(fmaf(v.x, c, v.y * s), fmaf(v.y, c, -v.x * s));
}

// Butterfly of an odd prime radix R > 7 (dft_odd) with compile-time cos / sin
// tables: (R - 1)^2 real FMAs, no twiddle memory.
template <int R>
struct OddRoots {
    float c[R], s[R];
    constexpr OddRoots() : c{}, s{} {
        for (int k = 0; k < R; k++) {
            c[k] = (float)cx::cos2pi(k, R);
            s[k] = (float)cx::sin2pi(k, R);
        }
    }
};
template <int R>
struct Dft {
    static_assert(R > 7 && R % 2 == 1, "radix 2, 3, 4, 5, 7 and 8 have their own butterflies");
    static SPFFT_HD void run(float2* v, bool inv) {
        constexpr OddRoots<R> t{};
        dft_odd<R, (R - 1) / 2>(v, inv, t.c, t.s);
    }
};

template <int L, int S, int Ns>
struct RegStages {
    static SPFFT_HD void run(float2* a, float2* b, bool inv) {
        constexpr cx::RadixList pl = cx::radix_list(L);
        if constexpr (S < pl.n) {
            constexpr int R = pl.r[S];
            constexpr int nb = L / R;
            cx::static_for<nb>([&](auto Jc) {
                constexpr int j = decltype(Jc)::value;
                constexpr int k = j % Ns;
                float2 v[R];
                cx::static_for<R>([&](auto Rc) {
                    constexpr int r = decltype(Rc)::value;
                    if constexpr (r == 0 || k == 0) v[r] = a[j + r * nb];
                    else v[r] = twc<Ns * R, (r * k) % (Ns * R)>(a[j + r * nb], inv);
                });
                Dft<R>::run(v, inv);
                cx::static_for<R>([&](auto Rc) {
                    constexpr int r = decltype(Rc)::value;
                    b[(j / Ns) * Ns * R + k + r * Ns] = v[r];
                });
            });
            RegStages<L, S + 1, Ns * R>::run(b, a, inv);
        } else if constexpr (S % 2 == 1) {
            // odd number of stages: the result sits in `a`, which is the caller's scratch
#pragma unroll
            for (int i = 0; i < L; i++) b[i] = a[i];
        }
    }
};

template <int L>
struct RegFft {
    static_assert(cx::radix_list(L).n >= 0, "RegFft length has no radix list");
    // In-place on v[0..L).
    static SPFFT_HD void run(float2* v, bool inv) {
        if constexpr (L > 1) {
            float2 tmp[L];
            RegStages<L, 0, 1>::run(v, tmp, inv);
        }
    }
};

}  // namespace spfft
