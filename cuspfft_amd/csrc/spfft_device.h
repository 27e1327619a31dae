// spfft_device.h -- device building blocks for the gfx950 kernels:
// complex helpers, exact integer phase reduction and the radix-2/3/4/5/7/8
// DFT butterflies (host+device, so the codelets can be checked on the CPU).
//
// Every FFT in this library uses the Stockham autosort ("expand") stage: for
// stage radix R and Ns = product of the radices already applied,
//   butterfly j in [0, L/R):  k = j mod Ns
//     v[r] = x[j + r*L/R] * W_{Ns*R}^{r*k}          (r < R)
//     v    = DFT_R(v)
//     y[(j/Ns)*Ns*R + k + r*Ns] = v[r]
// which leaves the result in natural order after the last stage.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "spfft_types.h"

#define SPFFT_HD __host__ __device__ __forceinline__

namespace spfft {

SPFFT_HD float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
SPFFT_HD float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
SPFFT_HD float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
// a * conj(b)
SPFFT_HD float2 cmulc(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -a.x * b.y));
}
SPFFT_HD float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
SPFFT_HD float2 mul_mi(float2 a, bool inv) {
    return inv ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// a mod d for 32-bit a (Lemire-Kaser-Kurz "fastmod"): magic = 2^64 / d + 1.
__device__ __forceinline__ uint32_t fastmod_u32(uint32_t a, uint64_t magic, uint32_t d) {
    const uint64_t low = magic * (uint64_t)a;
    return (uint32_t)__umul64hi(low, (uint64_t)d);
}

// q = a / d for a < 2^24 using a float reciprocal with one-step correction.
__device__ __forceinline__ int fdiv_small(int a, int d, float inv_d) {
    int q = (int)((float)a * inv_d);
    int r = a - q * d;
    if (r < 0) { q -= 1; } else if (r >= d) { q += 1; }
    return q;
}

// ---------------------------------------------------------------------------
// Radix butterflies.  Forward: y[p] = sum_q v[q] exp(-2 pi i p q / R).
// ---------------------------------------------------------------------------
template <int R> struct Dft;

template <> struct Dft<2> {
    static SPFFT_HD void run(float2* v, bool) {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <> struct Dft<4> {
    static SPFFT_HD void run(float2* v, bool inv) {
        const float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        const float2 s13 = cadd(v[1], v[3]), d13 = mul_mi(csub(v[1], v[3]), inv);
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    }
};

template <> struct Dft<8> {
    static SPFFT_HD void run(float2* v, bool inv) {
        const float h = 0.70710678118654752f;
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4>::run(e, inv);
        Dft<4>::run(o, inv);
        // W8^1 = (1 - i)/sqrt2 fwd, (1 + i)/sqrt2 inv; W8^2 = -i / +i; W8^3 = (-1 -+ i)/sqrt2
        const float2 o1 = inv ? make_float2(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
                              : make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
        const float2 o2 = mul_mi(o[2], inv);
        const float2 o3 = inv ? make_float2(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
                              : make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

// Odd prime radix via symmetric pairs: with s_q = v_q + v_{R-q},
// d_q = v_q - v_{R-q}, C_k = cos(2 pi k/R), S_k = sin(2 pi k/R):
//   y_0 = v_0 + sum s_q
//   A_p = v_0 + sum_q s_q C_{pq},  B_p = sum_q d_q S_{pq}
//   forward: y_p = A_p - i B_p, y_{R-p} = A_p + i B_p (inverse swaps the signs).
template <int R, int H>
SPFFT_HD void dft_odd(float2* v, bool inv, const float* Cs, const float* Sn) {
    float2 s[H], d[H];
#pragma unroll
    for (int q = 1; q <= H; q++) {
        s[q - 1] = cadd(v[q], v[R - q]);
        d[q - 1] = csub(v[q], v[R - q]);
    }
    float2 y0 = v[0];
#pragma unroll
    for (int q = 0; q < H; q++) y0 = cadd(y0, s[q]);
    float2 out[R];
    out[0] = y0;
#pragma unroll
    for (int p = 1; p <= H; p++) {
        float2 A = v[0], B = make_float2(0.f, 0.f);
#pragma unroll
        for (int q = 1; q <= H; q++) {
            const int k = (p * q) % R;
            A.x = fmaf(s[q - 1].x, Cs[k], A.x);
            A.y = fmaf(s[q - 1].y, Cs[k], A.y);
            B.x = fmaf(d[q - 1].x, Sn[k], B.x);
            B.y = fmaf(d[q - 1].y, Sn[k], B.y);
        }
        const float2 iB = mul_mi(B, inv);  // -iB forward, +iB inverse
        out[p] = cadd(A, iB);
        out[R - p] = csub(A, iB);
    }
#pragma unroll
    for (int p = 0; p < R; p++) v[p] = out[p];
}

template <> struct Dft<3> {
    static SPFFT_HD void run(float2* v, bool inv) {
        const float Cs[3] = {1.f, -0.5f, -0.5f};
        const float Sn[3] = {0.f, 0.86602540378443865f, -0.86602540378443865f};
        dft_odd<3, 1>(v, inv, Cs, Sn);
    }
};

template <> struct Dft<5> {
    static SPFFT_HD void run(float2* v, bool inv) {
        const float Cs[5] = {1.f, 0.30901699437494745f, -0.80901699437494742f,
                             -0.80901699437494742f, 0.30901699437494745f};
        const float Sn[5] = {0.f, 0.95105651629515357f, 0.58778525229247314f,
                             -0.58778525229247314f, -0.95105651629515357f};
        dft_odd<5, 2>(v, inv, Cs, Sn);
    }
};

template <> struct Dft<7> {
    static SPFFT_HD void run(float2* v, bool inv) {
        const float Cs[7] = {1.f, 0.62348980185873353f, -0.22252093395631440f, -0.90096886790241913f,
                             -0.90096886790241913f, -0.22252093395631440f, 0.62348980185873353f};
        const float Sn[7] = {0.f, 0.78183148246802981f, 0.97492791218182361f, 0.43388373911755812f,
                             -0.43388373911755812f, -0.97492791218182361f, -0.78183148246802981f};
        dft_odd<7, 3>(v, inv, Cs, Sn);
    }
};

}  // namespace spfft
