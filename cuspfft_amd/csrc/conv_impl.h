// conv_impl.h -- k_conv: for every (u, k1) row of T, the length-M2 FFT, the
// product with B^[k1 + M1 k2] / M, and the inverse length-M2 FFT, through LDS.
//
// Fast path k_conv_fast<R1, R2>: M2 = R1 * R2 with both radices compile-time
// register codelets; three passes (forward R1; forward R2, the B^ product and
// inverse R2 in one pass; inverse R1).  The first stage reads T straight
// from HBM and the last writes it straight back, so T is read and written once
// and LDS holds the row between the stages.  Twiddles between the stages come
// from an LDS copy of the W_M2 table.
//
// Fallback k_conv_generic: the same transform for any 7-smooth M2 <= 4096 with
// radix-2/3/4/5/7/8 stages ping-ponged between two LDS buffers.
#pragma once
#include "rowpass_internal.h"

namespace spfft {

enum StageIO { kGlobalToLds = 0, kLdsToLds = 1, kLdsToGlobal = 2 };

// LDS row layout of k_conv_fast<R1, R2>.  Its LDS accesses are the stage-1 writes
// j R1 + r, the stage-2 reads j + r R1, the pivot writes j R2 + r and the stage-3
// reads j + r R2: lane strides of R1 or R2 entries, which for even radices land
// on few banks.  With D = gcd(R1, R2) even, entry i of a row is stored at
// i + floor(i / D) (row pitch L + L / D): every one of those accesses is still a
// per-thread base plus a compile-time offset (D divides R1, R2 and each base),
// and the strides become R1 + R1 / D and R2 + R2 / D.  (With D odd the padded
// strides conflict more, so those rows stay unpadded.)
constexpr int conv_gcd(int a, int b) { return b == 0 ? a : conv_gcd(b, a % b); }
constexpr int conv_pad_div(int r1, int r2) {
    return r2 > 1 && conv_gcd(r1, r2) % 2 == 0 ? conv_gcd(r1, r2) : 0;
}
constexpr int conv_row_pitch(int r1, int r2) {
    return conv_pad_div(r1, r2) ? r1 * r2 + r1 * r2 / conv_pad_div(r1, r2) : r1 * r2;
}
template <int D>
__device__ __forceinline__ int conv_lds_idx(int i) { return D ? i + i / D : i; }

// One Stockham stage of radix R over G2 rows of length L held in LDS (or
// streamed from / to T).  NS = product of the radices of the earlier stages.
// PIVOT: the last forward stage (NS = L / R), the product with B^, and the
// first inverse stage on the same R values (outputs j R + r), one LDS round
// trip for both; the inverse then runs its radices in reverse order.
template <int R, int L, int NS, int IO, bool INV, bool MUL, bool PIVOT = false, int D = 0>
__device__ __forceinline__ void cstage(const RowPassArgs& a, float2* buf, const float2* s_tw,
                                       const int* s_k1, int row0) {
    constexpr int MB = MaxBfly<R>::value;
    constexpr int nb = L / R;
    constexpr int LP = D ? L + L / D : L;  // LDS row pitch
    static_assert(!D || (nb % D == 0 && R % D == 0), "padded rows need D | nb and D | R");
    constexpr int step = L / (NS * R);
    const int total = a.G2 * nb;
    float2 v[MB][R];
    int gi[MB], ji[MB];
#pragma unroll
    for (int b = 0; b < MB; b++) {
        const int t = threadIdx.x + b * kThreads;
        gi[b] = -1;
        ji[b] = 0;
        if (t < total) {
            const int g = t / nb;
            const int j = t - g * nb;
            gi[b] = g;
            ji[b] = j;
            if (IO == kGlobalToLds) {
                const float2* src = a.T + (size_t)(row0 + g) * L;
                const bool live = s_k1[g] >= 0;
#pragma unroll
                for (int r = 0; r < R; r++) v[b][r] = live ? src[j + r * nb] : make_float2(0.f, 0.f);
            } else {
                const float2* src = buf + g * LP + conv_lds_idx<D>(j);
#pragma unroll
                for (int r = 0; r < R; r++) v[b][r] = src[conv_lds_idx<D>(r * nb)];
            }
        }
    }
    if (IO == kLdsToLds) __syncthreads();  // every read of the stage before any write
#pragma unroll
    for (int b = 0; b < MB; b++) {
        if (gi[b] < 0) continue;
        const int g = gi[b], j = ji[b];
        const int jq = j / NS;
        const int k = j - jq * NS;
        if (NS > 1 && k != 0) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                const float2 w = s_tw[r * k * step];
                v[b][r] = INV ? cmulc(v[b][r], w) : cmul(v[b][r], w);
            }
        }
        RegFft<R>::run(v[b], INV);
        int base = jq * NS * R + k;
        int stride = NS;
        if (MUL) {
            const int k1 = s_k1[g];
            const float2* bh = a.bhat + (size_t)(k1 < 0 ? 0 : k1) * L;
#pragma unroll
            for (int r = 0; r < R; r++) v[b][r] = cmul(v[b][r], bh[base + r * NS]);
        }
        if (PIVOT) {
            static_assert(!PIVOT || (MUL && !INV && NS * R == L), "the pivot is the last forward stage");
            RegFft<R>::run(v[b], true);
            base = j * R;
            stride = 1;
        }
        if constexpr (IO == kLdsToGlobal) {
            if (s_k1[g] >= 0) {
                float2* dst = a.T + (size_t)(row0 + g) * L;
#pragma unroll
                for (int r = 0; r < R; r++) dst[base + r * stride] = v[b][r];
            }
        } else {
            // padded rows: the LDS writes are the stage-1 and pivot ones, stride 1 from
            // a base j R with D | R
            static_assert(!D || NS == 1 || PIVOT, "padded LDS writes: stride 1 from a multiple of D");
            float2* dst = buf + g * LP + conv_lds_idx<D>(base);
#pragma unroll
            for (int r = 0; r < R; r++) dst[conv_lds_idx<D>(r * stride)] = v[b][r];
        }
    }
    if (IO != kLdsToGlobal) __syncthreads();
}

__device__ __forceinline__ void conv_setup(const RowPassArgs& a, float2* s_tw, int* s_k1, int row0) {
    const int L = a.M2;
    const int nrows = a.tile_rows * a.M1;
    for (int i = threadIdx.x; i < L; i += kThreads) s_tw[i] = a.tw2[i];
    for (int g = threadIdx.x; g < a.G2; g += kThreads)
        s_k1[g] = (row0 + g < nrows) ? (row0 + g) % a.M1 : -1;
    __syncthreads();
}

template <int R1, int R2>
__global__ __launch_bounds__(kThreads) void k_conv_fast(RowPassArgs a) {
    extern __shared__ float2 smem[];
    constexpr int L = R1 * R2;
    constexpr int D = conv_pad_div(R1, R2);
    float2* s_tw = smem;
    float2* buf = s_tw + L;
    int* s_k1 = reinterpret_cast<int*>(buf + (size_t)a.G2 * conv_row_pitch(R1, R2));
    const int row0 = blockIdx.x * a.G2;
    conv_setup(a, s_tw, s_k1, row0);
    if constexpr (R2 == 1) {
        cstage<R1, L, 1, kGlobalToLds, false, true>(a, buf, s_tw, s_k1, row0);
        cstage<R1, L, 1, kLdsToGlobal, true, false>(a, buf, s_tw, s_k1, row0);
    } else {
        cstage<R1, L, 1, kGlobalToLds, false, false, false, D>(a, buf, s_tw, s_k1, row0);
        cstage<R2, L, R1, kLdsToLds, false, true, true, D>(a, buf, s_tw, s_k1, row0);
        cstage<R1, L, R2, kLdsToGlobal, true, false, false, D>(a, buf, s_tw, s_k1, row0);
    }
}

// ---------------------------------------------------------------------------
// Generic fallback: radix-2/3/4/5/7/8 stages ping-ponged between two LDS buffers.
// ---------------------------------------------------------------------------
// One Stockham stage over G sequences of length L stored back to back in LDS
// (sequence g at src + g*L).  tw: W_L^i, i < L (forward; conjugated for inverse).
template <int R>
__device__ __forceinline__ void lds_stage(const float2* __restrict__ src, float2* __restrict__ dst,
                                          int L, int G, int Ns, const float2* __restrict__ tw,
                                          bool inv) {
    const int nb = L / R;
    const int total = G * nb;
    const int step = L / (Ns * R);
    const float inv_nb = 1.f / (float)nb;
    const float inv_ns = 1.f / (float)Ns;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int g = fdiv_small(t, nb, inv_nb);
        const int j = t - g * nb;
        const int jq = fdiv_small(j, Ns, inv_ns);
        const int k = j - jq * Ns;
        const float2* s = src + g * L;
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = s[j + r * nb];
        if (k != 0) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                const float2 w = tw[r * k * step];
                v[r] = inv ? cmulc(v[r], w) : cmul(v[r], w);
            }
        }
        Dft<R>::run(v, inv);
        float2* d = dst + g * L + jq * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) d[r * Ns] = v[r];
    }
}

// Batched LDS FFT: G sequences of length f.L in buffer a (scratch b of the
// same size).  Caller must __syncthreads() before.  Returns the buffer holding
// the result (a or b); ends with a barrier.
__device__ __forceinline__ float2* lds_fft(float2* a, float2* b, int G, const FftStages& f,
                                           const float2* tw, bool inv) {
    int Ns = 1;
    for (int s = 0; s < f.nstages; s++) {
        const int R = f.radix[s];
        switch (R) {
            case 2: lds_stage<2>(a, b, f.L, G, Ns, tw, inv); break;
            case 3: lds_stage<3>(a, b, f.L, G, Ns, tw, inv); break;
            case 4: lds_stage<4>(a, b, f.L, G, Ns, tw, inv); break;
            case 5: lds_stage<5>(a, b, f.L, G, Ns, tw, inv); break;
            case 7: lds_stage<7>(a, b, f.L, G, Ns, tw, inv); break;
            default: lds_stage<8>(a, b, f.L, G, Ns, tw, inv); break;
        }
        __syncthreads();
        float2* t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}



#ifdef SPFFT_DEFINE_CONV_GENERIC  // defined by the one translation unit that launches it
__global__ __launch_bounds__(kThreads) void k_conv_generic(RowPassArgs a) {
    extern __shared__ float2 smem[];
    const int L = a.M2, G = a.G2;
    float2* s_tw = smem;
    float2* bufA = s_tw + L;
    float2* bufB = bufA + (size_t)G * L;
    int* s_k1 = reinterpret_cast<int*>(bufB + (size_t)G * L);
    const int row0 = blockIdx.x * G;
    conv_setup(a, s_tw, s_k1, row0);
    const float inv_l = 1.f / (float)L;
    for (int q = threadIdx.x; q < G * L; q += kThreads) {
        const int g = fdiv_small(q, L, inv_l);
        const int n2 = q - g * L;
        bufA[q] = s_k1[g] >= 0 ? a.T[(size_t)(row0 + g) * L + n2] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    float2* f = lds_fft(bufA, bufB, G, a.f2, s_tw, false);
    float2* o = (f == bufA) ? bufB : bufA;
    for (int q = threadIdx.x; q < G * L; q += kThreads) {
        const int g = fdiv_small(q, L, inv_l);
        const int k2 = q - g * L;
        if (s_k1[g] >= 0) f[q] = cmul(f[q], a.bhat[(size_t)s_k1[g] * L + k2]);
    }
    __syncthreads();
    const float2* y = lds_fft(f, o, G, a.f2, s_tw, true);
    for (int q = threadIdx.x; q < G * L; q += kThreads) {
        const int g = fdiv_small(q, L, inv_l);
        const int n2 = q - g * L;
        if (s_k1[g] >= 0) a.T[(size_t)(row0 + g) * L + n2] = y[q];
    }
}
#endif

// Lengths M2 = R1 * R2 with a fast k_conv kernel, grouped by translation unit.
#define SPFFT_CONV1_LIST(X) \
    X(2, 1) X(3, 1) X(4, 1) X(5, 1) X(6, 1) X(7, 1) X(8, 1) X(9, 1) X(10, 1) X(12, 1) X(14, 1) \
    X(15, 1) X(16, 1) X(18, 1) X(20, 1) X(21, 1) X(24, 1) X(25, 1) X(27, 1) X(28, 1) X(30, 1) \
    X(32, 1) X(36, 1)
#define SPFFT_CONV2_LIST_A(X) \
    X(8, 8) X(9, 8) X(9, 9) X(12, 8) X(12, 9) X(12, 12) X(16, 8) X(16, 9) X(16, 12) X(16, 16) \
    X(18, 8) X(18, 9) X(18, 12) X(18, 16) X(18, 18)
#define SPFFT_CONV2_LIST_B(X) \
    X(24, 8) X(24, 9) X(24, 12) X(24, 16) X(24, 18) X(24, 24) X(25, 8) X(25, 9) X(25, 12) \
    X(25, 16) X(25, 18) X(25, 24) X(25, 25)
#define SPFFT_CONV2_LIST_C(X) \
    X(27, 8) X(27, 9) X(27, 12) X(27, 16) X(27, 18) X(27, 24) X(27, 25) X(27, 27) X(28, 8) \
    X(28, 9) X(28, 12) X(28, 16) X(28, 18) X(28, 24) X(28, 25) X(28, 27) X(28, 28)
#define SPFFT_CONV2_LIST_D(X) \
    X(32, 8) X(32, 9) X(32, 12) X(32, 16) X(32, 18) X(32, 24) X(32, 25) X(32, 27) X(32, 28) \
    X(32, 32) X(36, 8) X(36, 9) X(36, 12) X(36, 16) X(36, 18) X(36, 24) X(36, 25) X(36, 27) \
    X(36, 28) X(36, 32) X(36, 36)

}  // namespace spfft
