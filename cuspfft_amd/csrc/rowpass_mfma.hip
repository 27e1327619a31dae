// rowpass_mfma.hip -- precision form of the length-M1 sub-FFTs (M1 = 16 or 32):
// the radix-M1 butterfly is a complex DFT-matrix product on the gfx950 matrix
// cores, in BF16 with a hi/lo split of both operands and FP32 accumulation.
//
// With x = xr + i xi, the length-L DFT y = F x is the real product
//     [yr]   [Fr  -Fi] [xr]        F[p][q] = W_L^{pq}
//     [yi] = [Fi   Fr] [xi]
// of a 2L x 2L matrix G and a 2L-vector.  A wave stacks 32 such vectors as the
// columns of B (2L x 32) and issues v_mfma_f32_32x32x16_bf16 over 2L/32 output
// tiles and 2L/16 k-steps.  Each operand is split into BF16 parts, v = vh + vl
// (vh = bf16(v), vl = bf16(v - vh), ~16 significant bits together), and
//     G B ~= Gh Bh + Gh Bl + Gl Bh        (the Gl Bl term, ~2^-18, is dropped)
// accumulates in FP32.  G's parts are split on the host from FP64 values.
//
// Operand layouts (verified on gfx950): A (32 x 16): lane l holds row l % 32,
// k = (l / 32) * 8 + j; B (16 x 32): lane l holds column l % 32, same k;
// D (32 x 32): lane l holds column l % 32, rows 8 g + (l / 32) * 4 + i in
// element 4 g + i.
//
// k_fft1_mfma<L>:  per (u, n2) vector: load x[M2 n1 + n2] (zero for c >= C),
//                  DFT_L, times W_M^{k1 n2}, store T[u][k1][n2].
// k_ifft1_mfma<L>: load T[u][k1][n2] times W_M^{-k1 n2}, inverse DFT_L, chirp,
//                  and the writes of row u and its mirror row R - u.
#include "rowpass_internal.h"

namespace spfft {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// Splits v into BF16 hi and lo parts.
__device__ __forceinline__ void split_bf16(float v, __bf16& hi, __bf16& lo) {
    hi = (__bf16)v;
    lo = (__bf16)(v - (float)hi);
}

// The G-operand fragments of one lane for the whole 2L x 2L matrix:
// frag[t][s] holds G[32 t + l % 32][16 s + 8 (l / 32) + j], j < 8.
template <int L>
struct GFrags {
    static constexpr int RT = 2 * L / 32;  // output row tiles
    static constexpr int KS = 2 * L / 16;  // k-steps
    bf16x8 hi[RT][KS], lo[RT][KS];

    // dft: hi/lo parts of cos(2 pi i / L) and sin(2 pi i / L), i < L, as
    // [4][L] floats: cos_hi, cos_lo, sin_hi, sin_lo.
    __device__ __forceinline__ void load(const float* dft, bool inv) {
        const int l = threadIdx.x & 63;
#pragma unroll
        for (int t = 0; t < RT; t++)
#pragma unroll
            for (int s = 0; s < KS; s++)
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int m = 32 * t + (l & 31);
                    const int k = 16 * s + 8 * (l >> 5) + j;
                    const int p = m % L, q = k % L;
                    const bool out_im = m >= L, in_im = k >= L;
                    const int idx = (p * q) % L;
                    // F = cos - i sin (forward) or cos + i sin (inverse)
                    // G = [[Fr, -Fi], [Fi, Fr]]: re/re and im/im -> cos; im-out/re-in -> Fi;
                    // re-out/im-in -> -Fi.  Fi = -sin (forward), +sin (inverse).
                    float h, o;
                    if (out_im == in_im) {
                        h = dft[idx];
                        o = dft[L + idx];
                    } else {
                        const float sg = ((out_im ? 1.f : -1.f) * (inv ? 1.f : -1.f));
                        h = sg * dft[2 * L + idx];
                        o = sg * dft[3 * L + idx];
                    }
                    hi[t][s][j] = (__bf16)h;
                    lo[t][s][j] = (__bf16)o;
                }
    }
};

// Runs y = G x for the 32 vectors of the wave.  xcomp[8 s + j] is the lane's
// share of its vector: component k = 16 s + 8 h + j (h = l / 32), the real
// (k < L) or imaginary part of x[k % L].  On return acc[t] holds output rows
// 32 t + 8 g + 4 h + i in element 4 g + i.
template <int L>
__device__ __forceinline__ void mfma_dft(const GFrags<L>& g, const float* xcomp, floatx16* acc) {
    constexpr int RT = GFrags<L>::RT, KS = GFrags<L>::KS;
    bf16x8 bh[KS], bl[KS];
#pragma unroll
    for (int s = 0; s < KS; s++)
#pragma unroll
        for (int j = 0; j < 8; j++) {
            __bf16 h, lo;
            split_bf16(xcomp[8 * s + j], h, lo);
            bh[s][j] = h;
            bl[s][j] = lo;
        }
#pragma unroll
    for (int t = 0; t < RT; t++) {
        floatx16 c;
#pragma unroll
        for (int i = 0; i < 16; i++) c[i] = 0.f;
#pragma unroll
        for (int s = 0; s < KS; s++) {
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(g.hi[t][s], bh[s], c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(g.hi[t][s], bl[s], c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(g.lo[t][s], bh[s], c, 0, 0, 0);
        }
        acc[t] = c;
    }
}

// Element e = 4 g + i of accumulator tile t holds output row m = 32 t + 8 g + 4 h + i.
// Rows m < L are real parts; the imaginary part of the same output sits at row
// m + L: in tile t + L/32 at the same element (L >= 32), or in the same tile at
// element e + 8 (L = 16).
template <int L>
struct OutMap {
    static constexpr int RE_T = L >= 32 ? L / 32 : 1;  // tiles holding real parts
    static constexpr int RE_G = L >= 32 ? 4 : 2;       // g values holding real parts
    static constexpr int IM_DT = L >= 32 ? L / 32 : 0;
    static constexpr int IM_DE = L >= 32 ? 0 : 8;
};

template <int L>
__global__ __launch_bounds__(kThreads) void k_fft1_mfma(RowPassArgs a) {
    constexpr int KS = GFrags<L>::KS, RT = GFrags<L>::RT;
    const int lane = threadIdx.x & 63, h = lane >> 5;
    const int n2 = blockIdx.x * 32 + (lane & 31);
    const int ul = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (ul >= a.tile_rows) return;  // whole wave leaves together
    GFrags<L> g;
    g.load(a.mfma_dft, false);
    const bool live = n2 < a.M2;
    const float2* x0 = a.X0 + (size_t)ul * a.ldx;
    // the lane's share: x[q] for q = 16 s + 8 h + j (s < L/16); its real part is
    // component q (k-step s), its imaginary part component q + L (k-step s + L/16)
    float xc[8 * KS];
#pragma unroll
    for (int s = 0; s < KS / 2; s++)
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int n1 = 16 * s + 8 * h + j;
            const int c = n2 + a.M2 * n1;
            const float2 x = (live && c < a.C) ? x0[c] : make_float2(0.f, 0.f);
            xc[8 * s + j] = x.x;
            xc[8 * (s + KS / 2) + j] = x.y;
        }
    floatx16 acc[RT];
    mfma_dft<L>(g, xc, acc);
    if (!live) return;
    using OM = OutMap<L>;
    float2* t0 = a.T + (size_t)ul * a.M + n2;
    const float2* tw = a.twM + n2;
#pragma unroll
    for (int t = 0; t < OM::RE_T; t++)
#pragma unroll
        for (int gg = 0; gg < OM::RE_G; gg++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int p = 32 * t + 8 * gg + 4 * h + i;
                const float2 y = make_float2(acc[t][4 * gg + i], acc[t + OM::IM_DT][4 * gg + i + OM::IM_DE]);
                t0[(size_t)p * a.M2] = p == 0 ? y : cmul(y, tw[(size_t)p * a.M2]);
            }
}

template <int L>
__global__ __launch_bounds__(kThreads) void k_ifft1_mfma(RowPassArgs a) {
    constexpr int KS = GFrags<L>::KS, RT = GFrags<L>::RT;
    const int lane = threadIdx.x & 63, h = lane >> 5;
    const int n2 = blockIdx.x * 32 + (lane & 31);
    const int ul = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (ul >= a.tile_rows) return;
    GFrags<L> g;
    g.load(a.mfma_dft, true);
    const bool live = n2 < a.M2;
    const float2* t0 = a.T + (size_t)ul * a.M + n2;
    const float2* tw = a.twM + n2;
    float xc[8 * KS];
#pragma unroll
    for (int s = 0; s < KS / 2; s++)
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k1 = 16 * s + 8 * h + j;
            float2 x = make_float2(0.f, 0.f);
            if (live) {
                x = t0[(size_t)k1 * a.M2];
                if (k1) x = cmulc(x, tw[(size_t)k1 * a.M2]);
            }
            xc[8 * s + j] = x.x;
            xc[8 * (s + KS / 2) + j] = x.y;
        }
    floatx16 acc[RT];
    mfma_dft<L>(g, xc, acc);
    if (!live) return;
    const int u = a.u_base + ul;
    const bool mirror = u > 0 && a.R - u > u;
    float2* o = a.out + (size_t)u * a.P;
    float2* om = mirror ? a.out_m + (size_t)(a.R - u) * a.P : a.out + (size_t)u * a.P;
    using OM = OutMap<L>;
#pragma unroll
    for (int t = 0; t < OM::RE_T; t++)
#pragma unroll
        for (int gg = 0; gg < OM::RE_G; gg++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int p = 32 * t + 8 * gg + 4 * h + i;
                const int w = n2 + a.M2 * p;
                if (w < a.C) {
                    const float2 y = make_float2(acc[t][4 * gg + i], acc[t + OM::IM_DT][4 * gg + i + OM::IM_DE]);
                    const float2 x = cmul(y, a.h[w]);
                    if (w < a.V) o[w] = x;
                    const int vm = w == 0 ? 0 : a.C - w;
                    if (mirror && vm < a.V) om[vm] = make_float2(x.x, -x.y);
                }
            }
}

}  // namespace

bool mfma_supports_m1(int m1) { return m1 == 16 || m1 == 32; }

hipError_t launch_fft1_mfma(const RowPassArgs& a, hipStream_t s, bool inverse) {
    const dim3 grid((a.M2 + 31) / 32, (a.tile_rows + 3) / 4);
    switch (a.M1) {
        case 16:
            if (inverse) hipLaunchKernelGGL((k_ifft1_mfma<16>), grid, dim3(kThreads), 0, s, a);
            else hipLaunchKernelGGL((k_fft1_mfma<16>), grid, dim3(kThreads), 0, s, a);
            break;
        case 32:
            if (inverse) hipLaunchKernelGGL((k_ifft1_mfma<32>), grid, dim3(kThreads), 0, s, a);
            else hipLaunchKernelGGL((k_fft1_mfma<32>), grid, dim3(kThreads), 0, s, a);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace spfft
