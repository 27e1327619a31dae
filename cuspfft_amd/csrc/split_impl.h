// split_impl.h -- k_splitcols<N1>: the second pass of the split row transform.
//
// A width C = N1 * N2 whose Bluestein length has no on-chip shape is transformed in
// two passes over a tile buffer T of C entries per row (Cooley-Tukey with
// c = n1 + N1 n2 and k = k2 + N2 k1):
//   Y[k2 + N2 k1] = sum_{n1} W_N1^{n1 k1} W_C^{n1 k2} Z_{n1}[k2],
//   Z_{n1}[k2]    = sum_{n2} x[n1 + N1 n2] W_N2^{n2 k2}.
// Pass A is the on-chip row kernel of a plan for width N2 run over the N1 sequences
// of every row (the column pass writes X0[u] as N1 contiguous runs x[n1 + N1 n2],
// n2 < N2), which leaves Z_{n1} at T[u][n1 P2 + k2] (P2: N2 rounded up to 16 entries,
// so every run starts a 128-byte line).  Pass B (this kernel) takes 256 columns k2 of
// one row u per workgroup, one per thread: it reads the N1 entries T[u][n1 P2 + k2] (a
// wave reads 64 consecutive k2: whole lines), multiplies them by
// W_C^{n1 k2} (exact integer exponent n1 k2 < C, two-level LDS table), runs the
// length-N1 DFT in registers and writes Y[k] to row u (k <= C/2) and conj(Y[k]) to
// the mirror row R - u at C - k (k > C/2, and k = 0) -- for fixed k1 a wave's 64
// outputs are consecutive entries of one row.
//
// Length-N1 DFTs: the register codelets (RegFft) for N1 <= 64 whose prime factors
// are at most 31; for odd primes 37 .. 127 the symmetric-pair direct DFT
//   s_q = v_q + v_{N-q},  d_q = v_q - v_{N-q}   (q = 1 .. H, H = (N - 1) / 2)
//   y_0 = v_0 + sum s_q,  y_p = A_p - i B_p,  y_{N-p} = A_p + i B_p,
//   A_p = v_0 + sum_q s_q cos(2 pi p q / N),  B_p = sum_q d_q sin(2 pi p q / N)
// with a runtime loop over p (its outputs are stored as they are formed, so the
// registers hold the N inputs and two sums) and the cosines and sines of row p read
// from a host table at a wave-uniform address.
#pragma once
#include "codelets.h"
#include "spfft_device.h"

namespace spfft {

// Instantiated lengths (N1): register codelets (two lists, one source file each), then
// the direct odd-prime DFTs.
#define SPFFT_SPLIT_REG_LIST_A(X)                                                                 \
    X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17)       \
    X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31)
#define SPFFT_SPLIT_REG_LIST_B(X)                                                                 \
    X(32) X(33) X(34) X(35) X(36) X(38) X(39) X(40) X(42) X(44) X(45) X(46) X(48) X(49) X(50)     \
    X(51) X(52) X(54) X(55) X(56) X(57) X(58) X(60) X(62) X(63) X(64)
#define SPFFT_SPLIT_DIRECT_LIST(X)                                                                \
    X(37) X(41) X(43) X(47) X(53) X(59) X(61) X(67) X(71) X(73) X(79) X(83) X(89) X(97) X(101)   \
    X(103) X(107) X(109) X(113) X(127)

constexpr int kSplitThreads = 256;
typedef float split_f32x2 __attribute__((ext_vector_type(2)));

// Y[k] of row u (and its mirror row) from the thread's column k2, k = k2 + N2 k1.
struct SplitOut {
    float2* o;   // row u
    float2* om;  // row R - u (valid when mirror)
    int k2, N2, C, V;
    bool mirror;
    __device__ __forceinline__ void put(int k1, float2 y) const {
        const int k = k2 + N2 * k1;
        const split_f32x2 y0 = {y.x, y.y}, y1 = {y.x, -y.y};
        if (k < V) __builtin_nontemporal_store(y0, reinterpret_cast<split_f32x2*>(o + k));
        const int vm = k == 0 ? 0 : C - k;
        if (mirror && vm < V) __builtin_nontemporal_store(y1, reinterpret_cast<split_f32x2*>(om + vm));
    }
};

// The table through the constant address space: a wave-uniform address becomes a
// scalar load (the cosines and sines arrive in SGPRs, no VGPRs or vector loads).
typedef const __attribute__((address_space(4))) float* split_const_ptr;

template <int N1>
__device__ __forceinline__ void split_dft_direct(float2* v, split_const_ptr tab, const SplitOut& out) {
    constexpr int H = (N1 - 1) / 2;
    float2 y0 = v[0];
#pragma unroll
    for (int q = 1; q <= H; q++) {
        const float2 a = v[q], b = v[N1 - q];
        v[q] = make_float2(a.x + b.x, a.y + b.y);       // s_q
        v[N1 - q] = make_float2(a.x - b.x, a.y - b.y);  // d_q
        y0.x += v[q].x;
        y0.y += v[q].y;
    }
    out.put(0, y0);
#pragma unroll 1
    for (int p = 1; p <= H; p++) {
        split_const_ptr row = tab + 2 * (p - 1) * H;  // (cos, sin)(2 pi p q / N1), q = 1 .. H
        float2 A = v[0], B = make_float2(0.f, 0.f);
#pragma unroll
        for (int q = 1; q <= H; q++) {
            const float c = row[2 * (q - 1)], sn = row[2 * (q - 1) + 1];
            A.x = fmaf(v[q].x, c, A.x);
            A.y = fmaf(v[q].y, c, A.y);
            B.x = fmaf(v[N1 - q].x, sn, B.x);
            B.y = fmaf(v[N1 - q].y, sn, B.y);
        }
        // forward: y_p = A - i B, y_{N-p} = A + i B
        out.put(p, make_float2(A.x + B.y, A.y - B.x));
        out.put(N1 - p, make_float2(A.x - B.y, A.y + B.x));
    }
}

template <int N1, bool DIRECT>
__global__ __launch_bounds__(kSplitThreads) void k_splitcols(RowPassArgs a) {
    extern __shared__ float2 s_w[];  // W_C^e: 256 low entries, then W_C^{256 i}
    const int N2 = a.split_n2, C = a.C;
    const int nhi = (C + 255) >> 8;
    for (int i = threadIdx.x; i < 256 + nhi; i += kSplitThreads) s_w[i] = a.split_w[i];
    const int ul = blockIdx.y;
    const int u = a.u_base + ul;
    const int k2 = blockIdx.x * kSplitThreads + threadIdx.x;
    const bool live = k2 < N2;
    float2 v[N1];
    // lanes past N2 read column 0 (in bounds) and store nothing
    const int P2 = a.split_p2;
    const float2* t = a.T + (size_t)ul * N1 * P2 + (live ? k2 : 0);
#pragma unroll
    for (int n1 = 0; n1 < N1; n1++) {
        const split_f32x2 x = __builtin_nontemporal_load(reinterpret_cast<const split_f32x2*>(t + (size_t)n1 * P2));
        v[n1] = make_float2(x[0], x[1]);
    }
    __syncthreads();
    if (!live) return;
#pragma unroll
    for (int n1 = 1; n1 < N1; n1++) {
        const int e = n1 * k2;  // < N1 N2 = C
        v[n1] = cmul(v[n1], cmul(s_w[256 + (e >> 8)], s_w[e & 255]));
    }
    SplitOut out;
    out.mirror = u > 0 && a.R - u > u;
    out.o = a.out + (size_t)u * a.P;
    out.om = a.out_m + (size_t)(out.mirror ? a.R - u : 0) * a.P;
    out.k2 = k2;
    out.N2 = N2;
    out.C = C;
    out.V = a.V;
    if constexpr (DIRECT) {
        split_dft_direct<N1>(v, (split_const_ptr)(a.split_dft), out);
    } else {
        RegFft<N1>::run(v, false);
#pragma unroll
        for (int k1 = 0; k1 < N1; k1++) out.put(k1, v[k1]);
    }
}

// launch helper of one instantiation list (rowpass_split_{a,b,c}.hip)
#define SPFFT_SPLIT_LAUNCH_CASE(n, d)                                                           \
    if (a.split_n1 == n) {                                                                      \
        hipLaunchKernelGGL((k_splitcols<n, d>), grid, dim3(kSplitThreads), lds, s, a);          \
        return hipGetLastError();                                                               \
    }

}  // namespace spfft
