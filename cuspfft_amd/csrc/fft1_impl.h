// fft1_impl.h -- k_fft1 / k_ifft1, the length-M1 sub-FFTs of the four-step
// Bluestein row pass (see rowpass.hip for the pipeline); each rowpass_fft1_*.hip
// instantiates a group of M1 lengths.
//
// M >= 2C - 1, so the entries w = n2 + M2 n1 < C have n1 < NZ = (M1 + 1) / 2:
// only those inputs of k_fft1 can be non-zero and only those outputs of
// k_ifft1 are written; the codelets are pruned to them at compile time (the
// files are built without signed zeros, so x + 0 folds to x).
//
// One (u, n2) column of the four-step matrix per thread; block = 256 threads =
// 64 consecutive n2 x 4 rows u, so a wave has one row.  Rows are buffers (see
// buffer_io.h): the per-lane part of an offset is n2, the strided part M2 n1 or
// M2 k1 a scalar offset, and the bounds of a row do the zero padding of x
// (n >= C), drop the writes of v > C/2 and of lanes past M2 (offset kOob).
#pragma once
#include "buffer_io.h"
#include "rowpass_internal.h"

namespace spfft {

namespace {

// W_M^{k1 n2} for k1 < M1 from A + B - 2 entries of the table twM[k1][n2] =
// W_M^{k1 n2}: with k1 = A i + j (j < A), W_M^{k1 n2} = twM[A i][n2] twM[j][n2],
// so a thread loads rows j < A and A i of its column and forms the rest with
// one complex product each (instead of one table load per k1).
template <int M1>
struct TwSplit {
    static constexpr int A = M1 >= 49 ? 8 : M1 >= 25 ? 6 : M1 >= 9 ? 4 : 2;
    static constexpr int B = (M1 + A - 1) / A;
    float2 lo[A], hi[B];
    __device__ __forceinline__ void load(bio::rsrc_t r_w, int vo, int M2) {
        lo[0] = hi[0] = make_float2(1.f, 0.f);
#pragma unroll
        for (int j = 1; j < A; j++) lo[j] = bio::load2(r_w, vo, M2 * j * 8);
#pragma unroll
        for (int i = 1; i < B; i++) hi[i] = bio::load2(r_w, vo, M2 * A * i * 8);
    }
    // W_M^{k1 n2}, k1 a compile-time index after unrolling
    __device__ __forceinline__ float2 at(int k1) const {
        const int i = k1 / A, j = k1 - (k1 / A) * A;
        return i == 0 ? lo[j] : j == 0 ? hi[i] : cmul(hi[i], lo[j]);
    }
};

template <int M1>
__global__ __launch_bounds__(kThreads) void k_fft1(RowPassArgs a) {
    const int n2 = blockIdx.x * 64 + (threadIdx.x & 63);
    const int ul = blockIdx.y * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (ul >= a.tile_rows) return;  // whole wave
    const int M2 = a.M2;
    const int vo = n2 < M2 ? n2 * 8 : bio::kOob;
    const bio::rsrc_t r_x = bio::make_rsrc(a.X0 + (size_t)ul * a.ldx, (uint32_t)a.C * 8u);
    const bio::rsrc_t r_t = bio::make_rsrc(a.T + (size_t)ul * a.M, (uint32_t)a.M * 8u);
    const bio::rsrc_t r_w = bio::make_rsrc(a.twM, (uint32_t)a.M * 8u);
    constexpr int NZ = (M1 + 1) / 2;
    float2 v[M1];
#pragma unroll
    for (int n1 = 0; n1 < M1; n1++)
        v[n1] = n1 < NZ ? bio::load2_stream(r_x, vo, M2 * n1 * 8) : make_float2(0.f, 0.f);
    TwSplit<M1> tw;
    tw.load(r_w, vo, M2);
    RegFft<M1>::run(v, false);
    bio::store2_stream(v[0], r_t, vo, 0);
#pragma unroll
    for (int k1 = 1; k1 < M1; k1++) bio::store2_stream(cmul(v[k1], tw.at(k1)), r_t, vo, M2 * k1 * 8);
}

// At least two workgroups (two waves per SIMD) per CU: the length-64 column and
// its loads otherwise take a few registers more and leave one wave per SIMD
// to wait on HBM.
template <int M1>
__global__ __launch_bounds__(kThreads, 2) void k_ifft1(RowPassArgs a) {
    const int n2 = blockIdx.x * 64 + (threadIdx.x & 63);
    const int ul = blockIdx.y * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (ul >= a.tile_rows) return;  // whole wave
    const int M2 = a.M2;
    const bool lane_ok = n2 < M2;
    const int vo = lane_ok ? n2 * 8 : bio::kOob;
    const bio::rsrc_t r_t = bio::make_rsrc(a.T + (size_t)ul * a.M, (uint32_t)a.M * 8u);
    const bio::rsrc_t r_w = bio::make_rsrc(a.twM, (uint32_t)a.M * 8u);
    TwSplit<M1> tw;
    tw.load(r_w, vo, M2);
    float2 v[M1];
#pragma unroll
    for (int k1 = 0; k1 < M1; k1++) v[k1] = bio::load2_stream(r_t, vo, M2 * k1 * 8);
#pragma unroll
    for (int k1 = 1; k1 < M1; k1++) v[k1] = cmulc(v[k1], tw.at(k1));
    RegFft<M1>::run(v, true);
    // X[u][w] = h[w] y[w] for every w < C.  The input is real, so
    // X[R-u][v] = conj(X[u][(C - v) mod C]): row u also yields its mirror row.
    // Row u: a buffer of C/2 + 1 entries.  Mirror row: X[R-u][C - w] at offset
    // M - 1 - w = M2 (M1 - 1 - n1) + (M2 - 1 - n2) from a base M - 1 - C entries
    // before it, a buffer ending at entry C/2 (writes of C - w > C/2 dropped);
    // w >= C takes kOob, w = 0 is written apart.
    const int u = a.u_base + ul;
    const bool mirror = u > 0 && a.R - u > u;
    const int dm = a.M - 1 - a.C;
    const bio::rsrc_t r_h = bio::make_rsrc(a.h, (uint32_t)a.C * 8u);
    const bio::rsrc_t r_o = bio::make_rsrc(a.out + (size_t)u * a.P, (uint32_t)a.V * 8u);
    const bio::rsrc_t r_m =
        bio::make_rsrc(mirror ? a.out_m + (size_t)(a.R - u) * a.P - dm : a.out, mirror ? (uint32_t)(dm + a.V) * 8u : 0u);
    const bio::rsrc_t r_m0 = bio::make_rsrc(mirror ? a.out_m + (size_t)(a.R - u) * a.P : a.out, mirror ? 8u : 0u);
    const int n1c = lane_ok ? (a.C - n2 + M2 - 1) / M2 : 0;  // columns w = n2 + M2 n1 < C
    const int vm0 = (M2 - 1 - n2) * 8;
#pragma unroll
    for (int n1 = 0; n1 < (M1 + 1) / 2; n1++) {
        const float2 x = cmul(v[n1], bio::load2(r_h, vo, M2 * n1 * 8));
        bio::store2_stream(x, r_o, vo, M2 * n1 * 8);
        bio::store2_stream(make_float2(x.x, -x.y), r_m, n1 < n1c ? vm0 : bio::kOob, M2 * (M1 - 1 - n1) * 8);
        if (n1 == 0 && n2 == 0) bio::store2_stream(make_float2(x.x, -x.y), r_m0, 0, 0);
    }
}

}  // namespace

}  // namespace spfft

#define SPFFT_FFT1_LAUNCH_CASE(m)                                                    \
    case m:                                                                          \
        if (inverse) hipLaunchKernelGGL((k_ifft1<m>), grid, dim3(kThreads), 0, s, a); \
        else hipLaunchKernelGGL((k_fft1<m>), grid, dim3(kThreads), 0, s, a);         \
        break;
