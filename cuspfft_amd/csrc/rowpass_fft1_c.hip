// rowpass_fft1_c.hip -- k_fft1 / k_ifft1 for the M1 lengths of group C
// (see rowpass.hip for the pipeline).
#include "rowpass_internal.h"

namespace spfft {

namespace {

// ---------------------------------------------------------------------------
// k_fft1 / k_ifft1: one (u, n2) column of the four-step matrix per thread.
// block = 256 threads = 64 consecutive n2 x 4 rows u.
// ---------------------------------------------------------------------------
template <int M1>
__global__ __launch_bounds__(kThreads) void k_fft1(RowPassArgs a) {
    const int n2 = blockIdx.x * 64 + (threadIdx.x & 63);
    const int ul = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (n2 >= a.M2 || ul >= a.tile_rows) return;
    const float2* x0 = a.X0 + (size_t)ul * a.ldx;
    float2 v[M1];
#pragma unroll
    for (int n1 = 0; n1 < M1; n1++) {
        const int c = n2 + a.M2 * n1;
        v[n1] = c < a.C ? x0[c] : make_float2(0.f, 0.f);
    }
    RegFft<M1>::run(v, false);
    float2* t = a.T + (size_t)ul * a.M + n2;
    const float2* tw = a.twM + n2;
    t[0] = v[0];
#pragma unroll
    for (int k1 = 1; k1 < M1; k1++) t[(size_t)k1 * a.M2] = cmul(v[k1], tw[(size_t)k1 * a.M2]);
}

template <int M1>
__global__ __launch_bounds__(kThreads) void k_ifft1(RowPassArgs a) {
    const int n2 = blockIdx.x * 64 + (threadIdx.x & 63);
    const int ul = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (n2 >= a.M2 || ul >= a.tile_rows) return;
    const float2* t = a.T + (size_t)ul * a.M + n2;
    const float2* tw = a.twM + n2;
    float2 v[M1];
    v[0] = t[0];
#pragma unroll
    for (int k1 = 1; k1 < M1; k1++) v[k1] = cmulc(t[(size_t)k1 * a.M2], tw[(size_t)k1 * a.M2]);
    RegFft<M1>::run(v, true);
    // X[u][w] = h[w] y[w] for every w < C.  The input is real, so
    // X[R-u][v] = conj(X[u][(C - v) mod C]): row u also yields its mirror row.
    const int u = a.u_base + ul;
    const bool mirror = u > 0 && a.R - u > u;
    float2* o = a.out + (size_t)u * a.P;
    float2* om = a.out + (size_t)(mirror ? a.R - u : u) * a.P;
#pragma unroll
    for (int n1 = 0; n1 < M1; n1++) {
        const int w = n2 + a.M2 * n1;
        if (w < a.C) {
            const float2 x = cmul(v[n1], a.h[w]);
            if (w < a.V) o[w] = x;
            const int vm = w == 0 ? 0 : a.C - w;
            if (mirror && vm < a.V) om[vm] = make_float2(x.x, -x.y);
        }
    }
}

}  // namespace

hipError_t launch_fft1_group_c(const RowPassArgs& a, dim3 grid, hipStream_t s, bool inverse) {
    switch (a.M1) {
#define SPFFT_CASE(m)                                                                   \
    case m:                                                                             \
        if (inverse) hipLaunchKernelGGL((k_ifft1<m>), grid, dim3(kThreads), 0, s, a);   \
        else hipLaunchKernelGGL((k_fft1<m>), grid, dim3(kThreads), 0, s, a);            \
        break;
        SPFFT_M1_GROUP_C(SPFFT_CASE)
#undef SPFFT_CASE
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace spfft
