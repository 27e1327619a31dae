// rowpass_fft1_a.hip -- k_fft1 / k_ifft1 for the M1 lengths of group A
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
    float2* o = a.out + (size_t)(a.u_base + ul) * a.P;
#pragma unroll
    for (int n1 = 0; n1 < M1; n1++) {
        const int vi = n2 + a.M2 * n1;
        if (vi < a.V) o[vi] = cmul(v[n1], a.h[vi]);
    }
}

}  // namespace

hipError_t launch_fft1_group_a(const RowPassArgs& a, dim3 grid, hipStream_t s, bool inverse) {
    switch (a.M1) {
#define SPFFT_CASE(m)                                                                   \
    case m:                                                                             \
        if (inverse) hipLaunchKernelGGL((k_ifft1<m>), grid, dim3(kThreads), 0, s, a);   \
        else hipLaunchKernelGGL((k_fft1<m>), grid, dim3(kThreads), 0, s, a);            \
        break;
        SPFFT_M1_GROUP_A(SPFFT_CASE)
#undef SPFFT_CASE
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace spfft
