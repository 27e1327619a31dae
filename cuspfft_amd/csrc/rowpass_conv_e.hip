// rowpass_conv_e.hip -- k_conv instantiations (SPFFT_CONV1_LIST); see conv_impl.h.
#define SPFFT_DEFINE_CONV_GENERIC
#include "conv_impl.h"

namespace spfft {

bool launch_conv_e(const RowPassArgs& a, hipStream_t s, hipError_t* err) {
    const dim3 grid(((size_t)a.tile_rows * a.M1 + a.G2 - 1) / a.G2);
    const size_t lds = conv_lds_bytes(a);
#define SPFFT_CASE(r1, r2)                                                                     \
    if (a.f2.nstages == ((r2) == 1 ? 1 : 2) && a.f2.radix[0] == (r1) &&                        \
        ((r2) == 1 || a.f2.radix[1] == (r2))) {                                                \
        hipLaunchKernelGGL((k_conv_fast<r1, r2>), grid, dim3(kThreads), lds, s, a);            \
        *err = hipGetLastError();                                                              \
        return true;                                                                           \
    }
    SPFFT_CONV1_LIST(SPFFT_CASE)
#undef SPFFT_CASE
    if (a.conv_kind == 0) {
        hipLaunchKernelGGL(k_conv_generic, grid, dim3(kThreads), lds, s, a);
        *err = hipGetLastError();
        return true;
    }
    return false;
}

size_t conv_lds_bytes(const RowPassArgs& a) {
    const size_t bufs = a.conv_kind == 0 ? 2 : 1;
    // k_conv_fast rows may be padded (conv_row_pitch)
    const size_t pitch = a.conv_kind == 1 && a.f2.nstages == 2 ? conv_row_pitch(a.f2.radix[0], a.f2.radix[1])
                                                                : (size_t)a.M2;
    return ((size_t)a.M2 + bufs * (size_t)a.G2 * pitch) * sizeof(float2) + (size_t)a.G2 * sizeof(int);
}

bool conv_fast_supported(int r1, int r2) {
#define SPFFT_CASE(x, y) if (r1 == (x) && r2 == (y)) return true;
    SPFFT_CONV1_LIST(SPFFT_CASE)
    SPFFT_CONV2_LIST_A(SPFFT_CASE)
    SPFFT_CONV2_LIST_B(SPFFT_CASE)
    SPFFT_CONV2_LIST_C(SPFFT_CASE)
    SPFFT_CONV2_LIST_D(SPFFT_CASE)
#undef SPFFT_CASE
    return false;
}

int conv_max_butterflies(int r) {
    switch (r) {
#define SPFFT_CASE(x, y) case x: return MaxBfly<x>::value;
        SPFFT_CONV1_LIST(SPFFT_CASE)
#undef SPFFT_CASE
        default: return 0;
    }
}

}  // namespace spfft
