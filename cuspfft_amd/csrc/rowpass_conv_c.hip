// rowpass_conv_c.hip -- k_conv instantiations (SPFFT_CONV2_LIST_C); see conv_impl.h.
#include "conv_impl.h"

namespace spfft {

bool launch_conv_c(const RowPassArgs& a, hipStream_t s, hipError_t* err) {
    const dim3 grid(((size_t)a.tile_rows * a.M1 + a.G2 - 1) / a.G2);
    const size_t lds = conv_lds_bytes(a);
#define SPFFT_CASE(r1, r2)                                                                     \
    if (a.f2.nstages == ((r2) == 1 ? 1 : 2) && a.f2.radix[0] == (r1) &&                        \
        ((r2) == 1 || a.f2.radix[1] == (r2))) {                                                \
        hipLaunchKernelGGL((k_conv_fast<r1, r2>), grid, dim3(kThreads), lds, s, a);            \
        *err = hipGetLastError();                                                              \
        return true;                                                                           \
    }
    SPFFT_CONV2_LIST_C(SPFFT_CASE)
#undef SPFFT_CASE
    return false;
}

}  // namespace spfft
