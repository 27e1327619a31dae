// rowpass_fft1_b.hip -- k_fft1 / k_ifft1 for the M1 lengths of group B
// (see fft1_impl.h and rowpass.hip for the pipeline).
#include "fft1_impl.h"

namespace spfft {

hipError_t launch_fft1_group_b(const RowPassArgs& a, dim3 grid, hipStream_t s, bool inverse) {
    switch (a.M1) {
        SPFFT_M1_GROUP_B(SPFFT_FFT1_LAUNCH_CASE)
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace spfft
