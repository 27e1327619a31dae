// rowpass_split_c.hip -- k_splitcols instantiations (see split_impl.h).
#include "spfft_kernels.h"
#include "split_impl.h"

namespace spfft {

hipError_t launch_split_cols_c(const RowPassArgs& a, dim3 grid, size_t lds, hipStream_t s) {
#define SPFFT_SPLIT_LAUNCH_REG(n) SPFFT_SPLIT_LAUNCH_CASE(n, false)
#define SPFFT_SPLIT_LAUNCH_DIR(n) SPFFT_SPLIT_LAUNCH_CASE(n, true)
    SPFFT_SPLIT_DIRECT_LIST(SPFFT_SPLIT_LAUNCH_DIR)
#undef SPFFT_SPLIT_LAUNCH_REG
#undef SPFFT_SPLIT_LAUNCH_DIR
    return hipErrorInvalidValue;
}

}  // namespace spfft
