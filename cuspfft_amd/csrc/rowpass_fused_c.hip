// rowpass_fused_c.hip -- k_rowfused instantiations of group C (see fused_impl.h).
#include "fused_impl.h"

namespace spfft {

hipError_t launch_fused_group_c(const RowPassArgs& a, hipStream_t s, bool prepare_only,
                                const FusedBatch* batch) {
    SPFFT_FUSED_GROUP_C(SPFFT_FUSED_LAUNCH_CASE)
    return hipErrorInvalidValue;
}

}  // namespace spfft
