// rowpass_fused_f.hip -- k_rowfused instantiations of group F (see fused_impl.h).
#include "fused_impl.h"

namespace spfft {

hipError_t launch_fused_group_f(const RowPassArgs& a, hipStream_t s, bool prepare_only,
                                const FusedBatch* batch) {
    SPFFT_FUSED_GROUP_F(SPFFT_FUSED_LAUNCH_CASE)
    return hipErrorInvalidValue;
}

}  // namespace spfft
