// rowpass_fused_e.hip -- k_rowfused instantiations of group E (see fused_impl.h).
#include "fused_impl.h"

namespace spfft {

hipError_t launch_fused_group_e(const RowPassArgs& a, hipStream_t s, bool prepare_only,
                                const FusedBatch* batch) {
    SPFFT_FUSED_GROUP_E(SPFFT_FUSED_LAUNCH_CASE)
    return hipErrorInvalidValue;
}

}  // namespace spfft
