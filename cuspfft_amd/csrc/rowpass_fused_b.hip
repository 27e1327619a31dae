// rowpass_fused_b.hip -- k_rowfused instantiations of group B (see fused_impl.h).
#include "fused_impl.h"

namespace spfft {

hipError_t launch_fused_group_b(const RowPassArgs& a, hipStream_t s, bool prepare_only,
                                const FusedBatch* batch) {
    SPFFT_FUSED_GROUP_B(SPFFT_FUSED_LAUNCH_CASE)
    return hipErrorInvalidValue;
}

bool fused_shape(int i, FusedShape* out) {
    static const FusedShape kShapes[] = {
#define SPFFT_SHAPE(m1, nb, rpw, nt, r1, r2, r3) \
    {m1, nb, rpw, nt, {r1, r2, r3}, fused_lds_bytes_of(m1, nb, rpw, r1, r2, r3)},
        SPFFT_FUSED_LIST(SPFFT_SHAPE)
#undef SPFFT_SHAPE
    };
    if (i < 0 || i >= (int)(sizeof(kShapes) / sizeof(kShapes[0]))) return false;
    *out = kShapes[i];
    return true;
}

static hipError_t launch_fused_any(const RowPassArgs& a, hipStream_t s, bool prepare_only, const FusedBatch* batch) {
    hipError_t (*const groups[])(const RowPassArgs&, hipStream_t, bool, const FusedBatch*) = {
        launch_fused_group_a, launch_fused_group_b, launch_fused_group_c, launch_fused_group_d,
        launch_fused_group_e, launch_fused_group_f, launch_fused_group_g, launch_fused_group_h};
    hipError_t e = hipErrorInvalidValue;
    for (auto* g : groups)
        if ((e = g(a, s, prepare_only, batch)) != hipErrorInvalidValue) break;
    return e;
}

hipError_t launch_fused(const RowPassArgs& a, hipStream_t s, bool prepare_only) {
    return launch_fused_any(a, s, prepare_only, nullptr);
}

hipError_t launch_fused_batch(const RowPassArgs& rep, const RowPassArgs* descs, const int2* work, int n,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const FusedBatch b{descs, work, n};
    return launch_fused_any(rep, s, false, &b);
}

}  // namespace spfft
