// rowpass_split.hip -- launcher of k_splitcols, the column pass B of the split row
// transform (split_impl.h); the instantiations are in rowpass_split_{a,b,c}.hip.
#include "spfft_kernels.h"
#include "split_impl.h"

namespace spfft {

hipError_t launch_split_cols_a(const RowPassArgs& a, dim3 grid, size_t lds, hipStream_t s);
hipError_t launch_split_cols_b(const RowPassArgs& a, dim3 grid, size_t lds, hipStream_t s);
hipError_t launch_split_cols_c(const RowPassArgs& a, dim3 grid, size_t lds, hipStream_t s);

bool split_supports_n1(int n1, bool* direct) {
#define SPFFT_SPLIT_REG(n) \
    if (n1 == n) {         \
        *direct = false;   \
        return true;       \
    }
#define SPFFT_SPLIT_DIR(n) \
    if (n1 == n) {         \
        *direct = true;    \
        return true;       \
    }
    SPFFT_SPLIT_REG_LIST_A(SPFFT_SPLIT_REG)
    SPFFT_SPLIT_REG_LIST_B(SPFFT_SPLIT_REG)
    SPFFT_SPLIT_DIRECT_LIST(SPFFT_SPLIT_DIR)
#undef SPFFT_SPLIT_REG
#undef SPFFT_SPLIT_DIR
    return false;
}

size_t split_lds_bytes(int cols) { return (size_t)(256 + (cols + 255) / 256) * sizeof(float2); }

hipError_t launch_split_cols(const RowPassArgs& a, hipStream_t s) {
    if (a.tile_rows <= 0) return hipSuccess;
    bool direct = false;
    if (!split_supports_n1(a.split_n1, &direct)) return hipErrorInvalidValue;
    const dim3 grid((a.split_n2 + kSplitThreads - 1) / kSplitThreads, a.tile_rows);
    const size_t lds = split_lds_bytes(a.C);
    if (direct) return launch_split_cols_c(a, grid, lds, s);
    return a.split_n1 < 32 ? launch_split_cols_a(a, grid, lds, s) : launch_split_cols_b(a, grid, lds, s);
}

}  // namespace spfft
