// csc_noalloc.hip -- spfft_csc_build_noalloc: the COO -> compact-CSC build of
// spfft_csc_build with no device allocation, for callers that cannot allocate
// (code under hipGraph capture, tight serving loops, harnesses that forbid
// hipMalloc).  The caller passes the COO input and the CSC output arrays; the
// build's temporaries (column counts, scan partials, radix-sort digit histograms
// and ping-pong keys / values) live in `scratch`, a file-scope device array of
// kNoallocScratchBytes that each GPU holds until the process ends.  The kernels
// are the ones of spfft_csc_build (csc_build.hip); only where the temporaries
// live differs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/spfft_c.h"
#include "spfft_kernels.h"
#include "spfft_plan_internal.h"

using namespace spfft_internal;

namespace {

constexpr size_t kNoallocScratchBytes = SPFFT_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device
// with that device current (each GPU holds its own copy of the array).
spfft_status scratch_ptr(unsigned char** out) {
    int dev = 0;
    SPFFT_HIP(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(SPFFT_ERR_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        SPFFT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) return fail(SPFFT_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
    return SPFFT_OK;
}

size_t up256(size_t n) { return (n + 255) / 256 * 256; }

// Bytes of scratch a build of (cols, nnz) takes, and the carve-up.
struct Layout {
    size_t counts, scan_tmp, hist, hist_off, keys0, keys1, vals0, vals1, total;
};
Layout layout(int cols, int nnz) {
    Layout l{};
    const size_t hist_n = spfft::csc_sort_hist_elems(nnz);
    const size_t scan_n = std::max(spfft::csc_scan_tmp_elems(cols), spfft::csc_scan_tmp_elems((int)hist_n));
    const bool sort_tmp = spfft::csc_sort_passes(cols) > 1;
    size_t at = 0;
    l.counts = at;   at += up256((size_t)(cols + 1) * sizeof(int));
    l.scan_tmp = at; at += up256(scan_n * sizeof(int2));
    l.hist = at;     at += up256(hist_n * sizeof(int));
    l.hist_off = at; at += up256((hist_n + 1) * sizeof(int));
    const size_t kv = sort_tmp ? up256((size_t)std::max(nnz, 1) * sizeof(int)) : 0;
    l.keys0 = at; at += kv;
    l.keys1 = at; at += kv;
    l.vals0 = at; at += kv;
    l.vals1 = at; at += kv;
    l.total = at;
    return l;
}

}  // namespace

extern "C" {

spfft_status spfft_csc_build_noalloc(int rows, int cols, const int* d_row, const int* d_col, int nnz,
                                     void* d_sparse_cols, int* d_col_ptr, void* d_row_idx, int* d_col_start,
                                     int* d_info, void* stream) {
    if (rows <= 0 || cols <= 0 || nnz < 0) return fail(SPFFT_ERR_INVALID_ARG, "bad matrix shape");
    if (cols > SPFFT_NOALLOC_MAX_COLS || nnz > SPFFT_NOALLOC_MAX_NNZ)
        return fail(SPFFT_ERR_INVALID_ARG,
                    "spfft_csc_build_noalloc: cols <= " + std::to_string(SPFFT_NOALLOC_MAX_COLS) + " and nnz <= " +
                        std::to_string(SPFFT_NOALLOC_MAX_NNZ) + " (the scratch array holds " +
                        std::to_string(kNoallocScratchBytes) + " bytes)");
    if (rows > (1 << 22)) return fail(SPFFT_ERR_UNSUPPORTED, "rows > 2^22 not supported");
    if ((nnz > 0 && (!d_row || !d_col || !d_row_idx)) || !d_sparse_cols || !d_col_ptr || !d_col_start || !d_info)
        return fail(SPFFT_ERR_INVALID_ARG, "NULL pointer");
    const Layout l = layout(cols, nnz);
    if (l.total > kNoallocScratchBytes) return fail(SPFFT_ERR_INVALID_ARG, "scratch array too small for this build");
    unsigned char* base = nullptr;
    spfft_status st = scratch_ptr(&base);
    if (st != SPFFT_OK) return st;
    spfft::CscBuildArgs a{};
    a.R = rows;
    a.C = cols;
    a.nnz = nnz;
    a.packed = rows <= 65536 && cols <= 65536;
    a.row = d_row;
    a.col = d_col;
    a.counts = reinterpret_cast<int*>(base + l.counts);
    a.colstart = d_col_start;
    a.sparse_cols = d_sparse_cols;
    a.col_ptr = d_col_ptr;
    a.row_idx = d_row_idx;
    a.totals = d_info;     // [0] stored nonzeros, [1] active columns
    a.err = d_info + 2;    // [2] nonzero if an index was out of range (cleared by the build)
    a.scan_tmp = reinterpret_cast<int2*>(base + l.scan_tmp);
    a.hist = reinterpret_cast<int*>(base + l.hist);
    a.hist_off = reinterpret_cast<int*>(base + l.hist_off);
    a.keys[0] = reinterpret_cast<int*>(base + l.keys0);
    a.keys[1] = reinterpret_cast<int*>(base + l.keys1);
    a.vals[0] = reinterpret_cast<int*>(base + l.vals0);
    a.vals[1] = reinterpret_cast<int*>(base + l.vals1);
    SPFFT_HIP(spfft::launch_csc_build(a, static_cast<hipStream_t>(stream)), "csc build launch");
    return SPFFT_OK;
}

}  // extern "C"
