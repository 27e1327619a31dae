// spfft_plan_internal.h -- the plan object and helpers shared by the C ABI
// translation units (spfft_plan.cpp, spfft_batch.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <memory>
#include <vector>

#include "../../include/spfft_c.h"
#include "spfft_types.h"

namespace spfft_internal {

spfft_status fail(spfft_status st, const std::string& msg);
spfft_status hip_fail(hipError_t e, const char* what);
spfft_status read_opts(const spfft_opts* in, spfft_opts* out);
spfft_status check_info(const spfft_plan_info* info);
spfft_status check_result(spfft_result* res);

#define SPFFT_HIP(call, what)                                                      \
    do {                                                                           \
        hipError_t e__ = (call);                                                   \
        if (e__ != hipSuccess) return spfft_internal::hip_fail(e__, what);         \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            reset();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n) {
        reset();
        if (n == 0) n = 16;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) { p = nullptr; return e; }
        bytes = n;
        return hipSuccess;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

template <class T>
hipError_t upload(DevBuf& b, const std::vector<T>& v) {
    hipError_t e = b.alloc(v.size() * sizeof(T));
    if (e != hipSuccess) return e;
    if (v.empty()) return hipSuccess;
    return hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}


}  // namespace spfft_internal

using spfft_internal::DevBuf;

struct spfft_plan_s {
    int R = 0, C = 0, V = 0, P = 0, M = 0, M1 = 0, M2 = 0;
    spfft_opts opts{};
    int t_rows = 0;  // four-step kernels: rows of the tile buffer T per sub-tile
    bool packed = false;
    bool colpass_full = false;
    spfft::RowPassArgs base{};  // kernel argument template (pointers to the tables filled in)
    DevBuf h, wr_hi, wr_lo, wr_full, tw2, twM, wm2, bhat, mfma_dft, mfma_frag, X0, T;
    DevBuf rader_k, colperm;  // Rader and split plans: g^-m mod p, X0 column order
    // split plans (C = N1 N2): the plan of the width-N2 row pass over the N1 runs of each
    // X0 row (its tables and shape; its own workspace is unused), W_C and DFT_N1 tables
    std::unique_ptr<spfft_plan_s> split_sub;
    spfft_opts split_opts{};  // the options of that plan
    int split_len = 0;        // its row-pass length (Bluestein M, N2 when direct, N2 - 1 by Rader)
    DevBuf split_w, split_dft;
    // or, for an odd prime N2 with a k_raderruns shape, Rader's algorithm over the runs:
    // its kernel arguments (tables: rader_k, rr_bhat, rr_tw2) in place of split_sub
    bool split_rr = false;
    spfft::RowPassArgs split_rr_args{};
    DevBuf rr_bhat, rr_tw2;
    // CSC workspace
    int nnz_cap = -1;
    int last_nnz = 0;
    bool have_csc = false;
    DevBuf counts, colstart, sparse_cols, col_ptr, totals, err, scan_tmp, hist, hist_off;
    DevBuf col_sched, col_sched_off;  // k_colpass column schedule
    int colpass_rows = 1;             // k_colpass rows per lane of the last CSC build
    DevBuf keys0, keys1, vals0, vals1, row_idx;
    size_t table_bytes = 0;
    // optional per-kernel event timing (spfft_profile_*)
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    // event groups recorded since the last read: kind 0 = CSC build (2 events),
    // kind 1 = row-pass tile (4 events)
    std::vector<std::pair<int, size_t>> ev_groups;
    // streamed output (spfft_execute_to_host): two device chunks of 2 * u_tile rows
    // (a tile's rows and their mirror rows), a copy stream, and per chunk an event
    // "written" (compute stream) and "copied" (copy stream)
    DevBuf chunk[2];
    hipStream_t copy_stream = nullptr;
    hipEvent_t chunk_ready[2] = {nullptr, nullptr}, chunk_free[2] = {nullptr, nullptr};
    hipEvent_t copies_done = nullptr;
    ~spfft_plan_s() {
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
        for (hipEvent_t e : {chunk_ready[0], chunk_ready[1], chunk_free[0], chunk_free[1], copies_done})
            if (e) (void)hipEventDestroy(e);
        if (copy_stream) (void)hipStreamDestroy(copy_stream);
    }
    hipEvent_t* take_events(int n) {
        while (ev_pool.size() < ev_used + n) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            ev_pool.push_back(e);
        }
        hipEvent_t* r = &ev_pool[ev_used];
        ev_used += n;
        return r;
    }
};


namespace spfft_internal {

spfft_status ensure_csc_capacity(spfft_plan_s& p, int nnz);
spfft_status do_csc_build(spfft_plan_s& p, const int* d_row, const int* d_col, int nnz, hipStream_t s);
// T / X0: workspace to use instead of the plan's own (NULL: the plan's own).
spfft_status do_rowpass(spfft_plan_s& p, void* d_out, int64_t pitch, hipStream_t s, float2* T,
                        float2* X0);
// The row pass with the output streamed to host memory h_out (pitch elements per row).
spfft_status do_rowpass_to_host(spfft_plan_s& p, void* h_out, int64_t pitch, hipStream_t s);
// The column pass pairs rows u and R/2 - u (SPFFT_PAIR_AUTO, R even, one tile of rows).
bool pair_eligible(const spfft_plan_s& p);
// Plan without its own row-pass workspace (T, X0) when own_workspace is false.
spfft_status plan_create_ex(int rows, int cols, const spfft_opts* opts, bool own_workspace,
                            spfft_plan_s** out);
// Bytes of the row-pass workspace: T (tile x M) and X0 (tile x ldx), complex64.
void workspace_bytes(const spfft_plan_s& p, size_t* t_bytes, size_t* x0_bytes);
size_t device_lds_limit();
void batch_descriptor(const spfft_plan_s& p, int64_t nnz, spfft::RowPassArgs& a);
// the automatic column-pass kernel (rows per lane, or spfft::kColpassMfma) for nnz nonzeros
int auto_colpass_rows(const spfft_plan_s& p, int64_t nnz);
// automatic W_R table choice: the whole table in LDS only up to this many rows
constexpr int kAutoFullTableRows = 2048;
// automatic column pass: the matrix-core kernel from this average column degree
constexpr int kAutoRows16Degree = 12;  // average column degree for 16 rows per lane (auto_colpass_rows)

}  // namespace spfft_internal
