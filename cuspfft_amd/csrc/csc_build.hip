// csc_build.hip -- COO -> compact CSC on the GPU, bit-exact with the
// reference's host builder make_compact_csc (include/sparse_fft_internal.cuh:
// 102-120, 200-222) and its uint16 packing (:122-124, :146-151).
//
// Reference semantics restated:
//   sparse_cols = sorted unique columns of the input (n_sp of them)
//   col_ptr[s]  = number of nonzeros in columns sparse_cols[0..s)
//   row_idx     = the rows, grouped by column, keeping the input order inside
//                 a column (a stable counting sort of the COO by column)
//   packed      = rows and columns stored as uint16 iff rows, cols <= 65536
//
// GPU method (integer work only, no floating point):
//   1. k_count_hist0     per-column counts with LDS-free global atomics (order
//                        free, so deterministic), range validation -> err flag;
//                        the same launch makes the first radix pass's histograms
//   2. k_scan_*          exclusive scan of (count, count > 0) over the C
//                        columns -> dense column starts (used by the column
//                        pass), compact slot of every active column ->
//                        sparse_cols and col_ptr
//   3. k_radix_*         stable LSD radix sort of (col, row) by col, 8-bit
//                        digits: per-block digit histograms, one exclusive
//                        scan in digit-major order, and a scatter in which
//                        every wave64 ranks its elements with ballot-based
//                        digit matching, so equal digits keep input order.
//                        The last pass writes row_idx (uint16 when packed).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "spfft_kernels.h"

namespace spfft {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanBlock = kScanThreads * kScanItems;  // 2048 elements per block
constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortIters = 8;
constexpr int kSortBlock = kSortThreads * kSortIters;  // 2048 elements per block

// Zeroes p[0, n) and q[0] (the column counts and the error flag) in one launch.
__global__ void k_zero_counts(int* __restrict__ p, long n, int* __restrict__ q) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
    if (i == 0) q[0] = 0;
}

// Block-wide exclusive scan of one int per thread; returns the block total.
__device__ int block_exclusive_scan(int x, int* s_warp, int& total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int incl = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) s_warp[wid] = incl;
    __syncthreads();
    int woff = 0, tot = 0;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; w++) {
        const int t = s_warp[w];
        if (w < wid) woff += t;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return woff + incl - x;
}

// Mode 0: plain exclusive scan of `in`.  Mode 1 (columns): also scans the
// active flag (in[c] > 0) in the second lane of an int2.
template <int MODE>
__global__ void k_scan_reduce(const int* __restrict__ in, int n, int2* __restrict__ bsum) {
    __shared__ int s_w[kScanThreads / 64];
    const int base = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
    int sx = 0, sy = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int i = base + k;
        const int v = i < n ? in[i] : 0;
        sx += v;
        sy += (MODE == 1 && v > 0) ? 1 : 0;
    }
    int tx, ty;
    block_exclusive_scan(sx, s_w, tx);
    block_exclusive_scan(sy, s_w, ty);
    if (threadIdx.x == 0) bsum[blockIdx.x] = make_int2(tx, ty);
}

// Single block: exclusive scan of the block sums in place; writes totals.
__global__ void k_scan_top(int2* __restrict__ bsum, int nb, int* __restrict__ totals) {
    __shared__ int s_w[kScanThreads / 64];
    __shared__ int s_carry[2];
    if (threadIdx.x == 0) { s_carry[0] = 0; s_carry[1] = 0; }
    __syncthreads();
    for (int base = 0; base < nb; base += kScanThreads) {
        const int i = base + threadIdx.x;
        const int2 v = i < nb ? bsum[i] : make_int2(0, 0);
        int tx, ty;
        const int ex = block_exclusive_scan(v.x, s_w, tx);
        const int ey = block_exclusive_scan(v.y, s_w, ty);
        const int cx = s_carry[0], cy = s_carry[1];
        if (i < nb) bsum[i] = make_int2(cx + ex, cy + ey);
        __syncthreads();
        if (threadIdx.x == 0) { s_carry[0] = cx + tx; s_carry[1] = cy + ty; }
        __syncthreads();
    }
    if (threadIdx.x == 0 && totals) { totals[0] = s_carry[0]; totals[1] = s_carry[1]; }
}

// Applies the scan.  Mode 0: out[i] = exclusive prefix.  Mode 1: out = dense
// column starts (C+1 entries); for every active column c with slot s:
// sparse_cols[s] = c, col_ptr[s] = start[c]; the block holding index n also
// writes out[n] and col_ptr[n_sp].
template <int MODE, bool PACKED>
__global__ void k_scan_apply(const int* __restrict__ in, int n, const int2* __restrict__ bsum,
                             int* __restrict__ out, void* __restrict__ sparse_cols,
                             int* __restrict__ col_ptr) {
    __shared__ int s_w[kScanThreads / 64];
    const int base = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
    int v[kScanItems];
    int sx = 0, sy = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int i = base + k;
        v[k] = i < n ? in[i] : 0;
        sx += v[k];
        sy += (MODE == 1 && v[k] > 0) ? 1 : 0;
    }
    int tx, ty;
    int ex = block_exclusive_scan(sx, s_w, tx);
    int ey = block_exclusive_scan(sy, s_w, ty);
    const int2 off = bsum[blockIdx.x];
    ex += off.x;
    ey += off.y;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int i = base + k;
        if (i < n) {
            out[i] = ex;
            if (MODE == 1 && v[k] > 0) {
                if (PACKED) ((uint16_t*)sparse_cols)[ey] = (uint16_t)i;
                else ((int*)sparse_cols)[ey] = i;
                col_ptr[ey] = ex;
                ey += 1;
            }
        } else if (i == n) {
            out[n] = ex;
            if (MODE == 1) col_ptr[ey] = ex;
        }
        ex += v[k];
    }
}

// Single-block scan for n + 1 <= kOneBlock elements: k_scan_reduce, k_scan_top
// and k_scan_apply in one launch (the same outputs, totals included).  Small
// scans are bound by launch latency; larger ones run faster on many blocks.
constexpr int kOneThreads = 1024;
constexpr int kOneItems = 4;
constexpr int kOneBlock = kOneThreads * kOneItems;

template <int MODE, bool PACKED>
__device__ __forceinline__ void scan_single_body(const int* __restrict__ in, int n, int* __restrict__ totals,
                                                 int* __restrict__ out, void* __restrict__ sparse_cols,
                                                 int* __restrict__ col_ptr) {
    __shared__ int s_w[kOneThreads / 64];
    const int base = threadIdx.x * kOneItems;
    int v[kOneItems];
    int sx = 0, sy = 0;
#pragma unroll
    for (int k = 0; k < kOneItems; k++) {
        const int i = base + k;
        v[k] = i < n ? in[i] : 0;
        sx += v[k];
        sy += (MODE == 1 && v[k] > 0) ? 1 : 0;
    }
    int tx, ty;
    int ex = block_exclusive_scan(sx, s_w, tx);
    int ey = block_exclusive_scan(sy, s_w, ty);
#pragma unroll
    for (int k = 0; k < kOneItems; k++) {
        const int i = base + k;
        if (i < n) {
            out[i] = ex;
            if (MODE == 1 && v[k] > 0) {
                if (PACKED) ((uint16_t*)sparse_cols)[ey] = (uint16_t)i;
                else ((int*)sparse_cols)[ey] = i;
                col_ptr[ey] = ex;
                ey += 1;
            }
        } else if (i == n) {
            out[n] = ex;
            if (MODE == 1) col_ptr[ey] = ex;
        }
        ex += v[k];
    }
    if (threadIdx.x == 0 && totals) {
        totals[0] = tx;
        totals[1] = ty;
    }
}

template <int MODE, bool PACKED>
__global__ __launch_bounds__(kOneThreads) void k_scan_single(const int* __restrict__ in, int n,
                                                             int* __restrict__ totals, int* __restrict__ out,
                                                             void* __restrict__ sparse_cols,
                                                             int* __restrict__ col_ptr) {
    scan_single_body<MODE, PACKED>(in, n, totals, out, sparse_cols, col_ptr);
}

// The column scan (block 0, as k_scan_single<1, PACKED>) and the scan of the first radix
// pass's histograms (block 1, as k_scan_single<0, false>) in one launch, when both fit
// one block.
template <bool PACKED>
__global__ __launch_bounds__(kOneThreads) void k_scan_single_pair(const int* __restrict__ counts, int C,
                                                                  int* __restrict__ totals, int* __restrict__ colstart,
                                                                  void* __restrict__ sparse_cols,
                                                                  int* __restrict__ col_ptr,
                                                                  const int* __restrict__ hist, int nh,
                                                                  int* __restrict__ hist_off) {
    if (blockIdx.x == 0) scan_single_body<1, PACKED>(counts, C, totals, colstart, sparse_cols, col_ptr);
    else scan_single_body<0, false>(hist, nh, nullptr, hist_off, nullptr, nullptr);
}

__device__ __forceinline__ int digit_of(int key, int shift) { return (int)(((uint32_t)key >> shift) & 255u); }

__device__ __forceinline__ void radix_hist_block(const int* __restrict__ keys, int nnz, int shift,
                                                 int* __restrict__ hist, int nblk, int b) {
    __shared__ int s_h[256];
    s_h[threadIdx.x] = 0;
    __syncthreads();
    const int base = b * kSortBlock;
    for (int k = 0; k < kSortIters; k++) {
        const int i = base + k * kSortThreads + threadIdx.x;
        if (i < nnz) atomicAdd(&s_h[digit_of(keys[i], shift)], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nblk + b] = s_h[threadIdx.x];
}

__global__ void k_radix_hist(const int* __restrict__ keys, int nnz, int shift, int* __restrict__ hist,
                             int nblk) {
    radix_hist_block(keys, nnz, shift, hist, nblk, blockIdx.x);
}

// The column counts (blocks [0, nb_count)) and the first radix pass's
// digit histograms (k_radix_hist of the input columns, blocks from nb_count on) in one
// launch: both only read the COO, so one launch does the work of two (small inputs are
// bound by launch latency).
__global__ __launch_bounds__(kSortThreads) void k_count_hist0(const int* __restrict__ row,
                                                              const int* __restrict__ col, int nnz, int R, int C,
                                                              int* __restrict__ counts, int* __restrict__ err,
                                                              int nb_count, int* __restrict__ hist, int nblk) {
    if ((int)blockIdx.x >= nb_count) {
        radix_hist_block(col, nnz, 0, hist, nblk, blockIdx.x - nb_count);
        return;
    }
    const int i = blockIdx.x * kSortThreads + threadIdx.x;
    if (i >= nnz) return;
    const int r = row[i], c = col[i];
    if (r < 0 || r >= R || c < 0 || c >= C) {
        atomicOr(err, 1);
        return;
    }
    atomicAdd(&counts[c], 1);
}

// Stable scatter.  Wave w of block b owns elements
// [b*kSortBlock + w*64*kSortIters, +64*kSortIters) in order; inside a wave the
// rank of a lane among equal digits is the popcount of the matching lanes
// below it, found with 8 ballots.  The final pass writes rows outside [0, R)
// as 0, so that the column pass never indexes its tables with an invalid row
// (such input is reported through the err flag; its output is undefined).
template <bool FINAL, bool PACKED>
__global__ void k_radix_scatter(const int* __restrict__ keys_in, const int* __restrict__ vals_in,
                                int nnz, int R, int shift, const int* __restrict__ offs, int nblk,
                                int* __restrict__ keys_out, int* __restrict__ vals_out,
                                void* __restrict__ final_out, const int* __restrict__ col_rbase = nullptr,
                                int C = 0) {
    __shared__ int s_cnt[kSortWaves][256];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kSortWaves * 256; i += kSortThreads) (&s_cnt[0][0])[i] = 0;
    __syncthreads();
    const int wbase = blockIdx.x * kSortBlock + wid * 64 * kSortIters;
    for (int k = 0; k < kSortIters; k++) {
        const int i = wbase + k * 64 + lane;
        if (i < nnz) atomicAdd(&s_cnt[wid][digit_of(keys_in[i], shift)], 1);
    }
    __syncthreads();
    {
        const int d = threadIdx.x;  // 256 threads = 256 digits
        int run = offs[d * nblk + blockIdx.x];
        for (int w = 0; w < kSortWaves; w++) {
            const int t = s_cnt[w][d];
            s_cnt[w][d] = run;
            run += t;
        }
    }
    __syncthreads();
    volatile int* cnt = s_cnt[wid];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int k = 0; k < kSortIters; k++) {
        const int i = wbase + k * 64 + lane;
        const bool valid = i < nnz;
        const int key = valid ? keys_in[i] : 0;
        const int d = digit_of(key, shift);
        uint64_t peer = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1;
            const uint64_t m = __ballot(bit);
            peer &= bit ? m : ~m;
        }
        const int base = valid ? cnt[d] : 0;
        const int rank = __popcll(peer & lt);
        const int leader = __ffsll((unsigned long long)peer) - 1;
        __builtin_amdgcn_wave_barrier();
        if (valid) {
            if (lane == leader) cnt[d] = base + __popcll(peer);
            const int pos = base + rank;
            int val = vals_in[i];
            if (FINAL) {
                // the batched build: rows of the key's (column's) own matrix
                if (col_rbase) val = (uint32_t)key < (uint32_t)C ? val - col_rbase[key] : 0;
                if ((uint32_t)val >= (uint32_t)R) val = 0;
                if (PACKED) ((uint16_t*)final_out)[pos] = (uint16_t)val;
                else ((int*)final_out)[pos] = val;
            } else {
                keys_out[pos] = key;
                vals_out[pos] = val;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

template <int MODE, bool PACKED>
hipError_t run_scan(const int* in, int n, int2* bsum, int* totals, int* out, void* sparse_cols,
                    int* col_ptr, hipStream_t s) {
    if ((long)n + 1 <= kOneBlock) {
        hipLaunchKernelGGL((k_scan_single<MODE, PACKED>), dim3(1), dim3(kOneThreads), 0, s, in, n, totals, out,
                           sparse_cols, col_ptr);
        return hipGetLastError();
    }
    // one extra slot so that index n (the total) is written by some block
    const int nb = cdiv((long)n + 1, kScanBlock);
    hipLaunchKernelGGL((k_scan_reduce<MODE>), dim3(nb), dim3(kScanThreads), 0, s, in, n, bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, s, bsum, nb, totals);
    hipLaunchKernelGGL((k_scan_apply<MODE, PACKED>), dim3(nb), dim3(kScanThreads), 0, s, in, n, bsum,
                       out, sparse_cols, col_ptr);
    return hipGetLastError();
}

}  // namespace

size_t csc_scan_tmp_elems(int n) { return (size_t)cdiv((long)n + 1, kScanBlock) + 1; }
size_t csc_sort_hist_elems(int nnz) { return (size_t)256 * (size_t)(cdiv((long)(nnz > 0 ? nnz : 1), kSortBlock)); }

int csc_sort_passes(int cols) {
    uint32_t m = cols > 1 ? (uint32_t)(cols - 1) : 0u;
    int p = 1;
    while (m > 255u) { m >>= 8; p++; }
    return p;
}

namespace {

// The batched build's input: element e of the concatenated COO (position coo_off[0] + e
// of the caller's arrays) belongs to the last matrix i with coo_off[i] <= coo_off[0] + e; its (r, c) becomes (r + r_off, c + c_off) of the
// block-diagonal matrix of all members (mat[i] = {R, C, r_off, c_off}).  A pair outside
// its own matrix flags err[i] and gets column -1, so that the build skips it; so do the
// pairs of matrices that are not members (mat[i].x < 0), with no flag.
__global__ void k_batch_coo(const int* __restrict__ row, const int* __restrict__ col,
                            const long long* __restrict__ coo_off, int count, const int4* __restrict__ mat,
                            int* __restrict__ mrow, int* __restrict__ mcol, int* __restrict__ err, long long nnz) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    const long long ea = coo_off[0] + e;  // position in the caller's arrays
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (coo_off[mid] <= ea) lo = mid;
        else hi = mid - 1;
    }
    const int4 m = mat[lo];
    if (m.x < 0) {  // not a member of the batched build: skipped, no error
        mrow[e] = 0;
        mcol[e] = -1;
        return;
    }
    const int r = row[ea], c = col[ea];
    if (r < 0 || r >= m.x || c < 0 || c >= m.y) {
        err[lo] = 1;
        mrow[e] = 0;
        mcol[e] = -1;
        return;
    }
    mrow[e] = r + m.z;
    mcol[e] = c + m.w;
}

__global__ void k_fill_int(int* __restrict__ p, long n, int v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

hipError_t launch_fill_int(int* p, long n, int v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_int, dim3(cdiv(n, 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_batch_coo(const int* row, const int* col, const long long* coo_off, int count, const int4* mat,
                            int* mrow, int* mcol, int* err, long long nnz, hipStream_t s) {
    if (nnz <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_coo, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, s, row, col, coo_off, count,
                       mat, mrow, mcol, err, nnz);
    return hipGetLastError();
}

hipError_t launch_csc_build(const CscBuildArgs& a, hipStream_t s) {
    const int R = a.R, C = a.C, nnz = a.nnz;
    // 1. counts + validation
    // (with the first radix pass's histograms of the input columns)
    hipLaunchKernelGGL(k_zero_counts, dim3(cdiv((long)C + 1, 256)), dim3(256), 0, s, a.counts, (long)C + 1, a.err);
    const int nblk = cdiv(nnz, kSortBlock);
    if (nnz > 0) {
        const int nb_count = cdiv(nnz, kSortThreads);
        hipLaunchKernelGGL(k_count_hist0, dim3(nb_count + nblk), dim3(kSortThreads), 0, s, a.row, a.col, nnz, R, C,
                           a.counts, a.err, nb_count, a.hist, nblk);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // 2. column scan + compaction (with the first radix pass's histogram scan when both
    // fit one block)
    const bool pair = nnz > 0 && (long)C + 1 <= kOneBlock && 256L * nblk + 1 <= kOneBlock;
    if (pair) {
        if (a.packed)
            hipLaunchKernelGGL((k_scan_single_pair<true>), dim3(2), dim3(kOneThreads), 0, s, a.counts, C, a.totals,
                               a.colstart, a.sparse_cols, a.col_ptr, a.hist, 256 * nblk, a.hist_off);
        else
            hipLaunchKernelGGL((k_scan_single_pair<false>), dim3(2), dim3(kOneThreads), 0, s, a.counts, C, a.totals,
                               a.colstart, a.sparse_cols, a.col_ptr, a.hist, 256 * nblk, a.hist_off);
        e = hipGetLastError();
    } else {
        e = a.packed ? run_scan<1, true>(a.counts, C, a.scan_tmp, a.totals, a.colstart, a.sparse_cols,
                                         a.col_ptr, s)
                     : run_scan<1, false>(a.counts, C, a.scan_tmp, a.totals, a.colstart, a.sparse_cols,
                                          a.col_ptr, s);
    }
    if (e != hipSuccess || nnz == 0) return e;
    // 3. stable radix sort of rows by column
    const int passes = csc_sort_passes(C);
    const int* kin = a.col;
    const int* vin = a.row;
    int* kout = a.keys[0];
    int* vout = a.vals[0];
    for (int p = 0; p < passes; p++) {
        const int shift = 8 * p;
        const bool fin = (p == passes - 1);
        if (p > 0)  // pass 0's histograms came with the column counts
            hipLaunchKernelGGL(k_radix_hist, dim3(nblk), dim3(kSortThreads), 0, s, kin, nnz, shift, a.hist, nblk);
        if (p > 0 || !pair) {  // pass 0's scan ran with the column scan when both fit one block
            e = run_scan<0, false>(a.hist, 256 * nblk, a.scan_tmp, nullptr, a.hist_off, nullptr, nullptr, s);
            if (e != hipSuccess) return e;
        }
        if (fin) {
            if (a.col_rbase)  // batched: uint16 rows of each column's own matrix (all of R <= 65536)
                hipLaunchKernelGGL((k_radix_scatter<true, true>), dim3(nblk), dim3(kSortThreads), 0, s, kin, vin,
                                   nnz, 65536, shift, a.hist_off, nblk, nullptr, nullptr, a.row_idx, a.col_rbase, C);
            else if (a.packed)
                hipLaunchKernelGGL((k_radix_scatter<true, true>), dim3(nblk), dim3(kSortThreads), 0, s, kin, vin,
                                   nnz, R, shift, a.hist_off, nblk, nullptr, nullptr, a.row_idx);
            else
                hipLaunchKernelGGL((k_radix_scatter<true, false>), dim3(nblk), dim3(kSortThreads), 0, s, kin, vin,
                                   nnz, R, shift, a.hist_off, nblk, nullptr, nullptr, a.row_idx);
        } else {
            hipLaunchKernelGGL((k_radix_scatter<false, false>), dim3(nblk), dim3(kSortThreads), 0, s, kin, vin,
                               nnz, R, shift, a.hist_off, nblk, kout, vout, nullptr);
            kin = kout;
            vin = vout;
            kout = (kout == a.keys[0]) ? a.keys[1] : a.keys[0];
            vout = (vout == a.vals[0]) ? a.vals[1] : a.vals[0];
        }
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace spfft
