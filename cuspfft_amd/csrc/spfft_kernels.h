// spfft_kernels.h -- host-visible launchers of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "spfft_types.h"

namespace spfft {

// ---- COO -> compact CSC (csc_build.hip) ----------------------------------
struct CscBuildArgs {
    int R, C, nnz;
    bool packed;
    const int* row;          // device COO, nnz entries each
    const int* col;
    int* counts;             // C+1
    int* colstart;           // C+1 dense column starts (output)
    void* sparse_cols;       // n_sp uint16 (packed) or int32 (output)
    int* col_ptr;            // n_sp+1 (output)
    void* row_idx;           // nnz uint16 (packed) or int32 (output)
    int* totals;             // [0] = valid nnz, [1] = n_sp (output)
    int* err;                // nonzero if any index is out of range (output)
    int2* scan_tmp;          // csc_scan_tmp_elems(max(C, 256*blocks))
    int* hist;               // csc_sort_hist_elems(nnz)
    int* hist_off;           // csc_sort_hist_elems(nnz) + 1
    int* keys[2];            // nnz each (only used when more than one pass)
    int* vals[2];
    const int* col_rbase;    // batched build (NULL otherwise): per column, the first row of its
                             // member matrix; row_idx then holds uint16 rows of that matrix
};

size_t csc_scan_tmp_elems(int n);
size_t csc_sort_hist_elems(int nnz);
int csc_sort_passes(int cols);
hipError_t launch_csc_build(const CscBuildArgs& a, hipStream_t s);
hipError_t launch_fill_int(int* p, long n, int v, hipStream_t s);
// the batched build's block-diagonal COO (mat[i] = {R, C, first row, first column})
hipError_t launch_batch_coo(const int* row, const int* col, const long long* coo_off, int count, const int4* mat,
                            int* mrow, int* mcol, int* err, long long nnz, hipStream_t s);

// ---- row pass (rowpass.hip) ----------------------------------------------
bool rowpass_supports_m1(int m1);        // k_fft1 / k_ifft1 instantiated for this M1
bool mfma_supports_m1(int m1);           // precision-form (BF16x3 MFMA) sub-FFT length
bool conv_fast_supported(int r1, int r2); // k_conv_fast<r1, r2> instantiated (r2 = 1: one stage)
int conv_max_butterflies(int r);         // butterflies per thread per k_conv_fast stage
// colpass_rows value of the matrix-core column pass (k_colpass_mx: 8 columns x 1024
// rows per workgroup, FP32 MFMA); the other values are k_colpass's rows per lane
constexpr int kColpassMfma = 32;
size_t colpass_lds_bytes(const RowPassArgs& a);   // needs a.colpass_rows and a.wr_full set
// k_colsched: the per-tile column schedule of k_colpass (once per CSC build);
// sizes in ints of sched (returned) and of off (*off_elems)
size_t colsched_elems(int cols, size_t* off_elems);
// (colperm: NULL, or the column of each X0 position, as RowPassArgs::colperm)
hipError_t launch_colsched(const int* colstart, const int* colperm, int cols, int rows_per_lane, int* sched,
                           int* off, hipStream_t s);
hipError_t colpass_prepare(size_t lds_limit);      // once per process, before launches
// ---- batched form: one launch per kernel instantiation for many matrices ----------
// descs: device array of the matrices' RowPassArgs; work: per workgroup, the matrix
// (descriptor index) and that matrix's block indices.
hipError_t launch_colsched_batch(const RowPassArgs* descs, const int2* work, int n, hipStream_t s);
// rep: one of the group's descriptors (its colpass_rows, pair_half and wr_full pick the
// kernel); lds: the group's largest colpass_lds_bytes
hipError_t launch_colpass_batch(const RowPassArgs& rep, bool packed, const RowPassArgs* descs, const int4* work,
                                int n, size_t lds, hipStream_t s);
// the (column tile, row block) grid of k_colpass for one tile of a matrix
void colpass_grid(const RowPassArgs& a, int* tiles_x, int* tiles_y);
int colpass_cols_per_tile(int rows_per_lane);
// row kernels (k_rowfused / k_rowrader / k_rowrader_mfma) of rep's shape; work[i] = (descriptor,
// row group); each matrix's tile_rows, u_base, out, ... come from its descriptor
hipError_t launch_fused_batch(const RowPassArgs& rep, const RowPassArgs* descs, const int2* work, int n,
                              hipStream_t s);
hipError_t launch_rader_batch(const RowPassArgs& rep, const RowPassArgs* descs, const int2* work, int n,
                              hipStream_t s);
size_t conv_lds_bytes(const RowPassArgs& a);
// Fused row kernel (k_rowfused) shapes: M1, LDS batches, rows per workgroup,
// threads, stage radices r[3] (1 = no stage).  Returns false past the last one.
struct FusedShape {
    int m1, nb, rpw, nt, r[3];
    size_t lds_bytes;
};
bool fused_shape(int i, FusedShape* out);
// prepare_only: raise the kernel's dynamic LDS limit (once per process, before launches)
hipError_t launch_fused(const RowPassArgs& a, hipStream_t s, bool prepare_only);
// Rader row kernel (k_rowrader, rader_impl.h) shapes for widths C = 2p (ns = 2) or C = p
// (ns = 1), p an odd prime: M1, threads, stage radices r[3] with p - 1 = M1 r[0] r[1] r[2].
struct RaderShape {
    int m1, nt, r[3];
    size_t lds_bytes;
    int mfma;  // 1: the precision form (k_rowrader_mfma, ns = 2 only), 0: FP32 codelets (k_rowrader)
    int ns;    // convolutions per row: 2 for C = 2p, 1 for C = p
};
bool rader_shape(int i, RaderShape* out);
hipError_t launch_rader(const RowPassArgs& a, hipStream_t s, bool prepare_only);
// ev: NULL, or 5 events recorded before k_colpass and after each of the 4 kernels
// (with the fused row kernel: after k_colpass, twice, after k_rowfused, twice).
hipError_t launch_rowpass_tile(const RowPassArgs& a, bool packed, hipStream_t s, hipEvent_t* ev);
// The column pass of one tile alone (k_colpass or k_colpass_mx as a.colpass_rows picks).
hipError_t launch_colpass_tile(const RowPassArgs& a, bool packed, hipStream_t s);
// Split row transform (rowpass_split*.hip): pass B, k_splitcols<a.split_n1> over the
// tile buffer a.T (C entries per row) into the output rows of the tile.
bool split_supports_n1(int n1, bool* direct);
// Pass A of the split transform for an odd prime run length p (rowpass_raderruns.hip):
// Rader's convolution of length L = p - 1 for a.fused_rpw runs per workgroup.  shape:
// M1, runs per workgroup, threads and stage radices r[3] for L, or false.
bool raderruns_shape(int L, int* m1, int* ng, int* nt, int* r);
size_t raderruns_lds_bytes(int L);  // of that shape's kernel (0: none)
hipError_t launch_raderruns(const RowPassArgs& a, hipStream_t s, bool prepare_only);
size_t split_lds_bytes(int cols);
hipError_t launch_split_cols(const RowPassArgs& a, hipStream_t s);

}  // namespace spfft
