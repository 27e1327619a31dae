// sparse_fft_shim.cpp -- the reference-shaped C++ entry points (sparse_fft.h,
// mtx_reader.h) over the C ABI.  Errors become std::runtime_error, as the
// reference's CUDA_CHECK / CUFFT_CHECK macros throw
// (reference include/sparse_fft_internal.cuh:16-38).
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "../../include/sparse_fft.h"
#include "../../include/spfft_c.h"

namespace {

enum class Variant { MixedRadix, MixedRadixGraph, Smooth7 };

SparseFFTResult run(const COOMatrix& coo, int u_tile, Variant v) {
    if (coo.nnz < 0 || (size_t)coo.nnz > coo.row_idx.size() || (size_t)coo.nnz > coo.col_idx.size())
        throw std::runtime_error("COOMatrix nnz does not match its index arrays");
    spfft_result r;
    r.struct_size = sizeof(r);
    const int* hr = coo.row_idx.data();
    const int* hc = coo.col_idx.data();
    spfft_status st;
    switch (v) {
        case Variant::Smooth7:
            st = spfft_csc_bluestein_cufft_smooth(coo.rows, coo.cols, coo.nnz, hr, hc, u_tile, &r);
            break;
        case Variant::MixedRadixGraph:
            st = spfft_csc_bluestein_mixed_radix_graph(coo.rows, coo.cols, coo.nnz, hr, hc, u_tile, &r);
            break;
        default:
            st = spfft_csc_bluestein_mixed_radix(coo.rows, coo.cols, coo.nnz, hr, hc, u_tile, &r);
    }
    if (st != SPFFT_OK) throw std::runtime_error(spfft_last_error());
    SparseFFTResult out;
    out.d_output = static_cast<hipFloatComplex*>(r.d_output);
    out.output_cols = r.output_cols;
    out.ms = r.ms;
    out.mem_bytes = r.mem_bytes;
    return out;
}

}  // namespace

SparseFFTResult sparse_fft_csc_bluestein_mixed_radix(const COOMatrix& coo, int u_tile) {
    return run(coo, u_tile, Variant::MixedRadix);
}

SparseFFTResult sparse_fft_csc_bluestein_mixed_radix_graph(const COOMatrix& coo, int u_tile) {
    return run(coo, u_tile, Variant::MixedRadixGraph);
}

SparseFFTResult sparse_fft_csc_bluestein_cufft_smooth(const COOMatrix& coo, int u_tile) {
    return run(coo, u_tile, Variant::Smooth7);
}

SparseFFTResult sparse_fft_csc_stockham_streaming(const COOMatrix& coo, int u_tile, int fft_backend) {
    if (coo.nnz < 0 || (size_t)coo.nnz > coo.row_idx.size() || (size_t)coo.nnz > coo.col_idx.size())
        throw std::runtime_error("COOMatrix nnz does not match its index arrays");
    if (u_tile <= 0) throw std::runtime_error("CSC tile size must be positive");
    if (coo.rows <= 0 || coo.cols <= 0) throw std::runtime_error("bad matrix shape");
    const int64_t pitch = ((int64_t)coo.cols / 2 + 1 + 3) / 4 * 4;
    const size_t n = (size_t)coo.rows * (size_t)pitch;
    hipFloatComplex* pinned = nullptr;
    if (hipHostMalloc(reinterpret_cast<void**>(&pinned), n * sizeof(hipFloatComplex), hipHostMallocDefault) != hipSuccess)
        throw std::runtime_error("hipHostMalloc failed for the streamed output");
    struct Pinned {
        void* p;
        ~Pinned() { (void)hipHostFree(p); }
    } guard{pinned};
    spfft_opts o;
    spfft_reference_opts(fft_backend == 3 ? SPFFT_SMOOTH_7 : SPFFT_SMOOTH_3, u_tile, &o);
    spfft_result r;
    r.struct_size = sizeof(r);
    if (spfft_run_host_coo_to_host(coo.rows, coo.cols, coo.nnz, coo.row_idx.data(), coo.col_idx.data(), &o,
                                   pinned, pitch, &r) != SPFFT_OK)
        throw std::runtime_error(spfft_last_error());
    SparseFFTResult out;
    out.d_output = nullptr;
    out.output_cols = r.output_cols;
    out.ms = r.ms;
    out.mem_bytes = r.mem_bytes;
    out.h_output.assign(pinned, pinned + n);
    return out;
}

SparseFFTResult sparse_fft_csc_bluestein_cufft_streaming(const COOMatrix& coo, int u_tile) {
    return sparse_fft_csc_stockham_streaming(coo, u_tile, 3);
}

COOMatrix read_mtx(const std::string& path) {
    spfft_host_coo h;
    const spfft_status st = spfft_read_mtx(path.c_str(), &h);
    if (st != SPFFT_OK) {
        spfft_coo_free(&h);
        throw std::runtime_error("Cannot read Matrix Market file: " + path);
    }
    COOMatrix c;
    c.rows = h.rows;
    c.cols = h.cols;
    c.nnz = h.nnz;
    c.row_idx.assign(h.row_idx, h.row_idx + h.nnz);
    c.col_idx.assign(h.col_idx, h.col_idx + h.nnz);
    spfft_coo_free(&h);
    return c;
}
