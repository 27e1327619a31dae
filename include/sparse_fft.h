// sparse_fft.h -- drop-in C++ entry points for the two reported variants of the
// reference (include/sparse_fft.h:7-13, 57, 70), implemented over the C ABI in
// spfft_c.h and hand-written gfx950 kernels.
#pragma once
#include <hip/hip_complex.h>
#include <vector>

#include "mtx_reader.h"

struct SparseFFTResult {
    hipFloatComplex* d_output;  // device pointer, rows * output_cols; caller frees with hipFree
    int output_cols;            // row pitch P = 4 * ceil((cols/2 + 1) / 4)
    float ms;                   // GPU time of the tile loop (the reference's timed region)
    size_t mem_bytes;           // device free-memory delta over the call
    std::vector<hipFloatComplex> h_output;  // streamed variants: rows * output_cols, host memory
};

// Every entry point here runs the reference's algorithm, the Bluestein row pass at
// its smooth length (spfft_reference_opts); the native C ABI (spfft_c.h) also
// offers the direct and Rader row transforms of the same half spectrum.
//
// Bluestein length next_3_smooth(2*cols - 1).  Replaces
// sparse_fft_csc_bluestein_mixed_radix (reference src/sparse_fft_bluestein.cu:2256-2392).
SparseFFTResult sparse_fft_csc_bluestein_mixed_radix(const COOMatrix& coo, int u_tile = 128);

// Same transform as sparse_fft_csc_bluestein_mixed_radix, with the CSC build and the
// whole tile loop captured into one hipGraph and launched once; ms is the GPU time of
// that launch.  Replaces sparse_fft_csc_bluestein_mixed_radix_graph (reference
// include/sparse_fft.h:63, src/sparse_fft_bluestein.cu:2402-2549).
SparseFFTResult sparse_fft_csc_bluestein_mixed_radix_graph(const COOMatrix& coo, int u_tile = 128);

// Bluestein length next_7_smooth(2*cols - 1), hand-written radix-2/3/4/5/7/8
// sub-FFTs in place of cuFFT.  Replaces sparse_fft_csc_bluestein_cufft_smooth
// (reference src/sparse_fft_bluestein.cu:2557-2719).
SparseFFTResult sparse_fft_csc_bluestein_cufft_smooth(const COOMatrix& coo, int u_tile = 128);

// Streamed output: only two chunks of 2 * u_tile output rows on the device; every
// tile's rows go to pinned host memory on a copy stream while the next tile
// computes.  d_output is NULL and h_output holds rows * output_cols entries; ms
// includes the copies.  fft_backend 3 takes the 7-smooth Bluestein length (the
// reference's cuFFT backend), any other value the 3-smooth one (its Stockham
// kernels).  Replaces sparse_fft_csc_stockham_streaming and
// sparse_fft_csc_bluestein_cufft_streaming (reference include/sparse_fft.h:35-37,
// src/sparse_fft_bluestein.cu:912-1149).
SparseFFTResult sparse_fft_csc_stockham_streaming(const COOMatrix& coo, int u_tile = 128, int fft_backend = 0);
SparseFFTResult sparse_fft_csc_bluestein_cufft_streaming(const COOMatrix& coo, int u_tile = 128);
