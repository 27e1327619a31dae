// mtx_reader.h -- the reference's COO input type and Matrix Market reader
// (reference include/mtx_reader.h:7-18), for the C++ entry points in
// sparse_fft.h.
#pragma once
#include <string>
#include <vector>

// Binary sparse matrix: every stored (row, col) pair has the value 1.
struct COOMatrix {
    int rows;
    int cols;
    int nnz;
    std::vector<int> row_idx;
    std::vector<int> col_idx;
};

// Reads a Matrix Market file (symmetric entries mirrored, 1-based indices made
// 0-based, duplicates removed, sorted by (row, col)); throws std::runtime_error
// on failure.  Replaces read_mtx (reference src/mtx_reader.cu:7-77).
COOMatrix read_mtx(const std::string& path);
