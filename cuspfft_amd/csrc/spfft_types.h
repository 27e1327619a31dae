// spfft_types.h -- plain structs shared by the host planner and the gfx950
// kernels (passed by value as kernel arguments, or as arrays in device memory
// for the grouped batch launches).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spfft {

constexpr int kMaxStages = 24;

// Radix sequence of one sub-FFT of length L (product of radix[0..nstages)).
struct FftStages {
    int L;
    int nstages;
    int radix[kMaxStages];
};

// Everything one row-pass launch (pass 1 + four-step Bluestein convolution +
// finalize) needs for one matrix and one u-tile.
//
// Four-step layout with M = M1*M2, n = M2*n1 + n2, k = k1 + M1*k2:
//   K1 (colpass_fft1): x[n] = h[c] * sum_{r in col c} W_R^{(u r) mod R} at
//      c = n < C (0 elsewhere); FFT over n1 (length M1) -> T[u][k1][n2]
//   K2 (conv_rows):   per (u, k1): *twM[k1][n2], FFT over n2 (M2),
//      *bhat[k1][k2] (= B^[k1 + M1 k2] / M), inverse FFT over k2, *conj(twM)
//   K3 (ifft1_finalize): per (u, n2): inverse FFT over k1 (length M1) ->
//      y[M2*n1 + n2]; out[u][v] = h[v] * y[v] for v < V.
struct RowPassArgs {
    int R, C, V, P;            // rows, cols, C/2+1, output pitch (elements)
    int M, M1, M2;
    int U, J;                  // K1/K3: rows and n2-slab width per workgroup
    int G2;                    // K2: (u, k1) rows per workgroup
    int u_base, tile_rows;     // first output row of this tile, rows in tile
    int hi_shift;              // two-level W_R table: W_R^p = hi[p >> s] * lo[p & (2^s - 1)]
    int n_hi;                  // entries in hi table
    uint64_t r_magic;          // fastmod magic for R
    FftStages f1, f2;          // sub-FFT plans for M1 and M2
    const int* colstart;       // C+1 dense column starts into row_idx
    const void* row_idx;       // uint16 (packed) or int32 rows, CSC order
    const float2* h;           // chirp h[k] = exp(-i pi k^2 / C), k < C
    const float2* wr_hi;
    const float2* wr_lo;
    const float2* tw1;         // W_M1^i, i < M1
    const float2* tw2;         // W_M2^i, i < M2
    const float2* twM;         // [k1*M2 + n2] = W_M^{k1 n2}
    const float2* bhat;        // [k1*M2 + k2] = B^[k1 + M1 k2] / M
    float2* T;                 // tile workspace [u_local][k1][n2]
    float2* out;               // output base (row 0), pitch P
};

}  // namespace spfft
