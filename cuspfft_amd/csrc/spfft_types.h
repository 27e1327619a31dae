// spfft_types.h -- plain structs shared by the host planner and the gfx950
// kernels (passed by value as kernel arguments).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spfft {

constexpr int kMaxStages = 8;

// Radix sequence of one sub-FFT of length L (product of radix[0..nstages)).
struct FftStages {
    int L;
    int nstages;
    int radix[kMaxStages];
    int ns_before[kMaxStages];  // product of the radices of the earlier stages
};

// Everything one u-tile of the column pass and the Bluestein row pass needs.
//
// Four-step layout with M = M1*M2, n = M2*n1 + n2, k = k1 + M1*k2:
//   k_colpass : X0[u][c] = h[c] * sum_{r in col c} W_R^{(u r) mod R}, c < C
//   k_fft1    : per (u, n2): FFT over n1 of x[M2 n1 + n2] (zero for n >= C),
//               times W_M^{k1 n2}            -> T[u][k1][n2]
//   k_conv    : per (u, k1): FFT over n2 (M2), times B^[k1 + M1 k2] / M,
//               inverse FFT over k2          -> T[u][k1][n2]
//   k_ifft1   : per (u, n2): times W_M^{-k1 n2}, inverse FFT over k1 -> y[M2 n1 + n2];
//               X[u][v] = h[v] * y[v], v < V
struct RowPassArgs {
    int R, C, V, P;            // rows, cols, C/2+1, output pitch (elements)
    int M, M1, M2;
    int u_base, tile_rows;     // first output row of this tile, rows in tile
    int sub_rows;              // four-step kernels: rows of T per sub-tile (0: the whole tile)
    int ldx;                   // row stride of X0 (elements)
    int G2;                    // k_conv: (u, k1) rows per workgroup
    int conv_kind;             // 1: k_conv_fast<f2.radix[0], f2.radix[1]>, 0: k_conv_generic
    int colpass_rows;          // k_colpass: rows per lane (1 .. 16; 64 * rows per workgroup)
    int pair_half;             // k_colpass: 1 = rows u <= R/4 each also give row R/2 - u (R even,
                               //   one tile of rows [0, R/2])
    int hi_shift;              // two-level W_R table: W_R^p = hi[p >> s] * lo[p & (2^s - 1)]
    int n_hi;                  // entries in hi table
    float inv_r;               // (1 - 2^-20) / R: a phase quotient estimate that never exceeds
    uint64_t r_magic;          // fastmod magic for R
    FftStages f2;              // radix stages of the length-M2 FFT in k_conv
    const int* colstart;       // C+1 dense column starts into row_idx
    const int* col_sched;      // k_colpass: per 64-column tile, its columns grouped by wave
    const int* col_sched_off;  // k_colpass: per tile, the groups' offsets (waves + 1)
    const void* row_idx;       // uint16 (packed) or int32 rows, CSC order
    const float2* h;           // chirp h[k] = exp(-i pi k^2 / C), k < C
    const float2* wr_hi;
    const float2* wr_lo;
    const float2* wr_full;     // W_R^p, p < R, when the whole table fits in LDS (else NULL)
    const float2* tw2;         // W_M2^i, i < M2
    const float2* twM;         // [k1*M2 + n2] = W_M^{k1 n2}
    const float2* wm2;         // W_M^e = wm2[256 + (e >> 8)] * wm2[e & 255], e < M (k_rowfused)
    const float2* bhat;        // [k1*M2 + k2] = B^[k1 + M1 k2] / M
    const float* mfma_dft;     // precision form: cos/sin(2 pi i / M1) hi and lo BF16 parts, [4][M1]
    const void* mfma_frag;     // precision form of k_rowrader: the lanes' BF16 A-operand fragments of
                               //   the 32-point DFT matrices C and S, [C/S][hi/lo][k-step][lane][8]
    int precision;             // 0: FP32 codelets, 1: BF16x3 MFMA length-M1 sub-FFTs
    int fused;                 // 1: k_rowfused<M1, fused_nb, fused_rpw, f2.radix...> replaces
                               //    k_fft1 + k_conv + k_ifft1 (no tile buffer T)
    int fused_nb, fused_rpw;   // k_rowfused: LDS batches per row, rows per workgroup
    int fused_nt;              // k_rowfused: threads per workgroup
    int direct;                // 1: k_rowfused runs the direct length-C FFT (M = C, no Bluestein)
    int fused_cus;             // k_rowfused: > 0: persistent launch on this many CUs (resident
                               //   workgroups loop over row groups); 0: one workgroup per group
    int rader;                 // 1: k_rowrader (C = 2p, p an odd prime, M = p - 1; fused = 1)
    int rader_p;               // k_rowrader: p
    const int* rader_k;        // k_rowrader: [n2 + M2 n1] = g^-q mod p, q = (M2 n1 + M1 n2) mod (p - 1)
                               //   (g a primitive root; the prime-factor order of the kernel)
    int split_n1, split_n2;    // split row transform (C = N1 N2, split_n1 > 0): pass A a row
                               //   kernel of width N2 over the N1 runs of each X0 row, pass B
                               //   k_splitcols<N1> over the tile buffer T
    int split_p2;              // T: run n1 of row u at T + (u N1 + n1) split_p2 (N2 rounded up to
                               //   16 entries: whole 128-byte lines)
    int vrow_n1, vrow_ld;      // row kernels, vrow_n1 > 0 (pass A of the split transform): tile row
                               //   r is run r % vrow_n1 of X0 row r / vrow_n1 (X0 rows of vrow_ld
                               //   entries, runs of C); 0: tile row r is X0 row r (ldx entries)
    uint32_t vrow_magic;       // ceil(2^32 / vrow_n1): r / vrow_n1 = umulhi(r, vrow_magic) for
                               //   r (vrow_n1 - 1) < 2^32 (tile rows r < 2^22, vrow_n1 <= 127)
    const float2* split_w;     // k_splitcols: W_C^e = split_w[256 + (e >> 8)] * split_w[e & 255]
    const float2* split_dft;   // k_splitcols, odd prime N1 > 31: (cos, sin)(2 pi p q / N1) at
                               //   [(p - 1) H + q - 1], p, q = 1 .. H = (N1 - 1) / 2
    const int* colperm;        // column pass: X0 column (position) i holds matrix column
                               //   colperm[i] (NULL: the identity; k_rowrader's order)
    float2* X0;                // tile pass-1 output [u_local][ldx]
    float2* T;                 // tile workspace [u_local][k1][n2]
    float2* out;               // output base (row 0), pitch P: row u at out + u P
    float2* out_m;             // mirror rows: row R - u at out_m + (R - u) P (= out, or a
                               //   streamed tile's chunk)
};

}  // namespace spfft
