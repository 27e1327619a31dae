// rowpass.hip -- column sparse-DFT pass and Bluestein row pass for gfx950.
//
// Replaces, per u-tile, the reference's memset + csc_build_bluestein_input
// + run_fft_mixed_23 (forward) + pointwise_mul_batched + run_fft_mixed_23
// (inverse) + scale + bluestein_finalize sequence
// (src/sparse_fft_bluestein.cu:2327-2370 and :2649-2696) with four kernels
// on a four-step split M = M1*M2 of the smooth Bluestein length:
//
//   k_colpass  pass 1.  A wave owns one column at a time and its 64 lanes own
//              64 consecutive output rows u, so the column is uniform over the
//              wave; its row indices are read 64 at a time and broadcast with
//              readlane, and every lane sums W_R^{(u r) mod R} with the exact
//              integer phase.  W_R comes from an LDS table.  The chirp h[c] is
//              applied and an LDS tile turns the column sums into whole X0 row
//              segments (the reference writes with stride M).
//   k_fft1<M1> one thread per (u, n2): the length-M1 FFT over n1 of
//              x[M2 n1 + n2] entirely in registers (compile-time codelet; the
//              zero padding c >= C is never loaded), times W_M^{k1 n2}.
//   k_conv     one workgroup per G2 rows (u, k1) of length M2: Stockham stages
//              of codelet radices through LDS; the first stage reads T from
//              HBM, the last forward stage multiplies by B^/M, the last
//              inverse stage writes T back.  One read and one write of T.
//   k_ifft1<M1> one thread per (u, n2): times W_M^{-k1 n2}, inverse length-M1
//              codelet, chirp post-multiply, and the half-spectrum write of
//              X[u][v] for v = M2 n1 + n2 <= C/2 at pitch P.
//
// When the plan has a k_rowfused shape for M (fused_impl.h), one kernel does
// k_fft1 + k_conv + k_ifft1 per row on chip and T is not used.
//
// Only rows u <= R/2 go through the four kernels: the input is real, so each
// column DFT is Hermitian and X[R-u][v] = conj(X[u][(C - v) mod C]).  The
// Bluestein convolution yields the full length-C spectrum of row u, and
// k_ifft1 writes both row u and its mirror row R - u from it -- half of the
// reference's pass-1 and row-FFT work for the same output.
#include "rowpass_internal.h"

namespace spfft {

namespace {

constexpr int kColThreads = 512;
constexpr int kColWaves = kColThreads / 64;
// k_colpass_mx: columns per wave (tiles of 8 * CPW columns)
constexpr int kColpassMxCpw = 1;

// Columns per k_colpass tile for U output rows per lane (64 U rows per tile).
__host__ __device__ constexpr int colpass_tile_cols(int rows_per_lane) {
    return rows_per_lane == kColpassMfma ? kColWaves * kColpassMxCpw
           : rows_per_lane == 1          ? 64
           : rows_per_lane == 2          ? 32
           : rows_per_lane == 4          ? 16
                                         : 8;  // U = 8, 16
}

// ---------------------------------------------------------------------------
// k_colsched: for every TC-column tile of k_colpass, shares its columns out to
// the kColWaves waves by longest-processing-time first: power-of-two classes of
// cost (nonzeros + 1) from the largest down, each column to the least-loaded
// wave.  One wave per tile; lane = column.  Output: the tile's columns grouped
// by wave, sched[tile][0 .. TC), and the groups' offsets off[tile][0 .. kColWaves].
// Columns past C are not listed.  With a column order (colperm), position c of
// the tile is matrix column colperm[c].
// ---------------------------------------------------------------------------
__device__ __forceinline__ void colsched_body(const int* __restrict__ colstart, const int* __restrict__ colperm,
                                              int C, int TC, int* __restrict__ sched, int* __restrict__ off,
                                              const int bx) {
    const int lane = threadIdx.x;
    const int c = bx * TC + lane;
    const int mc = c < C && colperm ? colperm[c] : c;  // the matrix column at position c
    const int cost = lane < TC && c < C ? colstart[mc + 1] - colstart[mc] + 1 : 0;
    const int cls = cost > 0 ? 31 - __clz(cost) : -1;
    int load[kColWaves] = {}, cnt[kColWaves] = {};
    int mine_w = 0, mine_k = 0;  // this lane's column: wave and position in its group
    for (int b = 30; b >= 0; b--) {
        uint64_t m = __ballot(cls == b);
        while (m) {
            const int l = __ffsll((unsigned long long)m) - 1;
            m &= m - 1;
            int w = 0;
#pragma unroll
            for (int k = 1; k < kColWaves; k++) w = load[k] < load[w] ? k : w;
            if (lane == l) {
                mine_w = w;
                mine_k = cnt[w];
            }
            load[w] += __builtin_amdgcn_readlane(cost, l);
            cnt[w] += 1;
        }
    }
    int start[kColWaves + 1];
    start[0] = 0;
#pragma unroll
    for (int k = 0; k < kColWaves; k++) start[k + 1] = start[k] + cnt[k];
    int sw = 0;
#pragma unroll
    for (int k = 0; k < kColWaves; k++) sw = mine_w == k ? start[k] : sw;
    if (cls >= 0) sched[bx * TC + sw + mine_k] = lane;
    if (lane <= kColWaves) {
        int v = 0;
#pragma unroll
        for (int k = 0; k <= kColWaves; k++) v = lane == k ? start[k] : v;
        off[bx * (kColWaves + 1) + lane] = v;
    }
}

__global__ __launch_bounds__(64) void k_colsched(const int* __restrict__ colstart,
                                                 const int* __restrict__ colperm, int C, int TC,
                                                 int* __restrict__ sched, int* __restrict__ off) {
    colsched_body(colstart, colperm, C, TC, sched, off, blockIdx.x);
}

// Batched form: work item (matrix descriptor, tile) per workgroup.
__global__ __launch_bounds__(64) void k_colsched_batch(const RowPassArgs* __restrict__ descs,
                                                       const int2* __restrict__ work) {
    const int2 w = work[blockIdx.x];
    const RowPassArgs& a = descs[w.x];
    colsched_body(a.colstart, a.colperm, a.C, colpass_tile_cols(a.colpass_rows), const_cast<int*>(a.col_sched),
                  const_cast<int*>(a.col_sched_off), w.y);
}

// acc[0] += w, acc[t] += w s^t (t < N): a lane's rows 64 apart share one phase
// lookup.  w s is one complex product; the further rows follow the three-term
// recurrence w s^(t+1) = 2 Re(s) w s^t - w s^(t-1) (|s| = 1), two FMAs a row instead
// of a complex product's four operations.  Its float32 error grows faster with t than
// the products' (8 rows: RMS 2.1e-7 against 1.0e-7 per term, at most 2e-6; 16 rows:
// 6.1e-7, at most 8e-6, over 20000 random phases of R = 40421).
template <int N>
__device__ __forceinline__ void accumulate_rows(float2 (&acc)[N], float2 w, float2 s) {
    acc[0].x += w.x;
    acc[0].y += w.y;
    if constexpr (N > 1) {
        float2 wp = w, wc = cmul(w, s);
        acc[1].x += wc.x;
        acc[1].y += wc.y;
        const float c2 = s.x + s.x;
#pragma unroll
        for (int t = 2; t < N; t++) {
            const float2 wn = make_float2(fmaf(c2, wc.x, -wp.x), fmaf(c2, wc.y, -wp.y));
            acc[t].x += wn.x;
            acc[t].y += wn.y;
            wp = wc;
            wc = wn;
        }
    }
}

// ---------------------------------------------------------------------------
// k_colpass: a tile of TC columns x 64*U rows per workgroup (8 waves; PAIR:
// 64*U rows u and their 64*U partner rows).
// The tile's columns are shared out to the waves by k_colsched, once per
// transform; a wave takes its columns one at a time, so the column is uniform over
// the wave; lane l takes rows u0 + l + 64t (t < U).  A column's row indices are
// read 64 at a time (one per lane) together with their per-nonzero constants,
// computed once per lane:
//   base = (u0 r) mod R, d64 = (64 r) mod R, w64 = W_R^{d64}
// and then broadcast one nonzero at a time with readlane, so the inner loop has
// no memory wait and no divergence.  Per nonzero and lane:
//   ph = (base + l r) mod R     (x = base + l r < 64 R; q = floor(x * inv_r) with inv_r
//                                 just below 1/R is floor(x / R) or one less, so
//                                 x - q R < 2 R needs one conditional subtraction)
//   w  = W_R^ph                 (LDS table: whole, or two-level hi[ph >> s] * lo[ph & mask])
//   acc[0] += w, acc[t] += w * w64^t
// The chirp h[c] is applied and the sums go through an LDS tile (re and im
// planes, odd row stride TC + 1: conflict-free both ways), so X0 is written in
// whole TC-column row segments (TC * 8 bytes).
//
// PAIR (R even, one tile of rows [0, H], H = R/2): W_R^H = -1, so row u + H
// sums (-1)^r W_R^{u r}.  A lane keeps the even-r and odd-r sums E, O of its
// rows u <= H/2 (each 64-nonzero chunk is permuted across lanes so that its
// even rows come first: two straight loops, no branch per nonzero) and emits
//   S[u] = E + O,   S[H - u] = conj(S[u + H]) = conj(E - O)
// (S[R - u'] = conj(S[u']) for a real input), through a second LDS tile.  One
// phase reduction and lookup per nonzero now serves twice the rows.
// ---------------------------------------------------------------------------
template <bool PACKED, bool FULL, int U, int TC, bool PAIR>
__device__ __forceinline__ void colpass_body(const RowPassArgs& a, const int bx, const int by) {
    extern __shared__ float2 smem[];
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr int TS = TC + 1;        // tile row stride
    const int c0 = bx * TC;
    const int u0l = by * 64 * U;
    const uint32_t R = (uint32_t)a.R;
    const int nlo = 1 << a.hi_shift;
    // the tables first, so that their offsets fit the LDS instructions' immediate field
    float2* s_lo = smem;  // FULL: the whole W_R table
    float2* s_hi = s_lo + nlo;
    constexpr int NTILE = PAIR ? 2 : 1;  // row tiles: rows u, and (PAIR) rows H - u
    float* t_re = reinterpret_cast<float*>(FULL ? s_lo + R : s_hi + a.n_hi);  // [NTILE][64U][TS]
    float* t_im = t_re + NTILE * 64 * U * TS;                                  // [NTILE][64U][TS]
    int* s_cs = reinterpret_cast<int*>(t_im + NTILE * 64 * U * TS);          // [TC] column starts
    int* s_ce = s_cs + TC;                                                    // [TC] column ends
    if (FULL) {
        for (int i = threadIdx.x; i < (int)R; i += kColThreads) s_lo[i] = a.wr_full[i];
    } else {
        for (int i = threadIdx.x; i < nlo; i += kColThreads) s_lo[i] = a.wr_lo[i];
        for (int i = threadIdx.x; i < a.n_hi; i += kColThreads) s_hi[i] = a.wr_hi[i];
    }
    // X0 position c0 + i holds matrix column colperm[c0 + i] (identity when NULL)
    for (int i = threadIdx.x; i < TC; i += kColThreads) {
        const int pos = c0 + i;
        int b = 0, e = 0;
        if (pos < a.C) {
            const int mc = a.colperm ? a.colperm[pos] : pos;
            b = a.colstart[mc];
            e = a.colstart[mc + 1];
        }
        s_cs[i] = b;
        s_ce[i] = e;
    }
    __syncthreads();

    const uint32_t mask = (uint32_t)nlo - 1u;
    const int hs = a.hi_shift;
    auto wr = [&](uint32_t ph) -> float2 {
        return FULL ? s_lo[ph] : cmul(s_hi[ph >> hs], s_lo[ph & mask]);
    };
    const uint32_t u0 = (uint32_t)(a.u_base + u0l);
    // this wave's columns (k_colsched): sched[blk][off[w] .. off[w+1])
    const int* sched = a.col_sched + bx * TC;
    const int* off = a.col_sched_off + bx * (kColWaves + 1);
    const int i0 = __builtin_amdgcn_readfirstlane(off[wid]), i1 = __builtin_amdgcn_readfirstlane(off[wid + 1]);
    for (int i = i0; i < i1; i++) {
        const int cc = __builtin_amdgcn_readfirstlane(sched[i]);  // column of the tile
        const int c = c0 + cc;
        float2 acc[U], acc_o[U];  // PAIR: acc sums even r, acc_o odd r
#pragma unroll
        for (int t = 0; t < U; t++) acc[t] = acc_o[t] = make_float2(0.f, 0.f);
        if (c < a.C) {
            const int beg = s_cs[cc], end = s_ce[cc];
            for (int p0 = beg; p0 < end; p0 += 64) {
                const int n = min(64, end - p0);
                // per-nonzero constants, one nonzero per lane
                uint32_t r = 0;
                if (lane < n)
                    r = PACKED ? (uint32_t)((const uint16_t*)a.row_idx)[p0 + lane]
                               : (uint32_t)((const int*)a.row_idx)[p0 + lane];
                int n_even = n;
                if (PAIR) {
                    // the chunk's even rows first, then its odd rows (one forward permute)
                    const uint64_t valid = n >= 64 ? ~0ull : (1ull << n) - 1;
                    const uint64_t om = __ballot((r & 1u) != 0) & valid, em = ~om & valid;
                    n_even = n - __popcll(om);
                    const uint64_t m = (r & 1u) ? om : em;
                    int pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                    pos = lane >= n ? lane : (r & 1u) ? n_even + pos : pos;
                    r = (uint32_t)__builtin_amdgcn_ds_permute(pos << 2, (int)r);
                }
                const uint32_t base = PACKED ? fastmod_u32(u0 * r, a.r_magic, R)
                                             : (uint32_t)(((uint64_t)u0 * r) % (uint64_t)R);
                const uint32_t d64 = fastmod_u32(r << 6, a.r_magic, R);
                float2 w64 = make_float2(1.f, 0.f);
                if (U > 1) w64 = wr(d64);
                // nonzeros jb .. je of the chunk into the lane's U row sums A
                auto run = [&](int jb, int je, float2(&A)[U]) {
#pragma unroll 2
                    for (int j = jb; j < je; j++) {
                        const uint32_t rj = (uint32_t)__builtin_amdgcn_readlane((int)r, j);
                        const uint32_t bj = (uint32_t)__builtin_amdgcn_readlane((int)base, j);
                        // (u0 + lane) r mod R = (bj + lane rj) mod R, bj + lane rj < 64 R
                        const uint32_t x = bj + __umul24((uint32_t)lane, rj);
                        const uint32_t q = (uint32_t)((float)x * a.inv_r);
                        const uint32_t y = x - __umul24(q, R);  // in [0, 2R)
                        const uint32_t ph = min(y, y - R);      // y - R wraps when y < R
                        float2 sj = make_float2(1.f, 0.f);
                        if (U > 1)
                            sj = make_float2(__builtin_bit_cast(float, __builtin_amdgcn_readlane(
                                                 __builtin_bit_cast(int, w64.x), j)),
                                             __builtin_bit_cast(float, __builtin_amdgcn_readlane(
                                                 __builtin_bit_cast(int, w64.y), j)));
                        accumulate_rows<U>(A, wr(ph), sj);
                    }
                };
                run(0, n_even, acc);
                if (PAIR) run(n_even, n, acc_o);
            }
            // the direct and Rader transforms have no chirp (h = 1): no product
            const bool chirp = !(a.direct || a.rader);
            const float2 hc = chirp ? a.h[c] : make_float2(1.f, 0.f);
#pragma unroll
            for (int t = 0; t < U; t++) {
                if (PAIR) {
                    const float2 e = acc[t], o = acc_o[t];
                    acc[t] = make_float2(e.x + o.x, e.y + o.y);
                    acc_o[t] = make_float2(e.x - o.x, o.y - e.y);
                    if (chirp) acc_o[t] = cmul(acc_o[t], hc);
                }
                if (chirp) acc[t] = cmul(acc[t], hc);
            }
        }
#pragma unroll
        for (int t = 0; t < U; t++) {
            t_re[(t * 64 + lane) * TS + cc] = acc[t].x;
            t_im[(t * 64 + lane) * TS + cc] = acc[t].y;
            if (PAIR) {
                t_re[((U + t) * 64 + lane) * TS + cc] = acc_o[t].x;
                t_im[((U + t) * 64 + lane) * TS + cc] = acc_o[t].y;
            }
        }
    }
    __syncthreads();
    // whole TC-column row segments of X0 (ldx is a multiple of 64, or C for the split
    // transform: positions past the row are not written)
    const int half = a.R >> 1, qmax = half >> 1;  // PAIR: rows u <= qmax and H - u
    for (int e = threadIdx.x; e < NTILE * 64 * U * TC; e += kColThreads) {
        const int row = e / TC, col = e % TC;
        int ul;
        bool ok;
        if (PAIR) {
            const bool partner = row >= 64 * U;
            const int u = u0l + (partner ? row - 64 * U : row);
            ul = partner ? half - u : u;
            ok = u <= qmax && (!partner || ul != u);  // H - u == u only for u = H/2: written once
        } else {
            ul = u0l + row;
            ok = ul < a.tile_rows;
        }
        if (ok && c0 + col < a.ldx) {  // X0 is read once, by the row pass: non-temporal
            typedef float f32x2 __attribute__((ext_vector_type(2)));
            const f32x2 x = {t_re[row * TS + col], t_im[row * TS + col]};
            __builtin_nontemporal_store(x, reinterpret_cast<f32x2*>(&a.X0[(size_t)ul * a.ldx + c0 + col]));
        }
    }
}

template <bool PACKED, bool FULL, int U, int TC, bool PAIR>
__global__ __launch_bounds__(kColThreads) void k_colpass(RowPassArgs a) {
    colpass_body<PACKED, FULL, U, TC, PAIR>(a, blockIdx.x, blockIdx.y);
}

// Batched form: one launch for the matrices of one instantiation; work item w =
// (matrix descriptor, column tile, row block) per workgroup.
template <bool PACKED, bool FULL, int U, int TC, bool PAIR>
__global__ __launch_bounds__(kColThreads) void k_colpass_batch(const RowPassArgs* __restrict__ descs,
                                                               const int4* __restrict__ work) {
    const int4 w = work[blockIdx.x];
    colpass_body<PACKED, FULL, U, TC, PAIR>(descs[w.x], w.y, w.z);
}

// ---------------------------------------------------------------------------
// k_colpass_mx: the column pass on the matrix cores in FP32 (colpass_rows
// kColpassMfma).  A workgroup (8 waves) takes 8 CPW columns x 1024 rows, CPW
// columns per wave.  With u = ub + 32 u1 + u0 (u1, u0 < 32) the phase splits as
//   W_R^{u r} = W_R^{(ub + 32 u1) r} * W_R^{u0 r}
// so a column's block of 1024 row sums is a complex product of a 32 x nnz and an
// nnz x 32 matrix.  v_mfma_f32_32x32x2f32 takes one nonzero per instruction, its
// two K slots holding the real and imaginary parts (lane l supplies A[l % 32][l / 32]
// and B[l / 32][l % 32]; element v of lane l is D[8 (v / 4) + 4 (l / 32) + v % 4][l % 32]):
//   re:  A = (a.x, -a.y), B = (b.x, b.y)  ->  sum a.x b.x - a.y b.y
//   im:  A = (a.x,  a.y), B = (b.y, b.x)  ->  sum a.x b.y + a.y b.x
// with a = W_R^{(ub + 32 u1) r} (u1 = l % 32) and b = W_R^{u0 r} (u0 = l % 32), both
// from the exact integer phase and the LDS table as in k_colpass.  Products and sums
// are FP32 (the matrix cores' FP32 path), so the result is the plain form's.  Per
// nonzero and 1024 rows: two table lookups per lane and two MFMAs, against a lookup
// and 7 complex products per lane and 512 rows in k_colpass.
// PAIR (R even, one tile of rows [0, H]): as in k_colpass, each 64-nonzero chunk is
// permuted so that its even rows come first; the even-r and odd-r nonzeros go into
// their own accumulators E and O (two straight loops), and rows u and H - u come from
// E + O and conj(E - O), written in two passes through the same LDS tile.
// ---------------------------------------------------------------------------
typedef float mx_f32x16 __attribute__((ext_vector_type(16)));

template <bool PACKED, bool FULL, bool PAIR, int CPW>
__device__ __forceinline__ void colpass_mx_body(const RowPassArgs& a, const int bx, const int by) {
    extern __shared__ float2 smem[];
    constexpr int TC = kColWaves * CPW, RB = 1024, TS = TC + 1;
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c0 = bx * TC;
    const int u0l = by * RB;
    const uint32_t R = (uint32_t)a.R;
    const int nlo = 1 << a.hi_shift;
    float2* s_lo = smem;  // FULL: the whole W_R table
    float2* s_hi = s_lo + nlo;
    float* t_re = reinterpret_cast<float*>(FULL ? s_lo + R : s_hi + a.n_hi);  // [RB][TS]
    float* t_im = t_re + RB * TS;                                              // [RB][TS]
    if (FULL) {
        for (int i = threadIdx.x; i < (int)R; i += kColThreads) s_lo[i] = a.wr_full[i];
    } else {
        for (int i = threadIdx.x; i < nlo; i += kColThreads) s_lo[i] = a.wr_lo[i];
        for (int i = threadIdx.x; i < a.n_hi; i += kColThreads) s_hi[i] = a.wr_hi[i];
    }
    __syncthreads();
    const uint32_t mask = (uint32_t)nlo - 1u;
    const int hs = a.hi_shift;
    auto wr = [&](uint32_t ph) -> float2 {
        return FULL ? s_lo[ph] : cmul(s_hi[ph >> hs], s_lo[ph & mask]);
    };
    // x < 32 R: q = floor(x inv_r) is floor(x / R) or one less
    auto modr = [&](uint32_t x) -> uint32_t {
        const uint32_t q = (uint32_t)((float)x * a.inv_r);
        const uint32_t y = x - __umul24(q, R);
        return min(y, y - R);
    };
    const uint32_t ub = (uint32_t)(a.u_base + u0l);
    const uint32_t i32 = (uint32_t)(lane & 31);
    const bool khi = lane >= 32;
    const bool chirp = !(a.direct || a.rader);
    // per column of this wave (positions c0 + wid + 8 k): E (or all) and O sums
    mx_f32x16 er[CPW], ei[CPW], orr[CPW], oi[CPW];
    float2 hc[CPW];
#pragma unroll
    for (int k = 0; k < CPW; k++) {
        er[k] = ei[k] = orr[k] = oi[k] = mx_f32x16{};
        hc[k] = make_float2(1.f, 0.f);
        const int pos = c0 + wid + kColWaves * k;  // this wave's X0 column position
        if (pos >= a.C) continue;
        const int mc = a.colperm ? a.colperm[pos] : pos;
        const int beg = __builtin_amdgcn_readfirstlane(a.colstart[mc]);
        const int end = __builtin_amdgcn_readfirstlane(a.colstart[mc + 1]);
        if (chirp) hc[k] = a.h[pos];
        for (int p0 = beg; p0 < end; p0 += 64) {
            const int n = min(64, end - p0);
            // per-nonzero constants, one nonzero per lane: r, (ub r) mod R, (32 r) mod R
            uint32_t r = 0;
            if (lane < n)
                r = PACKED ? (uint32_t)((const uint16_t*)a.row_idx)[p0 + lane]
                           : (uint32_t)((const int*)a.row_idx)[p0 + lane];
            int n_even = n;
            if (PAIR) {
                // the chunk's even rows first, then its odd rows (one forward permute), so
                // that each accumulator pair takes one straight loop
                const uint64_t valid = n >= 64 ? ~0ull : (1ull << n) - 1;
                const uint64_t om = __ballot((r & 1u) != 0) & valid, em = ~om & valid;
                n_even = n - __popcll(om);
                const uint64_t m = (r & 1u) ? om : em;
                int dst = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                dst = lane >= n ? lane : (r & 1u) ? n_even + dst : dst;
                r = (uint32_t)__builtin_amdgcn_ds_permute(dst << 2, (int)r);
            }
            const uint32_t base = PACKED ? fastmod_u32(ub * r, a.r_magic, R)
                                         : (uint32_t)(((uint64_t)ub * r) % (uint64_t)R);
            const uint32_t s32 = fastmod_u32(r << 5, a.r_magic, R);
            auto run = [&](int jb, int je, mx_f32x16& accr, mx_f32x16& acci) {
                for (int j = jb; j < je; j++) {
                    const uint32_t rj = (uint32_t)__builtin_amdgcn_readlane((int)r, j);
                    const uint32_t bj = (uint32_t)__builtin_amdgcn_readlane((int)base, j);
                    const uint32_t sj = (uint32_t)__builtin_amdgcn_readlane((int)s32, j);
                    const float2 wa = wr(modr(bj + __umul24(i32, sj)));  // row block u1 = l % 32
                    const float2 wb = wr(modr(__umul24(i32, rj)));       // row u0 = l % 32
                    const float are = khi ? -wa.y : wa.x, aim = khi ? wa.y : wa.x;
                    const float bre = khi ? wb.y : wb.x, bim = khi ? wb.x : wb.y;
                    accr = __builtin_amdgcn_mfma_f32_32x32x2f32(are, bre, accr, 0, 0, 0);
                    acci = __builtin_amdgcn_mfma_f32_32x32x2f32(aim, bim, acci, 0, 0, 0);
                }
            };
            run(0, n_even, er[k], ei[k]);
            if (PAIR) run(n_even, n, orr[k], oi[k]);
        }
    }
    const int half = a.R >> 1, qmax = half >> 1;  // PAIR: rows u <= qmax and H - u
    const int base_row = 4 * (lane >> 5);
    for (int pass = 0; pass < (PAIR ? 2 : 1); pass++) {
        // the wave's 1024-row sums (pass 1 of PAIR: the partner rows) into its tile columns
#pragma unroll
        for (int k = 0; k < CPW; k++) {
            const int col = wid + kColWaves * k;
#pragma unroll
            for (int v = 0; v < 16; v++) {
                float2 x = make_float2(er[k][v], ei[k][v]);
                if (PAIR) {
                    x = pass == 0 ? make_float2(er[k][v] + orr[k][v], ei[k][v] + oi[k][v])
                                  : make_float2(er[k][v] - orr[k][v], oi[k][v] - ei[k][v]);
                }
                if (chirp) x = cmul(x, hc[k]);
                const int row = 32 * (8 * (v >> 2) + base_row + (v & 3)) + (lane & 31);
                t_re[row * TS + col] = x.x;
                t_im[row * TS + col] = x.y;
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < RB * TC; e += kColThreads) {
            const int row = e / TC, col = e % TC;
            const int u = u0l + row;
            int ul;
            bool ok;
            if (PAIR) {
                ul = pass == 0 ? u : half - u;
                ok = u <= qmax && (pass == 0 || ul != u);  // H - u == u only for u = H/2: written once
            } else {
                ul = u;
                ok = u < a.tile_rows;
            }
            if (ok && c0 + col < a.ldx) {  // X0 is read once, by the row pass: non-temporal
                typedef float f32x2 __attribute__((ext_vector_type(2)));
                const f32x2 x = {t_re[row * TS + col], t_im[row * TS + col]};
                __builtin_nontemporal_store(x, reinterpret_cast<f32x2*>(&a.X0[(size_t)ul * a.ldx + c0 + col]));
            }
        }
        if (PAIR && pass == 0) __syncthreads();
    }
}

template <bool PACKED, bool FULL, bool PAIR>
__global__ __launch_bounds__(kColThreads) void k_colpass_mx(RowPassArgs a) {
    colpass_mx_body<PACKED, FULL, PAIR, kColpassMxCpw>(a, blockIdx.x, blockIdx.y);
}

template <bool PACKED, bool FULL, bool PAIR>
__global__ __launch_bounds__(kColThreads) void k_colpass_mx_batch(const RowPassArgs* __restrict__ descs,
                                                                  const int4* __restrict__ work) {
    const int4 w = work[blockIdx.x];
    colpass_mx_body<PACKED, FULL, PAIR, kColpassMxCpw>(descs[w.x], w.y, w.z);
}

}  // namespace

bool rowpass_supports_m1(int m1) {
    switch (m1) {
#define SPFFT_CASE(m) case m: return true;
        SPFFT_M1_LIST(SPFFT_CASE)
#undef SPFFT_CASE
        default: return false;
    }
}

size_t colpass_lds_bytes(const RowPassArgs& a) {
    const int U = a.colpass_rows > 0 ? a.colpass_rows : 1;
    const int TC = colpass_tile_cols(U);
    // (paired: U/2 rows per lane and their U/2 partner rows: the same tile; the
    // matrix-core kernel: 1024 rows, partner rows in a second pass)
    const size_t tile = U == kColpassMfma ? (size_t)2 * 1024 * (TC + 1) * sizeof(float)
                                          : (size_t)2 * 64 * U * (TC + 1) * sizeof(float) + (size_t)2 * TC * sizeof(int);
    if (a.wr_full) return tile + (size_t)a.R * sizeof(float2);
    return tile + ((size_t)(1 << a.hi_shift) + a.n_hi) * sizeof(float2);
}

size_t colsched_elems(int cols, size_t* off_elems) {
    // sized for the narrowest tile (8 columns)
    const size_t tiles = ((size_t)cols + 7) / 8;
    *off_elems = tiles * (kColWaves + 1);
    return tiles * 8;
}

hipError_t launch_colsched(const int* colstart, const int* colperm, int cols, int rows_per_lane, int* sched,
                           int* off, hipStream_t s) {
    if (cols <= 0) return hipSuccess;
    const int TC = colpass_tile_cols(rows_per_lane);
    hipLaunchKernelGGL(k_colsched, dim3((cols + TC - 1) / TC), dim3(64), 0, s, colstart, colperm, cols, TC, sched,
                       off);
    return hipGetLastError();
}

// The instantiated (rows per lane U, tile columns TC) pairs.
#define SPFFT_COLPASS_SHAPES(X) X(1, 64) X(2, 32) X(4, 16) X(8, 8) X(16, 8)
// The paired kernels: output rows per lane U (U/2 rows and their partners), same tiles.
#define SPFFT_COLPASS_PAIR_SHAPES(X) X(2, 32) X(4, 16) X(8, 8) X(16, 8)

hipError_t colpass_prepare(size_t lds_limit) {
    // raise the dynamic LDS limit of every k_colpass instantiation (the tile is
    // 2 * 64U * (TC + 1) floats: past 64 KB for some shapes)
    const void* ks[] = {
#define SPFFT_COLPASS_KERNELS(K, u, tc)                                    \
    reinterpret_cast<const void*>(&K<true, true, u, tc, false>),            \
        reinterpret_cast<const void*>(&K<true, false, u, tc, false>),       \
        reinterpret_cast<const void*>(&K<false, true, u, tc, false>),       \
        reinterpret_cast<const void*>(&K<false, false, u, tc, false>),
#define SPFFT_COLPASS_SINGLE(u, tc) SPFFT_COLPASS_KERNELS(k_colpass, u, tc)
#define SPFFT_COLPASS_BATCH(u, tc) SPFFT_COLPASS_KERNELS(k_colpass_batch, u, tc)
        SPFFT_COLPASS_SHAPES(SPFFT_COLPASS_SINGLE) SPFFT_COLPASS_SHAPES(SPFFT_COLPASS_BATCH)
#undef SPFFT_COLPASS_SINGLE
#undef SPFFT_COLPASS_BATCH
#undef SPFFT_COLPASS_KERNELS
#define SPFFT_COLPASS_KERNELS(K, u, tc)                                    \
    reinterpret_cast<const void*>(&K<true, true, u / 2, tc, true>),         \
        reinterpret_cast<const void*>(&K<true, false, u / 2, tc, true>),    \
        reinterpret_cast<const void*>(&K<false, true, u / 2, tc, true>),    \
        reinterpret_cast<const void*>(&K<false, false, u / 2, tc, true>),
#define SPFFT_COLPASS_SINGLE(u, tc) SPFFT_COLPASS_KERNELS(k_colpass, u, tc)
#define SPFFT_COLPASS_BATCH(u, tc) SPFFT_COLPASS_KERNELS(k_colpass_batch, u, tc)
        SPFFT_COLPASS_PAIR_SHAPES(SPFFT_COLPASS_SINGLE) SPFFT_COLPASS_PAIR_SHAPES(SPFFT_COLPASS_BATCH)
#undef SPFFT_COLPASS_SINGLE
#undef SPFFT_COLPASS_BATCH
#undef SPFFT_COLPASS_KERNELS
    };
    const void* kmx[] = {
#define SPFFT_COLPASS_MX(K)                                                                                   \
    reinterpret_cast<const void*>(&K<true, true, false>), reinterpret_cast<const void*>(&K<true, false, false>),   \
        reinterpret_cast<const void*>(&K<false, true, false>), reinterpret_cast<const void*>(&K<false, false, false>), \
        reinterpret_cast<const void*>(&K<true, true, true>), reinterpret_cast<const void*>(&K<true, false, true>),     \
        reinterpret_cast<const void*>(&K<false, true, true>), reinterpret_cast<const void*>(&K<false, false, true>),
        SPFFT_COLPASS_MX(k_colpass_mx) SPFFT_COLPASS_MX(k_colpass_mx_batch)
#undef SPFFT_COLPASS_MX
    };
    for (const void* k : ks) {
        const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit);
        if (e != hipSuccess) return e;
    }
    for (const void* k : kmx) {
        const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_colsched_batch(const RowPassArgs* descs, const int2* work, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_colsched_batch, dim3(n), dim3(64), 0, s, descs, work);
    return hipGetLastError();
}

template <bool PACKED, bool FULL>
static hipError_t launch_colpass_batch_u(const RowPassArgs& rep, const RowPassArgs* descs, const int4* work, int n,
                                         size_t lds, hipStream_t s) {
    const int U = rep.colpass_rows;
    const int TC = colpass_tile_cols(U);
    if (U == kColpassMfma) {
        if (rep.pair_half)
            hipLaunchKernelGGL((k_colpass_mx_batch<PACKED, FULL, true>), dim3(n), dim3(kColThreads), lds, s, descs, work);
        else
            hipLaunchKernelGGL((k_colpass_mx_batch<PACKED, FULL, false>), dim3(n), dim3(kColThreads), lds, s, descs, work);
        return hipGetLastError();
    }
    if (rep.pair_half) {
#define SPFFT_COLPASS_LAUNCH(u, tc)                                                                          \
    if (U == u && TC == tc) {                                                                                \
        hipLaunchKernelGGL((k_colpass_batch<PACKED, FULL, u / 2, tc, true>), dim3(n), dim3(kColThreads), lds, s, \
                           descs, work);                                                                     \
        return hipGetLastError();                                                                            \
    }
        SPFFT_COLPASS_PAIR_SHAPES(SPFFT_COLPASS_LAUNCH)
#undef SPFFT_COLPASS_LAUNCH
        return hipErrorInvalidValue;
    }
#define SPFFT_COLPASS_LAUNCH(u, tc)                                                                             \
    if (U == u && TC == tc) {                                                                                   \
        hipLaunchKernelGGL((k_colpass_batch<PACKED, FULL, u, tc, false>), dim3(n), dim3(kColThreads), lds, s, descs, \
                           work);                                                                               \
        return hipGetLastError();                                                                               \
    }
    SPFFT_COLPASS_SHAPES(SPFFT_COLPASS_LAUNCH)
#undef SPFFT_COLPASS_LAUNCH
    return hipErrorInvalidValue;
}

hipError_t launch_colpass_batch(const RowPassArgs& rep, bool packed, const RowPassArgs* descs, const int4* work,
                                int n, size_t lds, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (packed) return rep.wr_full ? launch_colpass_batch_u<true, true>(rep, descs, work, n, lds, s)
                                   : launch_colpass_batch_u<true, false>(rep, descs, work, n, lds, s);
    return rep.wr_full ? launch_colpass_batch_u<false, true>(rep, descs, work, n, lds, s)
                       : launch_colpass_batch_u<false, false>(rep, descs, work, n, lds, s);
}

void colpass_grid(const RowPassArgs& a, int* tiles_x, int* tiles_y) {
    const int U = a.colpass_rows;
    const int TC = colpass_tile_cols(U);
    *tiles_x = (a.C + TC - 1) / TC;
    if (U == kColpassMfma) {
        *tiles_y = ((a.pair_half ? (a.R >> 2) + 1 : a.tile_rows) + 1023) / 1024;
        return;
    }
    *tiles_y = a.pair_half ? ((a.R >> 2) + 1 + 64 * (U / 2) - 1) / (64 * (U / 2))
                           : (a.tile_rows + 64 * U - 1) / (64 * U);
}

int colpass_cols_per_tile(int rows_per_lane) { return colpass_tile_cols(rows_per_lane); }

template <bool PACKED, bool FULL>
static void launch_colpass_u(const RowPassArgs& a, hipStream_t s) {
    const int U = a.colpass_rows;
    const int TC = colpass_tile_cols(U);
    const size_t lds = colpass_lds_bytes(a);
    if (U == kColpassMfma) {
        // rows u < tile_rows, or (paired) u <= R/4 and their partners R/2 - u
        const int rows = a.pair_half ? (a.R >> 2) + 1 : a.tile_rows;
        const dim3 gp((a.C + TC - 1) / TC, (rows + 1023) / 1024);
        if (a.pair_half)
            hipLaunchKernelGGL((k_colpass_mx<PACKED, FULL, true>), gp, dim3(kColThreads), lds, s, a);
        else
            hipLaunchKernelGGL((k_colpass_mx<PACKED, FULL, false>), gp, dim3(kColThreads), lds, s, a);
        return;
    }
    if (a.pair_half) {
        // rows u <= R/4, U/2 per lane, each also giving row R/2 - u; the tile is rows [0, R/2]
        const int rows = (a.R >> 2) + 1, per_blk = 64 * (U / 2);
        const dim3 gp((a.C + TC - 1) / TC, (rows + per_blk - 1) / per_blk);
#define SPFFT_COLPASS_LAUNCH(u, tc) \
    if (U == u && TC == tc) hipLaunchKernelGGL((k_colpass<PACKED, FULL, u / 2, tc, true>), gp, dim3(kColThreads), lds, s, a);
        SPFFT_COLPASS_PAIR_SHAPES(SPFFT_COLPASS_LAUNCH)
#undef SPFFT_COLPASS_LAUNCH
        return;
    }
    const dim3 gp((a.C + TC - 1) / TC, (a.tile_rows + 64 * U - 1) / (64 * U));
#define SPFFT_COLPASS_LAUNCH(u, tc) \
    if (U == u && TC == tc) hipLaunchKernelGGL((k_colpass<PACKED, FULL, u, tc, false>), gp, dim3(kColThreads), lds, s, a);
    SPFFT_COLPASS_SHAPES(SPFFT_COLPASS_LAUNCH)
#undef SPFFT_COLPASS_LAUNCH
}

static hipError_t launch_conv(const RowPassArgs& a, hipStream_t s) {
    hipError_t e = hipErrorInvalidValue;
    if (a.conv_kind == 1) {
        if (launch_conv_a(a, s, &e) || launch_conv_b(a, s, &e) || launch_conv_c(a, s, &e) ||
            launch_conv_d(a, s, &e) || launch_conv_e(a, s, &e))
            return e;
        return hipErrorInvalidValue;
    }
    launch_conv_e(a, s, &e);
    return e;
}

static hipError_t launch_fft1(const RowPassArgs& a, dim3 grid, hipStream_t s, bool inverse) {
    if (a.precision == 1) return launch_fft1_mfma(a, s, inverse);
    hipError_t e = launch_fft1_group_a(a, grid, s, inverse);
    if (e == hipErrorInvalidValue) e = launch_fft1_group_b(a, grid, s, inverse);
    if (e == hipErrorInvalidValue) e = launch_fft1_group_c(a, grid, s, inverse);
    return e;
}

hipError_t launch_colpass_tile(const RowPassArgs& a, bool packed, hipStream_t s) {
    if (a.tile_rows <= 0) return hipSuccess;
    if (packed) {
        if (a.wr_full) launch_colpass_u<true, true>(a, s);
        else launch_colpass_u<true, false>(a, s);
    } else {
        if (a.wr_full) launch_colpass_u<false, true>(a, s);
        else launch_colpass_u<false, false>(a, s);
    }
    return hipGetLastError();
}

hipError_t launch_rowpass_tile(const RowPassArgs& a, bool packed, hipStream_t s, hipEvent_t* ev) {
    if (a.tile_rows <= 0) return hipSuccess;
    hipError_t e;
    if (ev && (e = hipEventRecord(ev[0], s)) != hipSuccess) return e;
    if ((e = launch_colpass_tile(a, packed, s)) != hipSuccess) return e;
    if (ev && (e = hipEventRecord(ev[1], s)) != hipSuccess) return e;
    if (a.fused) {
        // one kernel for the whole row pass; the fft1 and ifft1 stages time as empty
        if (ev && (e = hipEventRecord(ev[2], s)) != hipSuccess) return e;
        if ((e = a.rader ? launch_rader(a, s, false) : launch_fused(a, s, false)) != hipSuccess) return e;
        if (ev && (e = hipEventRecord(ev[3], s)) != hipSuccess) return e;
        if (ev && (e = hipEventRecord(ev[4], s)) != hipSuccess) return e;
        return hipSuccess;
    }
    // the four-step kernels, sub_rows rows of the tile at a time: T (sub_rows x M) is
    // written by k_fft1, rewritten by k_conv and read by k_ifft1 while it still sits in
    // the last-level cache.  With several sub-tiles the events time the three kernels
    // together (reported as the convolution stage).
    const int step = a.sub_rows > 0 ? a.sub_rows : a.tile_rows;
    const bool split = step < a.tile_rows;
    if (split && ev && (e = hipEventRecord(ev[2], s)) != hipSuccess) return e;
    for (int r0 = 0; r0 < a.tile_rows; r0 += step) {
        RowPassArgs b = a;
        b.u_base = a.u_base + r0;
        b.tile_rows = min(step, a.tile_rows - r0);
        b.X0 = a.X0 + (size_t)r0 * a.ldx;
        const dim3 g1((b.M2 + 63) / 64, (b.tile_rows + 3) / 4);
        if ((e = launch_fft1(b, g1, s, false)) != hipSuccess) return e;
        if (!split && ev && (e = hipEventRecord(ev[2], s)) != hipSuccess) return e;
        if ((e = launch_conv(b, s)) != hipSuccess) return e;
        if (!split && ev && (e = hipEventRecord(ev[3], s)) != hipSuccess) return e;
        if ((e = launch_fft1(b, g1, s, true)) != hipSuccess) return e;
    }
    if (split && ev && (e = hipEventRecord(ev[3], s)) != hipSuccess) return e;
    if (ev && (e = hipEventRecord(ev[4], s)) != hipSuccess) return e;
    return hipSuccess;
}

}  // namespace spfft
