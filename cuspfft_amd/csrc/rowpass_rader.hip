// rowpass_rader.hip -- k_rowrader: the row pass of an even width C = 2p, p an odd
// prime, by Rader's algorithm instead of a Bluestein convolution.
//
// The row DFT of length 2p is a 2 x p Good-Thomas (prime-factor) transform: with
// x_s[j] = X0[u][(p s + 2 j) mod 2p] (s = 0, 1; the even columns, then the odd ones in
// the order p, p + 2, ...), and E_s = DFT_p(x_s),
//   Y[k] = E_0[k mod p] + (-1)^k E_1[k mod p],
// with no twiddle factors.  Each length-p DFT of a prime p is, with a primitive root
// g of p, a cyclic convolution of length L = p - 1 (Rader):
//   a_s[q] = x_s[g^q],  b[q] = W_p^{g^-q}   (q < L)
//   E_s[g^-m] = x_s[0] + (a_s (*) b)[m],    E_s[0] = x_s[0] + sum_q a_s[q].
// The convolution runs as DFT_L, a product with B^ = DFT_L(b) / L, and the inverse
// DFT_L, each again a prime-factor transform L = M1 x M2 (gcd(M1, M2) = 1): entry
// (n1, n2) is a_s[(M2 n1 + M1 n2) mod L] and frequency (k1, k2) is
// (M2 (M2^-1 mod M1) k1 + M1 (M1^-1 mod M2) k2) mod L, so that
//   DFT_L = DFT_M1 over n1 (x) DFT_M2 over n2
// with no twiddles between the two.  The column pass writes X0 in this order
// (position s L + n2 + M2 n1 holds the column of a_s at (n1, n2); positions 2L and
// 2L + 1 hold columns 0 and p), the host stores B^ and the scatter indices g^-m in
// it, so the kernel only sees contiguous runs.  One workgroup per output row u runs
// both convolutions on chip as the two row groups of k_rowfused's split (phase A:
// length-M1 register codelet; phase B: the Stockham stages of length M2 through LDS
// with B^ at the pivot; phase C: the inverse codelet), then scatters E_s to natural
// order in LDS and writes
//   row u:      X[u][k] = E_0[k] + (-1)^k E_1[k]      (k < p),  X[u][p] = E_0[0] - E_1[0]
//   row R - u:  X[R-u][p - k] = conj(E_0[k] - (-1)^k E_1[k]),   X[R-u][0] = conj(X[u][0])
// For c3's width 16706 = 2 * 8353 the convolutions have length 8352 = 29 * 288
// (two of them per row) against one of length 34992 for Bluestein.
//
// Replaces, for these widths, the Bluestein row pass of the reference
// (src/sparse_fft_bluestein.cu:2327-2370); the output is the same half spectrum.
#include "fused_impl.h"

namespace spfft {

namespace {

// LDS entries (float2) of the row buffer: the two row groups of the convolution,
// or afterwards E_0 and E_1 in natural order (2p = 2L + 2 entries)
constexpr int rader_buf(int m1, int m2, int lp) { return 2 * m1 * lp > 2 * m1 * m2 + 2 ? 2 * m1 * lp : 2 * m1 * m2 + 2; }
constexpr int rader_gcd(int a, int b) { return b ? rader_gcd(b, a % b) : a; }
constexpr size_t rader_lds_bytes(int m1, int nt, int r1, int r2, int r3, int pad, bool tp) {
    const int m2 = r1 * r2 * r3, L = m1 * m2;
    const int lp = pad == 2 ? m2 + m2 / 8 : pad ? m2 + m2 / 32 : m2;
    const int tabs = tp ? fused_tp_a(r1, r2, r3) + fused_tp_b(r1, r2, r3) : 0;
    (void)L;
    // W_M2, the r-major tables, the row buffer, the partial sums
    return (size_t)(m2 + tabs + rader_buf(m1, m2, lp) + 2 * (nt / 64)) * sizeof(float2);
}
// Row padding: 1 (entry i at i + i/32) where every stage keeps its compile-time
// offsets, else 2 (i + i/8: radix-8 stages write at a stride of 9 entries instead of
// 8, some accesses compute their index per element) for rows of a multiple of 8
// entries with a radix-8 stage, else none; only if the padded rows fit.
constexpr int rader_pad(int m1, int nt, int r1, int r2, int r3) {
    const int m2 = r1 * r2 * r3;
    if (fused_pad_ok(r1, r2, r3) && rader_lds_bytes(m1, nt, r1, r2, r3, 1, false) <= 160 * 1024) return 1;
    if (m2 % 8 == 0 && (r1 == 8 || r2 == 8 || r3 == 8) && rader_lds_bytes(m1, nt, r1, r2, r3, 2, false) <= 160 * 1024)
        return 2;
    return 0;
}
constexpr bool rader_tp(int m1, int nt, int r1, int r2, int r3) {
    return fused_tp_a(r1, r2, r3) > 0 &&
           rader_lds_bytes(m1, nt, r1, r2, r3, rader_pad(m1, nt, r1, r2, r3), true) <= 160 * 1024;
}
constexpr size_t rader_lds_bytes_of(int m1, int nt, int r1, int r2, int r3) {
    return rader_lds_bytes(m1, nt, r1, r2, r3, rader_pad(m1, nt, r1, r2, r3), rader_tp(m1, nt, r1, r2, r3));
}
// L2 prefetch of the pass-1 row that the workgroup kRaderAhead places later takes (the
// same XCD: workgroups go to the 8 XCDs in turn), by direct-to-LDS loads of one 4-byte
// word per 128-byte line into a scratch of 64 words that all waves share (its contents
// are never read), when that fits in LDS.  A workgroup's row loads run at the rate its
// CU can keep lines in flight (about 17 GB/s per CU from HBM); from L2 they arrive
// sooner.  No registers wait for these loads, and they are issued after the row's last
// global load and barrier, so only the final wait of the workgroup includes them.
// The distance keeps the prefetched rows of an XCD (kRaderAhead / 8 rows of 134 KB on
// c3) well inside its 4 MB L2: c3 row kernel 32: 0.976 ms, 48: 0.908, 64: 0.911,
// 80: 0.933, 128: 0.932, 256: 0.946 (at 256 a fifth more lines are fetched twice).
// k_rowrader_mfma issues its prefetch late, after phase C (kRaderAheadMfma).
#ifndef SPFFT_RADER_AHEAD
#define SPFFT_RADER_AHEAD 56
#endif
#ifndef SPFFT_RADER_AHEAD_MFMA
#define SPFFT_RADER_AHEAD_MFMA 256
#endif
constexpr int kRaderAhead = SPFFT_RADER_AHEAD;
constexpr int kRaderAheadMfma = SPFFT_RADER_AHEAD_MFMA;
constexpr size_t rader_pf_bytes(int) { return 64 * sizeof(float); }
constexpr bool rader_pf(int m1, int nt, int r1, int r2, int r3) {
    return rader_lds_bytes_of(m1, nt, r1, r2, r3) + rader_pf_bytes(nt) <= 160 * 1024;
}
constexpr size_t rader_launch_lds(int m1, int nt, int r1, int r2, int r3) {
    return rader_lds_bytes_of(m1, nt, r1, r2, r3) + (rader_pf(m1, nt, r1, r2, r3) ? rader_pf_bytes(nt) : 0);
}

// The L2 prefetch of pass-1 row bx + AHEAD of the tile.
template <int NW, int AHEAD>
__device__ __forceinline__ void rader_prefetch(const RowPassArgs& a, int bx, int wave, int lane, float* scr) {
    const int nb = bx + AHEAD;
    if (nb >= a.tile_rows) return;
    const char* rowp = reinterpret_cast<const char*>(a.X0 + (size_t)nb * a.ldx);
    const int lines = (a.C * 8 + 127) >> 7;
    for (int i = __builtin_amdgcn_readfirstlane(wave); i * 64 < lines; i += NW) {
        const int line = i * 64 + lane;
        if (line < lines)
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(rowp + (size_t)line * 128),
                                             (__attribute__((address_space(3))) void*)scr, 4, 0, 0);
    }
}

// The 2 x p combination from E_0 | E_1 (natural order in buf): row u and its
// mirror row R - u, with the sign (-1)^k of the prime-factor split.
template <int NT>
__device__ __forceinline__ void rader_combine(const RowPassArgs& a, const int bx, const float2* buf) {
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    const int p = a.rader_p;
    const int u = a.u_base + bx;
    const bool mirror = u > 0 && a.R - u > u;
    f32x2* o = reinterpret_cast<f32x2*>(a.out + (size_t)u * a.P);
    f32x2* om = reinterpret_cast<f32x2*>(mirror ? a.out_m + (size_t)(a.R - u) * a.P : a.out);
    for (int k = threadIdx.x; k < p; k += NT) {
        const float2 e0 = buf[k], f1 = buf[p + k];
        const float2 e1 = (k & 1) ? make_float2(-f1.x, -f1.y) : f1;
        const f32x2 y0 = {e0.x + e1.x, e0.y + e1.y};
        const f32x2 y1 = {e0.x - e1.x, e0.y - e1.y};
        __builtin_nontemporal_store(y0, o + k);
        if (k == 0) __builtin_nontemporal_store(y1, o + p);
        if (mirror) {
            const f32x2 c1 = {y1[0], -y1[1]};
            __builtin_nontemporal_store(c1, om + (p - k));
            if (k == 0) {
                const f32x2 c0 = {y0[0], -y0[1]};
                __builtin_nontemporal_store(c0, om);
            }
        }
    }
}

template <int M1, int NT, int R1, int R2, int R3, bool PF>
__device__ __forceinline__ void rowrader_body(const RowPassArgs& a, const int bx) {
    constexpr int M2 = R1 * R2 * R3;
    constexpr int L = M1 * M2;  // p - 1
    constexpr int KB = M1;      // one batch: the M1 rows k1 of each row group
    constexpr int G = 2 * KB;   // row group s = 0 (even columns), 1 (odd columns)
    constexpr int NW = NT / 64;
    static_assert(NT % 64 == 0 && 2 * M2 <= NT, "one column n2 of one row group per thread");
    static_assert(R1 > 1 && (R2 > 1 || R3 == 1), "stage radices first");
    static_assert(rader_gcd(M1, M2) == 1, "the prime-factor split L = M1 x M2 needs coprime factors");
    constexpr int P = rader_pad(M1, NT, R1, R2, R3);
    constexpr int LP = kRowPitch<M2, P>;
    constexpr int TP = rader_tp(M1, NT, R1, R2, R3) ? 1 : 0;
    constexpr int RA = fused_tp_radix(R2, R3), NA = TP ? fused_tp_a(R1, R2, R3) : 0;
    constexpr int NBT = TP ? fused_tp_b(R1, R2, R3) : 0;
    constexpr int NBUF = rader_buf(M1, M2, LP);
    extern __shared__ float2 smem[];
    float2* s_tw = smem;                    // W_M2^i, i < M2
    float2* s_ta = s_tw + M2;               // TP: r-major tables of the pivot stage ...
    float2* s_tb = NBT ? s_ta + NA : s_ta;  // ... and of the last inverse stage
    float2* buf = s_ta + NA + NBT;          // [s][k1][n2] rows, then E_0 | E_1
    float2* s_red = buf + NBUF;             // [wave][s]: partial sums of a_s
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int s = t / M2, n2 = t - s * M2;  // row group and column (t < 2 M2)
    const bool live = t < 2 * M2;
    const int p = a.rader_p;
    const bio::rsrc_t r_bh = bio::make_rsrc(a.bhat, (uint32_t)L * 8u);

    // phase A: the entries (n1, n2), n1 < M1, of a_s; the tables while they load
    const bio::rsrc_t r_x = bio::make_rsrc(a.X0 + (size_t)bx * a.ldx, (uint32_t)a.C * 8u);
    float2 v[M1];
    if (live) {
#pragma unroll
        for (int n1 = 0; n1 < M1; n1++) v[n1] = bio::load2_stream(r_x, (s * L + n2) * 8, M2 * n1 * 8);
    }
    for (int i = t; i < M2; i += NT) s_tw[i] = a.tw2[i];
    if constexpr (TP) {
        constexpr int nsa = NA ? M2 / RA : 1, nsb = NBT ? M2 / R1 : 1;
        for (int i = t; i < NA; i += NT) s_ta[i] = a.tw2[(i / nsa + 1) * (i % nsa)];
        for (int i = t; i < NBT; i += NT) s_tb[i] = a.tw2[(i / nsb + 1) * (i % nsb)];
    }
    lds_barrier();
    // after the row's own loads (whose waits it does not join); the barriers below are
    // LDS-only, so it stays in flight until the pivot stage's loads of B^ wait for it
    if constexpr (PF) rader_prefetch<NW, kRaderAhead>(a, bx, wave, lane, reinterpret_cast<float*>(s_red + 2 * NW));
    if (live) RegFft<M1>::run(v, false);
    {
        // sum_q a_s[q] = sum over the columns of the codelet's output 0: per wave and row
        // group in a fixed order (no atomics), added up in phase C
        float2 p0 = live && s == 0 ? v[0] : make_float2(0.f, 0.f);
        float2 p1 = live && s == 1 ? v[0] : make_float2(0.f, 0.f);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            p0.x += __shfl_xor(p0.x, off, 64);
            p0.y += __shfl_xor(p0.y, off, 64);
            p1.x += __shfl_xor(p1.x, off, 64);
            p1.y += __shfl_xor(p1.y, off, 64);
        }
        if (lane == 0) {
            s_red[2 * wave] = p0;
            s_red[2 * wave + 1] = p1;
        }
    }

    // phase B (no twiddles between the length-M1 and length-M2 DFTs: prime-factor split): both row groups' rows k1 into LDS, the convolution with B^
    if (live) {
        float2* col = buf + s * KB * LP + lds_idx<P>(n2);
#pragma unroll
        for (int k = 0; k < KB; k++) col[k * LP] = v[k];
    }
    lds_barrier();
    fused_conv<M2, G, KB, NT, R1, R2, R3, P, TP>(r_bh, buf, s_tw, 0, s_ta, s_tb);

    // phase C: the inverse length-M1 codelet gives x_s[0] + c_s[m], m = (M2 n1 + M1 n2)
    // mod L (x_s[0] added to the codelet's input 0 reaches each of its outputs)
    int kk[M1];  // g^-m mod p
    float2 x0 = make_float2(0.f, 0.f);
    if (live) {
        x0 = bio::load2(r_x, (2 * L + s) * 8, 0);
        const float2* col = buf + s * KB * LP + lds_idx<P>(n2);
#pragma unroll
        for (int k = 0; k < KB; k++) v[k] = col[k * LP];
        v[0] = cadd(v[0], x0);
        RegFft<M1>::run(v, true);
#pragma unroll
        for (int n1 = 0; n1 < M1; n1++) kk[n1] = a.rader_k[n2 + M2 * n1];
    }
    lds_barrier();  // the rows are read: buf now takes E_0 | E_1 in natural order
    if (live) {
        float2* e = buf + s * p;
#pragma unroll
        for (int n1 = 0; n1 < M1; n1++) e[kk[n1]] = v[n1];
        if (n2 == 0) {
            float2 sum = x0;
            for (int w = 0; w < NW; w++) sum = cadd(sum, s_red[2 * w + s]);
            e[0] = sum;
        }
    }
    lds_barrier();
    rader_combine<NT>(a, bx, buf);
    // the prefetch writes this workgroup's LDS: it lands before the workgroup ends
    if constexpr (PF) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}


template <int M1, int NT, int R1, int R2, int R3>
__global__ __launch_bounds__(NT) void k_rowrader(RowPassArgs a) {
    rowrader_body<M1, NT, R1, R2, R3, rader_pf(M1, NT, R1, R2, R3)>(a, blockIdx.x);
}

// Batched form: work item (matrix descriptor, row of its tile) per workgroup.
template <int M1, int NT, int R1, int R2, int R3>
__global__ __launch_bounds__(NT) void k_rowrader_batch(const RowPassArgs* __restrict__ descs,
                                                       const int2* __restrict__ work) {
    const int2 w = work[blockIdx.x];
    rowrader_body<M1, NT, R1, R2, R3, false>(descs[w.x], w.y);
}

// ---- precision form: the length-M1 DFTs on the matrix cores ------------------
//
// k_rowrader_mfma is k_rowrader with phase A's forward and phase C's inverse
// length-M1 DFTs (M1 <= 32, e.g. 29: the codelet that costs O(M1^2) VALU
// operations) done as complex DFT-matrix products on the matrix cores.  With
// W_M1^{kn} = C[k][n] - i S[k][n] (zero rows and columns from M1 to 32), the DFT of
// 32 columns x (32 x 32, n1 x column) is
//     forward  Yr = C Xr + S Xi,   Yi = C Xi - S Xr
//     inverse  Yr = C Xr - S Xi,   Yi = C Xi + S Xr
// four real 32 x 32 x 32 products, each two v_mfma_f32_32x32x16_bf16 k-steps.  Every
// operand is split into BF16 parts v = vh + vl (vh = RNE(v), vl = RNE(v - vh), about
// 16 significant bits together) and A B ~= Ah Bh + Ah Bl + Al Bh accumulates in FP32
// (the dropped Al Bl term is ~2^-18 relative).  C and S are split on the host from
// FP64 values.  A wave takes 32 columns q = s M2 + n2 of both row groups at a time
// (tile), lane l holding column q = 32 tile + l % 32 and, per k-step, the 8 entries
// k = 16 ks + 8 (l / 32) + j; the product leaves output rows 8 g + 4 (l / 32) + i in
// accumulator element 4 g + i (operand and accumulator layouts of gfx950).

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// The lane's A-operand fragments of C and S: [part hi / lo][k-step].
struct DftFrags {
    bf16x8 c[2][2], s[2][2];
};

// Table layout (host-built, spfft_plan.cpp): [matrix C, S][part][k-step][lane], 8 BF16 each.
__device__ __forceinline__ void load_dft_frags(const bf16x8* tab, int lane, DftFrags& f) {
#pragma unroll
    for (int part = 0; part < 2; part++)
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            f.c[part][ks] = tab[((0 * 2 + part) * 2 + ks) * 64 + lane];
            f.s[part][ks] = tab[((1 * 2 + part) * 2 + ks) * 64 + lane];
        }
}

// x = hi + lo in BF16 (round to nearest even both times).
__device__ __forceinline__ void split_bf16x8(const float* x, bf16x8& hi, bf16x8& lo) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const __bf16 h = (__bf16)x[j];
        hi[j] = h;
        lo[j] = (__bf16)(x[j] - (float)h);
    }
}

__device__ __forceinline__ bf16x8 neg_bf16x8(bf16x8 v) {
    u32x4 u = __builtin_bit_cast(u32x4, v);
    u ^= 0x80008000u;
    return __builtin_bit_cast(bf16x8, u);
}

// The lane's share of its column: BF16 parts of the real and imaginary entries.
struct ColFrags {
    bf16x8 r[2][2], i[2][2];  // [part][k-step]
};

__device__ __forceinline__ void split_col(const float2* x, ColFrags& c) {
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
        float re[8], im[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            re[j] = x[8 * ks + j].x;
            im[j] = x[8 * ks + j].y;
        }
        split_bf16x8(re, c.r[0][ks], c.r[1][ks]);
        split_bf16x8(im, c.i[0][ks], c.i[1][ks]);
    }
}

#define SPFFT_MFMA3(acc, A, B)                                                        \
    do {                                                                              \
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[0][ks], B[0][ks], acc, 0, 0, 0); \
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[0][ks], B[1][ks], acc, 0, 0, 0); \
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[1][ks], B[0][ks], acc, 0, 0, 0); \
    } while (0)

// y = F x (forward) or conj(F) x (inverse) for the wave's 32 columns.
template <bool INV>
__device__ __forceinline__ void dft32_mfma(const DftFrags& f, const ColFrags& x, floatx16& yr, floatx16& yi) {
#pragma unroll
    for (int e = 0; e < 16; e++) yr[e] = yi[e] = 0.f;
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
        // the S term's operand with its sign: forward -Xr (into Yi), inverse -Xi (into Yr)
        bf16x8 n[2][2];
        n[0][ks] = neg_bf16x8(INV ? x.i[0][ks] : x.r[0][ks]);
        n[1][ks] = neg_bf16x8(INV ? x.i[1][ks] : x.r[1][ks]);
        SPFFT_MFMA3(yr, f.c, x.r);
        SPFFT_MFMA3(yi, f.c, x.i);
        if (INV) {
            SPFFT_MFMA3(yr, f.s, n);
            SPFFT_MFMA3(yi, f.s, x.r);
        } else {
            SPFFT_MFMA3(yr, f.s, x.i);
            SPFFT_MFMA3(yi, f.s, n);
        }
    }
}
#undef SPFFT_MFMA3

template <int M1, int NT, int R1, int R2, int R3, bool PF>
__device__ __forceinline__ void rowrader_mfma_body(const RowPassArgs& a, const int bx) {
    constexpr int M2 = R1 * R2 * R3;
    constexpr int L = M1 * M2;  // p - 1
    constexpr int KB = M1;
    constexpr int G = 2 * KB;
    constexpr int NW = NT / 64;
    constexpr int NQ = 2 * M2;              // columns of both row groups
    constexpr int NTILE = (NQ + 31) / 32;   // 32-column tiles
    constexpr int TPW = (NTILE + NW - 1) / NW;  // tiles per wave (at most)
    static_assert(M1 <= 32 && M1 > 16, "one 32-point DFT matrix, two k-steps");
    static_assert(NT % 64 == 0, "whole waves");
    static_assert(rader_gcd(M1, M2) == 1, "the prime-factor split L = M1 x M2 needs coprime factors");
    constexpr int P = rader_pad(M1, NT, R1, R2, R3);
    constexpr int LP = kRowPitch<M2, P>;
    constexpr int TP = rader_tp(M1, NT, R1, R2, R3) ? 1 : 0;
    constexpr int RA = fused_tp_radix(R2, R3), NA = TP ? fused_tp_a(R1, R2, R3) : 0;
    constexpr int NBT = TP ? fused_tp_b(R1, R2, R3) : 0;
    constexpr int NBUF = rader_buf(M1, M2, LP);
    extern __shared__ float2 smem[];
    float2* s_tw = smem;
    float2* s_ta = s_tw + M2;
    float2* s_tb = NBT ? s_ta + NA : s_ta;
    float2* buf = s_ta + NA + NBT;
    float2* s_red = buf + NBUF;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6, h = lane >> 5, jl = lane & 31;
    const int p = a.rader_p;
    const bio::rsrc_t r_bh = bio::make_rsrc(a.bhat, (uint32_t)L * 8u);
    const bio::rsrc_t r_x = bio::make_rsrc(a.X0 + (size_t)bx * a.ldx, (uint32_t)a.C * 8u);
    const bf16x8* frag_tab = reinterpret_cast<const bf16x8*>(a.mfma_frag);

    // phase A: the wave's tiles of pass-1 entries (rows n1 of its columns), the tables
    float2 xa[TPW][16];
#pragma unroll
    for (int k = 0; k < TPW; k++) {
        const int tile = wave + k * NW;
        const int q = tile * 32 + jl;
        const bool lv = tile < NTILE && q < NQ;
        const int s = q >= M2 ? 1 : 0, n2 = q - s * M2;
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const int n1 = 16 * (e >> 3) + 8 * h + (e & 7);
            xa[k][e] = bio::load2_stream(r_x, lv && n1 < M1 ? (s * L + n2) * 8 : bio::kOob, M2 * n1 * 8);
        }
    }
    for (int i = t; i < M2; i += NT) s_tw[i] = a.tw2[i];
    if constexpr (TP) {
        constexpr int nsa = NA ? M2 / RA : 1, nsb = NBT ? M2 / R1 : 1;
        for (int i = t; i < NA; i += NT) s_ta[i] = a.tw2[(i / nsa + 1) * (i % nsa)];
        for (int i = t; i < NBT; i += NT) s_tb[i] = a.tw2[(i / nsb + 1) * (i % nsb)];
    }
    DftFrags f;
    load_dft_frags(frag_tab, lane, f);
    lds_barrier();
    float2 psum0 = make_float2(0.f, 0.f), psum1 = make_float2(0.f, 0.f);
#pragma unroll
    for (int k = 0; k < TPW; k++) {
        const int tile = wave + k * NW;
        if (tile >= NTILE) break;  // wave-uniform
        const int q = tile * 32 + jl;
        const bool lv = q < NQ;
        const int s = q >= M2 ? 1 : 0, n2 = q - s * M2;
        ColFrags xc;
        split_col(xa[k], xc);
        floatx16 yr, yi;
        dft32_mfma<false>(f, xc, yr, yi);
        // sum_q a_s[q] = sum over the columns of output row 0 (lanes h = 0, element 0)
        if (lv && h == 0) {
            const float2 y0 = make_float2(yr[0], yi[0]);
            if (s == 0) psum0 = cadd(psum0, y0);
            else psum1 = cadd(psum1, y0);
        }
        // rows k1 = 4 (2 g + h) + i (no twiddles: prime-factor split)
        float2* col = buf + s * KB * LP + lds_idx<P>(n2);
#pragma unroll
        for (int g = 0; g < 4; g++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int k1 = 8 * g + 4 * h + i;
                if (lv && k1 < M1) col[k1 * LP] = make_float2(yr[4 * g + i], yi[4 * g + i]);
            }
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        psum0.x += __shfl_xor(psum0.x, off, 64);
        psum0.y += __shfl_xor(psum0.y, off, 64);
        psum1.x += __shfl_xor(psum1.x, off, 64);
        psum1.y += __shfl_xor(psum1.y, off, 64);
    }
    if (lane == 0) {
        s_red[2 * wave] = psum0;
        s_red[2 * wave + 1] = psum1;
    }
    lds_barrier();

    // phase B: the convolution of both row groups with B^
    fused_conv<M2, G, KB, NT, R1, R2, R3, P, TP>(r_bh, buf, s_tw, 0, s_ta, s_tb);

    // phase C: the wave's columns back from LDS, split into BF16 parts (held across
    // the barrier: buf then takes E_0 | E_1 in natural order)
    ColFrags xcs[TPW];
#pragma unroll
    for (int k = 0; k < TPW; k++) {
        const int tile = wave + k * NW;
        const int q = tile * 32 + jl;
        const bool lv = tile < NTILE && q < NQ;
        const int s = q >= M2 ? 1 : 0, n2 = q - s * M2;
        float2 xc[16];
        if (tile < NTILE) {
            // rows k1 = 8 (2 ks + h) + j
            const float2* col = buf + s * KB * LP + lds_idx<P>(n2);
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int ks = e >> 3, j = e & 7;
                const int k1 = 16 * ks + 8 * h + j;
                xc[e] = lv && k1 < M1 ? col[k1 * LP] : make_float2(0.f, 0.f);
            }
        } else {
#pragma unroll
            for (int e = 0; e < 16; e++) xc[e] = make_float2(0.f, 0.f);
        }
        split_col(xc, xcs[k]);
        __builtin_amdgcn_sched_barrier(0);  // one tile's loads in flight at a time
    }
    lds_barrier();
#pragma unroll
    for (int k = 0; k < TPW; k++) {
        const int tile = wave + k * NW;
        if (tile >= NTILE) break;  // wave-uniform
        const int q = tile * 32 + jl;
        const bool lv = q < NQ;
        const int s = q >= M2 ? 1 : 0, n2 = q - s * M2;
        const float2 x0 = bio::load2(r_x, lv ? (2 * L + s) * 8 : bio::kOob, 0);
        floatx16 yr, yi;
        dft32_mfma<true>(f, xcs[k], yr, yi);
        // c_s[m], m = (M2 m1 + M1 n2) mod L (m1 = 8 g + 4 h + i): E_s[g^-m] = x_s[0] + c_s[m]
        float2* e = buf + s * p;
#pragma unroll
        for (int g = 0; g < 4; g++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int m1 = 8 * g + 4 * h + i;
                if (lv && m1 < M1) e[a.rader_k[n2 + M2 * m1]] = cadd(x0, make_float2(yr[4 * g + i], yi[4 * g + i]));
            }
        if (lv && n2 == 0 && h == 0) {
            float2 sum = x0;
            for (int w = 0; w < NW; w++) sum = cadd(sum, s_red[2 * w + s]);
            e[0] = sum;
        }
    }
    lds_barrier();

    if constexpr (PF) rader_prefetch<NW, kRaderAheadMfma>(a, bx, wave, lane, reinterpret_cast<float*>(s_red + 2 * NW));

    rader_combine<NT>(a, bx, buf);
    if constexpr (PF) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

template <int M1, int NT, int R1, int R2, int R3>
__global__ __launch_bounds__(NT) void k_rowrader_mfma(RowPassArgs a) {
    rowrader_mfma_body<M1, NT, R1, R2, R3, rader_pf(M1, NT, R1, R2, R3)>(a, blockIdx.x);
}

template <int M1, int NT, int R1, int R2, int R3>
__global__ __launch_bounds__(NT) void k_rowrader_mfma_batch(const RowPassArgs* __restrict__ descs,
                                                            const int2* __restrict__ work) {
    const int2 w = work[blockIdx.x];
    rowrader_mfma_body<M1, NT, R1, R2, R3, false>(descs[w.x], w.y);
}

}  // namespace

// Instantiated shapes X(M1, NT, R1, R2, R3), L = p - 1 = M1 R1 R2 R3:
//   8352 = 29 * (36 * 8)   (width 16706 = 2 * 8353; 704 threads, rows padded by one
//                           entry every 8)
//   2088 = 29 * (8 * 9)    (width 4178 = 2 * 2089)
//    928 = 29 * (4 * 8)    (width 1858 = 2 * 929)
// For 8352, measured on c3 (row kernel per launch, one MI355X), with the padding where
// it applies: 29 * (36 * 8) at 704 threads 1.090 ms, at 768 1.093, at 640 1.106, at
// 576 1.118; 36 * (29 * 8) at 512 threads 1.150 (1.176 unpadded: the radix-8 pivot
// writes at a stride of 8 entries; bank-conflict cycles 9.3e7 -> 7.8e7 per launch), and
// its radix-29 stages run 576 butterflies on 512 threads, a second round for one wave;
// 36 * (8 * 29) 1.223; 29 * (32 * 9) 1.306; 29 * (18 * 16) 1.307; 29 * (16 * 18) 1.319
// (its radix-16 first stage writes at a stride of 16 entries: 1.9e8 conflict cycles).
// Radix-29 LDS stages at 576+ threads (168 VGPRs), radix 12 / 24 stages, or M1 = 32 or
// 48 spill 50-175 VGPRs.
#define SPFFT_RADER_LIST(X) X(29, 704, 36, 8, 1) X(29, 192, 8, 9, 1) X(29, 64, 4, 8, 1)
// The precision form (k_rowrader_mfma, M1 <= 32 padded to one 32-point DFT matrix).
#define SPFFT_RADER_MFMA_LIST(X) X(29, 704, 36, 8, 1) X(29, 576, 36, 8, 1) X(29, 192, 8, 9, 1) X(29, 64, 4, 8, 1)

bool rader_shape(int i, RaderShape* out) {
    static const RaderShape kShapes[] = {
#define SPFFT_SHAPE(m1, nt, r1, r2, r3) {m1, nt, {r1, r2, r3}, rader_lds_bytes_of(m1, nt, r1, r2, r3), 0},
        SPFFT_RADER_LIST(SPFFT_SHAPE)
#undef SPFFT_SHAPE
#define SPFFT_SHAPE(m1, nt, r1, r2, r3) {m1, nt, {r1, r2, r3}, rader_lds_bytes_of(m1, nt, r1, r2, r3), 1},
        SPFFT_RADER_MFMA_LIST(SPFFT_SHAPE)
#undef SPFFT_SHAPE
    };
    if (i < 0 || i >= (int)(sizeof(kShapes) / sizeof(kShapes[0]))) return false;
    *out = kShapes[i];
    return true;
}

hipError_t launch_rader_batch(const RowPassArgs& a, const RowPassArgs* descs, const int2* work, int n,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
#define SPFFT_CASE(K, prec, m1, nt, r1, r2, r3)                                                          \
    if (a.precision == prec && a.M1 == m1 && a.fused_nt == nt && a.f2.radix[0] == r1 &&                  \
        (a.f2.nstages > 1 ? a.f2.radix[1] : 1) == r2 && (a.f2.nstages > 2 ? a.f2.radix[2] : 1) == r3) {   \
        hipLaunchKernelGGL((K<m1, nt, r1, r2, r3>), dim3(n), dim3(nt), rader_lds_bytes_of(m1, nt, r1, r2, r3), \
                           s, descs, work);                                                              \
        return hipGetLastError();                                                                        \
    }
#define SPFFT_CASE_MFMA(m1, nt, r1, r2, r3) SPFFT_CASE(k_rowrader_mfma_batch, 1, m1, nt, r1, r2, r3)
#define SPFFT_CASE_FP32(m1, nt, r1, r2, r3) SPFFT_CASE(k_rowrader_batch, 0, m1, nt, r1, r2, r3)
    SPFFT_RADER_MFMA_LIST(SPFFT_CASE_MFMA)
    SPFFT_RADER_LIST(SPFFT_CASE_FP32)
#undef SPFFT_CASE_MFMA
#undef SPFFT_CASE_FP32
#undef SPFFT_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_rader(const RowPassArgs& a, hipStream_t s, bool prepare_only) {
#define SPFFT_CASE(m1, nt, r1, r2, r3)                                                                   \
    if (a.precision == 1 && a.M1 == m1 && a.fused_nt == nt && a.f2.radix[0] == r1 &&                     \
        (a.f2.nstages > 1 ? a.f2.radix[1] : 1) == r2 && (a.f2.nstages > 2 ? a.f2.radix[2] : 1) == r3) {   \
        constexpr size_t lds = rader_lds_bytes_of(m1, nt, r1, r2, r3);                                   \
        constexpr size_t lds_pf = rader_launch_lds(m1, nt, r1, r2, r3);                                  \
        const void* k = reinterpret_cast<const void*>(&k_rowrader_mfma<m1, nt, r1, r2, r3>);              \
        if (prepare_only) {                                                                              \
            const hipError_t e0 = hipFuncSetAttribute(                                                   \
                reinterpret_cast<const void*>(&k_rowrader_mfma_batch<m1, nt, r1, r2, r3>),               \
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                                   \
            if (e0 != hipSuccess) return e0;                                                             \
            return hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_pf);      \
        }                                                                                                \
        hipLaunchKernelGGL((k_rowrader_mfma<m1, nt, r1, r2, r3>), dim3(a.tile_rows), dim3(nt), lds_pf, s, a); \
        return hipGetLastError();                                                                        \
    }
    SPFFT_RADER_MFMA_LIST(SPFFT_CASE)
#undef SPFFT_CASE
#define SPFFT_CASE(m1, nt, r1, r2, r3)                                                                   \
    if (a.precision == 0 && a.M1 == m1 && a.fused_nt == nt && a.f2.radix[0] == r1 &&                     \
        (a.f2.nstages > 1 ? a.f2.radix[1] : 1) == r2 && (a.f2.nstages > 2 ? a.f2.radix[2] : 1) == r3) {   \
        constexpr size_t lds = rader_lds_bytes_of(m1, nt, r1, r2, r3);                                   \
        constexpr size_t lds_pf = rader_launch_lds(m1, nt, r1, r2, r3);                                  \
        const void* k = reinterpret_cast<const void*>(&k_rowrader<m1, nt, r1, r2, r3>);                   \
        if (prepare_only) {                                                                              \
            const hipError_t e0 = hipFuncSetAttribute(                                                   \
                reinterpret_cast<const void*>(&k_rowrader_batch<m1, nt, r1, r2, r3>),                    \
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                                   \
            if (e0 != hipSuccess) return e0;                                                             \
            return hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_pf);      \
        }                                                                                                \
        hipLaunchKernelGGL((k_rowrader<m1, nt, r1, r2, r3>), dim3(a.tile_rows), dim3(nt), lds_pf, s, a); \
        return hipGetLastError();                                                                        \
    }
    SPFFT_RADER_LIST(SPFFT_CASE)
#undef SPFFT_CASE
    return hipErrorInvalidValue;
}

}  // namespace spfft
