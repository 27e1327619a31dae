// rader_impl.h -- k_rowrader (FP32 codelets): the row pass of a width C = 2p or
// C = p, p an odd prime, by Rader's algorithm (shared by rowpass_rader.hip and the
// generated shape lists of rowpass_rader_gen_*.hip).  The algorithm:
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
// An odd prime width C = p (NS = 1) is the same with one row group: x_0 = the row,
// X0 position n2 + M2 n1 holds column g^((M2 n1 + M1 n2) mod L), position L column 0,
// and the workgroup writes row u: X[u][k] = E_0[k] and row R - u: conj(E_0[(p - k) mod p])
// for k <= (p - 1) / 2.
//
// Replaces, for these widths, the Bluestein row pass of the reference
// (src/sparse_fft_bluestein.cu:2327-2370); the output is the same half spectrum.
#pragma once
#include "fused_impl.h"

namespace spfft {

namespace {


// LDS entries (float2) of the row buffer: the NS row groups of the convolution, or
// afterwards E_0 (and E_1) in natural order (NS p = NS (L + 1) entries)
constexpr int rader_buf(int m1, int m2, int lp, int ns = 2) {
    return ns * m1 * lp > ns * (m1 * m2 + 1) ? ns * m1 * lp : ns * (m1 * m2 + 1);
}
constexpr int rader_gcd(int a, int b) { return b ? rader_gcd(b, a % b) : a; }
constexpr size_t rader_lds_bytes(int m1, int nt, int r1, int r2, int r3, int pad, bool tp, int ns = 2) {
    const int m2 = r1 * r2 * r3;
    const int lp = pad == 2 ? m2 + m2 / 8 : pad ? m2 + m2 / 32 : m2;
    const int tabs = tp ? fused_tp_a(r1, r2, r3) + fused_tp_b(r1, r2, r3) : 0;
    // W_M2, the r-major tables, the row buffer, the partial sums
    return (size_t)(m2 + tabs + rader_buf(m1, m2, lp, ns) + ns * (nt / 64)) * sizeof(float2);
}
// Row padding: 1 (entry i at i + i/32) where every stage keeps its compile-time
// offsets, else 2 (i + i/8: radix-8 stages write at a stride of 9 entries instead of
// 8, some accesses compute their index per element) for rows of a multiple of 8
// entries with a radix-8 stage, else none; only if the padded rows fit.
constexpr int rader_pad(int m1, int nt, int r1, int r2, int r3, int ns = 2) {
    const int m2 = r1 * r2 * r3;
    if (fused_pad_ok(r1, r2, r3) && rader_lds_bytes(m1, nt, r1, r2, r3, 1, false, ns) <= 160 * 1024) return 1;
    if (m2 % 8 == 0 && (r1 == 8 || r2 == 8 || r3 == 8) &&
        rader_lds_bytes(m1, nt, r1, r2, r3, 2, false, ns) <= 160 * 1024)
        return 2;
    return 0;
}
constexpr bool rader_tp(int m1, int nt, int r1, int r2, int r3, int ns = 2) {
    return fused_tp_a(r1, r2, r3) > 0 &&
           rader_lds_bytes(m1, nt, r1, r2, r3, rader_pad(m1, nt, r1, r2, r3, ns), true, ns) <= 160 * 1024;
}
constexpr size_t rader_lds_bytes_of(int m1, int nt, int r1, int r2, int r3, int ns = 2) {
    return rader_lds_bytes(m1, nt, r1, r2, r3, rader_pad(m1, nt, r1, r2, r3, ns), rader_tp(m1, nt, r1, r2, r3, ns),
                           ns);
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
constexpr bool rader_pf(int m1, int nt, int r1, int r2, int r3, int ns = 2) {
    return rader_lds_bytes_of(m1, nt, r1, r2, r3, ns) + rader_pf_bytes(nt) <= 160 * 1024;
}
constexpr size_t rader_launch_lds(int m1, int nt, int r1, int r2, int r3, int ns = 2) {
    return rader_lds_bytes_of(m1, nt, r1, r2, r3, ns) + (rader_pf(m1, nt, r1, r2, r3, ns) ? rader_pf_bytes(nt) : 0);
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

// NS = 2: the 2 x p combination from E_0 | E_1 (natural order in buf): row u and its
// mirror row R - u, with the sign (-1)^k of the prime-factor split.  NS = 1 (C = p):
// row u is E_0[k], k <= (p - 1) / 2, and the mirror row conj(E_0[(p - k) mod p]).
template <int NT, int NS = 2>
__device__ __forceinline__ void rader_combine(const RowPassArgs& a, const int bx, const float2* buf) {
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    const int p = a.rader_p;
    const int u = a.u_base + bx;
    const bool mirror = u > 0 && a.R - u > u;
    f32x2* o = reinterpret_cast<f32x2*>(a.out + (size_t)u * a.P);
    f32x2* om = reinterpret_cast<f32x2*>(mirror ? a.out_m + (size_t)(a.R - u) * a.P : a.out);
    if constexpr (NS == 1) {
        const int V = (p + 1) / 2;
        for (int k = threadIdx.x; k < V; k += NT) {
            const float2 e = buf[k];
            __builtin_nontemporal_store((f32x2){e.x, e.y}, o + k);
            if (mirror) {
                const float2 f = buf[k == 0 ? 0 : p - k];
                __builtin_nontemporal_store((f32x2){f.x, -f.y}, om + k);
            }
        }
        return;
    }
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

// NS = 2: width C = 2p, two row groups (the even and the odd columns); NS = 1: an odd
// prime width C = p, one row group (x_0 = the row), X0 position n2 + M2 n1 holds the
// column of a_0 at (n1, n2) and position L holds column 0.
template <int M1, int NT, int R1, int R2, int R3, bool PF, int NS = 2>
__device__ __forceinline__ void rowrader_body(const RowPassArgs& a, const int bx) {
    constexpr int M2 = R1 * R2 * R3;
    constexpr int L = M1 * M2;  // p - 1
    constexpr int KB = M1;      // one batch: the M1 rows k1 of each row group
    constexpr int G = NS * KB;  // row group s = 0 (even columns), 1 (odd columns)
    constexpr int NW = NT / 64;
    static_assert(NS == 1 || NS == 2, "one or two row groups");
    static_assert(NT % 64 == 0 && NS * M2 <= NT, "one column n2 of one row group per thread");
    static_assert(R1 > 1 && (R2 > 1 || R3 == 1), "stage radices first");
    static_assert(rader_gcd(M1, M2) == 1, "the prime-factor split L = M1 x M2 needs coprime factors");
    constexpr int P = rader_pad(M1, NT, R1, R2, R3, NS);
    constexpr int LP = kRowPitch<M2, P>;
    constexpr int TP = rader_tp(M1, NT, R1, R2, R3, NS) ? 1 : 0;
    constexpr int RA = fused_tp_radix(R2, R3), NA = TP ? fused_tp_a(R1, R2, R3) : 0;
    constexpr int NBT = TP ? fused_tp_b(R1, R2, R3) : 0;
    constexpr int NBUF = rader_buf(M1, M2, LP, NS);
    extern __shared__ float2 smem[];
    float2* s_tw = smem;                    // W_M2^i, i < M2
    float2* s_ta = s_tw + M2;               // TP: r-major tables of the pivot stage ...
    float2* s_tb = NBT ? s_ta + NA : s_ta;  // ... and of the last inverse stage
    float2* buf = s_ta + NA + NBT;          // [s][k1][n2] rows, then E_0 | E_1
    float2* s_red = buf + NBUF;             // [wave][s]: partial sums of a_s
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int s = t / M2, n2 = t - s * M2;  // row group and column (t < NS M2)
    const bool live = t < NS * M2;
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
    if constexpr (PF) rader_prefetch<NW, kRaderAhead>(a, bx, wave, lane, reinterpret_cast<float*>(s_red + NS * NW));
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
            if constexpr (NS == 2) {
                p1.x += __shfl_xor(p1.x, off, 64);
                p1.y += __shfl_xor(p1.y, off, 64);
            }
        }
        if (lane == 0) {
            s_red[NS * wave] = p0;
            if constexpr (NS == 2) s_red[2 * wave + 1] = p1;
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
        x0 = bio::load2(r_x, (NS * L + s) * 8, 0);
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
            for (int w = 0; w < NW; w++) sum = cadd(sum, s_red[NS * w + s]);
            e[0] = sum;
        }
    }
    lds_barrier();
    rader_combine<NT, NS>(a, bx, buf);
    // the prefetch writes this workgroup's LDS: it lands before the workgroup ends
    if constexpr (PF) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}


template <int M1, int NT, int R1, int R2, int R3, int NS = 2>
__global__ __launch_bounds__(NT) void k_rowrader(RowPassArgs a) {
    rowrader_body<M1, NT, R1, R2, R3, rader_pf(M1, NT, R1, R2, R3, NS), NS>(a, blockIdx.x);
}

// Batched form: work item (matrix descriptor, row of its tile) per workgroup.
template <int M1, int NT, int R1, int R2, int R3, int NS = 2>
__global__ __launch_bounds__(NT) void k_rowrader_batch(const RowPassArgs* __restrict__ descs,
                                                       const int2* __restrict__ work) {
    const int2 w = work[blockIdx.x];
    rowrader_body<M1, NT, R1, R2, R3, false, NS>(descs[w.x], w.y);
}

// Launch of one FP32 shape (NS row groups) for a plan or a batch group; prepare_only
// raises both kernels' dynamic LDS limit instead.
template <int M1, int NT, int R1, int R2, int R3, int NS>
hipError_t launch_rader_fp32(const RowPassArgs& a, hipStream_t s, bool prepare_only) {
    constexpr size_t lds = rader_lds_bytes_of(M1, NT, R1, R2, R3, NS);
    constexpr size_t lds_pf = rader_launch_lds(M1, NT, R1, R2, R3, NS);
    const void* k = reinterpret_cast<const void*>(&k_rowrader<M1, NT, R1, R2, R3, NS>);
    if (prepare_only) {
        const hipError_t e0 = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rowrader_batch<M1, NT, R1, R2, R3, NS>),
                                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e0 != hipSuccess) return e0;
        return hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_pf);
    }
    hipLaunchKernelGGL((k_rowrader<M1, NT, R1, R2, R3, NS>), dim3(a.tile_rows), dim3(NT), lds_pf, s, a);
    return hipGetLastError();
}

template <int M1, int NT, int R1, int R2, int R3, int NS>
hipError_t launch_rader_fp32_batch(const RowPassArgs* descs, const int2* work, int n, hipStream_t s) {
    hipLaunchKernelGGL((k_rowrader_batch<M1, NT, R1, R2, R3, NS>), dim3(n), dim3(NT),
                       rader_lds_bytes_of(M1, NT, R1, R2, R3, NS), s, descs, work);
    return hipGetLastError();
}

// Does the plan or batch representative a run the FP32 shape (M1, NT, R1, R2, R3, NS)?
__host__ inline bool rader_shape_is(const RowPassArgs& a, int m1, int nt, int r1, int r2, int r3, int ns) {
    return a.precision == 0 && a.M1 == m1 && a.fused_nt == nt && a.f2.radix[0] == r1 &&
           (a.f2.nstages > 1 ? a.f2.radix[1] : 1) == r2 && (a.f2.nstages > 2 ? a.f2.radix[2] : 1) == r3 &&
           (a.C == 2 * a.rader_p ? 2 : 1) == ns;
}

}  // namespace

}  // namespace spfft
