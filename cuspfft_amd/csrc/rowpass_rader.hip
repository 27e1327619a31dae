// rowpass_rader.hip -- k_rowrader: the row pass of an even width C = 2p, p an odd
// prime, by Rader's algorithm instead of a Bluestein convolution.
//
// The row DFT splits by column parity (x_s[j] = X0[u][2j + s], s = 0, 1):
//   Y[v] = E_0[v mod p] + W_C^v E_1[v mod p],   E_s = DFT_p(x_s),
// and each length-p DFT of a prime p is, with a primitive root g of p, a cyclic
// convolution of length L = p - 1 (Rader):
//   a_s[q] = x_s[g^q],  b[q] = W_p^{g^-q}   (q < L)
//   E_s[g^-m] = x_s[0] + (a_s (*) b)[m],    E_s[0] = x_s[0] + sum_q a_s[q].
// The column pass writes X0 in Rader order (position s L + q holds column
// 2 g^q + s, positions 2L and 2L + 1 hold columns 0 and 1), so a_s is one
// contiguous run of the pass-1 row.  One workgroup per output row u runs both
// convolutions on chip as the two row groups of k_rowfused's four-step split
// L = M1 M2 (phase A: length-M1 register codelet and W_L twiddles; phase B: the
// Stockham stages through LDS with B^ = DFT_L(b) / L at the pivot; phase C: the
// inverse), then scatters E_s to natural order in LDS and writes
//   row u:      X[u][k] = E_0[k] + W_C^k E_1[k]      (k < p),  X[u][p] = E_0[0] - E_1[0]
//   row R - u:  X[R-u][p - k] = conj(E_0[k] - W_C^k E_1[k]),   X[R-u][0] = conj(X[u][0])
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
// two-level tables of W_n^e, e < n: 256 low entries and ceil(n / 256) high ones
constexpr int rader_w2(int n) { return 256 + (n + 255) / 256; }
constexpr size_t rader_lds_bytes(int m1, int nt, int r1, int r2, int r3, int pad, bool tp) {
    const int m2 = r1 * r2 * r3, L = m1 * m2;
    const int lp = pad == 2 ? m2 + m2 / 8 : pad ? m2 + m2 / 32 : m2;
    const int tabs = tp ? fused_tp_a(r1, r2, r3) + fused_tp_b(r1, r2, r3) : 0;
    // W_M2, W_L, the r-major tables, the row buffer, W_C (C = 2p = 2L + 2), the partial sums
    return (size_t)(m2 + rader_w2(L) + tabs + rader_buf(m1, m2, lp) + rader_w2(L + 1) + 2 * (nt / 64)) *
           sizeof(float2);
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

template <int M1, int NT, int R1, int R2, int R3>
__global__ __launch_bounds__(NT) void k_rowrader(RowPassArgs a) {
    constexpr int M2 = R1 * R2 * R3;
    constexpr int L = M1 * M2;  // p - 1
    constexpr int KB = M1;      // one batch: the M1 rows k1 of each row group
    constexpr int G = 2 * KB;   // row group s = 0 (even columns), 1 (odd columns)
    constexpr int NW = NT / 64;
    static_assert(NT % 64 == 0 && 2 * M2 <= NT, "one column n2 of one row group per thread");
    static_assert(R1 > 1 && (R2 > 1 || R3 == 1), "stage radices first");
    constexpr int P = rader_pad(M1, NT, R1, R2, R3);
    constexpr int LP = kRowPitch<M2, P>;
    constexpr int TP = rader_tp(M1, NT, R1, R2, R3) ? 1 : 0;
    constexpr int RA = fused_tp_radix(R2, R3), NA = TP ? fused_tp_a(R1, R2, R3) : 0;
    constexpr int NBT = TP ? fused_tp_b(R1, R2, R3) : 0;
    constexpr int NBUF = rader_buf(M1, M2, LP);
    extern __shared__ float2 smem[];
    float2* s_tw = smem;                    // W_M2^i, i < M2
    float2* s_wm = s_tw + M2;               // W_L^e: 256 low + (L + 255) / 256 high entries
    float2* s_ta = s_wm + rader_w2(L);      // TP: r-major tables of the pivot stage ...
    float2* s_tb = NBT ? s_ta + NA : s_ta;  // ... and of the last inverse stage
    float2* buf = s_ta + NA + NBT;          // [s][k1][n2] rows, then E_0 | E_1
    float2* s_wc = buf + NBUF;              // W_C^e: 256 low + (p + 255) / 256 high entries
    float2* s_red = s_wc + rader_w2(L + 1); // [wave][s]: partial sums of a_s
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int s = t / M2, n2 = t - s * M2;  // row group and column (t < 2 M2)
    const bool live = t < 2 * M2;
    const int p = a.rader_p;
    const bio::rsrc_t r_bh = bio::make_rsrc(a.bhat, (uint32_t)L * 8u);

    // phase A: a_s[n2 + M2 n1] (n1 < M1) and x_s[0]; the tables while they load
    const bio::rsrc_t r_x = bio::make_rsrc(a.X0 + (size_t)blockIdx.x * a.ldx, (uint32_t)a.C * 8u);
    float2 v[M1];
    float2 x0 = make_float2(0.f, 0.f);
    if (live) {
#pragma unroll
        for (int n1 = 0; n1 < M1; n1++) v[n1] = bio::load2_stream(r_x, (s * L + n2) * 8, M2 * n1 * 8);
        x0 = bio::load2_stream(r_x, (2 * L + s) * 8, 0);
    }
    for (int i = t; i < M2; i += NT) s_tw[i] = a.tw2[i];
    for (int i = t; i < 256 + (L + 255) / 256; i += NT) s_wm[i] = a.wm2[i];
    for (int i = t; i < 256 + (p + 255) / 256; i += NT) s_wc[i] = a.wc2[i];
    if constexpr (TP) {
        constexpr int nsa = NA ? M2 / RA : 1, nsb = NBT ? M2 / R1 : 1;
        for (int i = t; i < NA; i += NT) s_ta[i] = a.tw2[(i / nsa + 1) * (i % nsa)];
        for (int i = t; i < NBT; i += NT) s_tb[i] = a.tw2[(i / nsb + 1) * (i % nsb)];
    }
    __syncthreads();
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
    if (live) fused_twiddle<M1, false>(v, n2, s_wm);

    // phase B: both row groups' rows k1 into LDS, the convolution with B^
    if (live) {
        float2* col = buf + s * KB * LP + lds_idx<P>(n2);
#pragma unroll
        for (int k = 0; k < KB; k++) col[k * LP] = v[k];
    }
    __syncthreads();
    fused_conv<M2, G, KB, NT, R1, R2, R3, P, TP>(r_bh, buf, s_tw, 0, s_ta, s_tb);

    // phase C: the inverse length-M1 codelet gives c_s[m], m = n2 + M2 n1
    int kk[M1];  // g^-m mod p
    if (live) {
        const float2* col = buf + s * KB * LP + lds_idx<P>(n2);
#pragma unroll
        for (int k = 0; k < KB; k++) v[k] = col[k * LP];
        fused_twiddle<M1, true>(v, n2, s_wm);
        RegFft<M1>::run(v, true);
#pragma unroll
        for (int n1 = 0; n1 < M1; n1++) kk[n1] = a.rader_k[n2 + M2 * n1];
    }
    __syncthreads();  // the rows are read: buf now takes E_0 | E_1 in natural order
    if (live) {
        float2* e = buf + s * p;
#pragma unroll
        for (int n1 = 0; n1 < M1; n1++) e[kk[n1]] = cadd(x0, v[n1]);
        if (n2 == 0) {
            float2 sum = x0;
            for (int w = 0; w < NW; w++) sum = cadd(sum, s_red[2 * w + s]);
            e[0] = sum;
        }
    }
    __syncthreads();

    // the radix-2 combination, row u and its mirror row R - u
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    const int u = a.u_base + (int)blockIdx.x;
    const bool mirror = u > 0 && a.R - u > u;
    f32x2* o = reinterpret_cast<f32x2*>(a.out + (size_t)u * a.P);
    f32x2* om = reinterpret_cast<f32x2*>(mirror ? a.out_m + (size_t)(a.R - u) * a.P : a.out);
    for (int k = t; k < p; k += NT) {
        const float2 e0 = buf[k], e1 = cmul(buf[p + k], cmul(s_wc[256 + (k >> 8)], s_wc[k & 255]));
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

bool rader_shape(int i, RaderShape* out) {
    static const RaderShape kShapes[] = {
#define SPFFT_SHAPE(m1, nt, r1, r2, r3) {m1, nt, {r1, r2, r3}, rader_lds_bytes_of(m1, nt, r1, r2, r3)},
        SPFFT_RADER_LIST(SPFFT_SHAPE)
#undef SPFFT_SHAPE
    };
    if (i < 0 || i >= (int)(sizeof(kShapes) / sizeof(kShapes[0]))) return false;
    *out = kShapes[i];
    return true;
}

hipError_t launch_rader(const RowPassArgs& a, hipStream_t s, bool prepare_only) {
#define SPFFT_CASE(m1, nt, r1, r2, r3)                                                                   \
    if (a.M1 == m1 && a.fused_nt == nt && a.f2.radix[0] == r1 && (a.f2.nstages > 1 ? a.f2.radix[1] : 1) == r2 && \
        (a.f2.nstages > 2 ? a.f2.radix[2] : 1) == r3) {                                                  \
        constexpr size_t lds = rader_lds_bytes_of(m1, nt, r1, r2, r3);                                   \
        const void* k = reinterpret_cast<const void*>(&k_rowrader<m1, nt, r1, r2, r3>);                   \
        if (prepare_only) return hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
        hipLaunchKernelGGL((k_rowrader<m1, nt, r1, r2, r3>), dim3(a.tile_rows), dim3(nt), lds, s, a);   \
        return hipGetLastError();                                                                        \
    }
    SPFFT_RADER_LIST(SPFFT_CASE)
#undef SPFFT_CASE
    return hipErrorInvalidValue;
}

}  // namespace spfft
