// rowpass_rader.hip -- the Rader row transform's hand-written shapes and dispatch.
//
// k_rowrader (rader_impl.h: the algorithm, for widths C = 2p and C = p, p an odd prime)
// runs here for the widths 2p with p - 1 = 29 m (c3's 16706 = 2 * 8353 among them); the
// shapes of every prime p with 7-smooth p - 1 are generated (rader_shapes.h) and
// instantiated in rowpass_rader_gen_{a,b,c,d}.hip.  This file also holds the precision
// form k_rowrader_mfma (the length-29 DFTs of phases A and C on the matrix cores) and the
// launch dispatch over all the shape lists.
//
// Replaces, for these widths, the Bluestein row pass of the reference
// (src/sparse_fft_bluestein.cu:2327-2370); the output is the same half spectrum.
#include "rader_impl.h"

namespace spfft {

namespace {

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

// The generated FP32 shapes (rader_shapes.h), one list per compilation unit
// (rowpass_rader_gen_{a,b,c,d}.hip).
#define SPFFT_GEN_DECL(x)                                                                                   \
    bool rader_gen_shape_##x(int i, RaderShape* out);                                                     \
    hipError_t launch_rader_gen_##x(const RowPassArgs& a, hipStream_t s, bool prepare_only, bool* found); \
    hipError_t launch_rader_gen_batch_##x(const RowPassArgs& a, const RowPassArgs* descs, const int2* work, int n, \
                                          hipStream_t s, bool* found);
SPFFT_GEN_DECL(a)
SPFFT_GEN_DECL(b)
SPFFT_GEN_DECL(c)
SPFFT_GEN_DECL(d)
#undef SPFFT_GEN_DECL

bool rader_shape(int i, RaderShape* out) {
    static const RaderShape kShapes[] = {
#define SPFFT_SHAPE(m1, nt, r1, r2, r3) {m1, nt, {r1, r2, r3}, rader_lds_bytes_of(m1, nt, r1, r2, r3), 0, 2},
        SPFFT_RADER_LIST(SPFFT_SHAPE)
#undef SPFFT_SHAPE
#define SPFFT_SHAPE(m1, nt, r1, r2, r3) {m1, nt, {r1, r2, r3}, rader_lds_bytes_of(m1, nt, r1, r2, r3), 1, 2},
        SPFFT_RADER_MFMA_LIST(SPFFT_SHAPE)
#undef SPFFT_SHAPE
    };
    constexpr int nk = (int)(sizeof(kShapes) / sizeof(kShapes[0]));
    if (i < 0) return false;
    if (i < nk) {
        *out = kShapes[i];
        return true;
    }
    i -= nk;
    for (auto gen : {rader_gen_shape_a, rader_gen_shape_b, rader_gen_shape_c, rader_gen_shape_d}) {
        int n = 0;
        RaderShape r;
        while (gen(n, &r)) n++;
        if (i < n) return gen(i, out);
        i -= n;
    }
    return false;
}

hipError_t launch_rader_batch(const RowPassArgs& a, const RowPassArgs* descs, const int2* work, int n,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
#define SPFFT_CASE(m1, nt, r1, r2, r3)                                                                     \
    if (a.precision == 1 && a.M1 == m1 && a.fused_nt == nt && a.f2.radix[0] == r1 &&                       \
        (a.f2.nstages > 1 ? a.f2.radix[1] : 1) == r2 && (a.f2.nstages > 2 ? a.f2.radix[2] : 1) == r3) {     \
        hipLaunchKernelGGL((k_rowrader_mfma_batch<m1, nt, r1, r2, r3>), dim3(n), dim3(nt),                 \
                           rader_lds_bytes_of(m1, nt, r1, r2, r3), s, descs, work);                        \
        return hipGetLastError();                                                                          \
    }
    SPFFT_RADER_MFMA_LIST(SPFFT_CASE)
#undef SPFFT_CASE
#define SPFFT_CASE(m1, nt, r1, r2, r3) \
    if (rader_shape_is(a, m1, nt, r1, r2, r3, 2)) return launch_rader_fp32_batch<m1, nt, r1, r2, r3, 2>(descs, work, n, s);
    SPFFT_RADER_LIST(SPFFT_CASE)
#undef SPFFT_CASE
    for (auto gen : {launch_rader_gen_batch_a, launch_rader_gen_batch_b, launch_rader_gen_batch_c,
                     launch_rader_gen_batch_d}) {
        bool found = false;
        const hipError_t e = gen(a, descs, work, n, s, &found);
        if (found) return e;
    }
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
#define SPFFT_CASE(m1, nt, r1, r2, r3) \
    if (rader_shape_is(a, m1, nt, r1, r2, r3, 2)) return launch_rader_fp32<m1, nt, r1, r2, r3, 2>(a, s, prepare_only);
    SPFFT_RADER_LIST(SPFFT_CASE)
#undef SPFFT_CASE
    for (auto gen : {launch_rader_gen_a, launch_rader_gen_b, launch_rader_gen_c, launch_rader_gen_d}) {
        bool found = false;
        const hipError_t e = gen(a, s, prepare_only, &found);
        if (found) return e;
    }
    return hipErrorInvalidValue;
}

}  // namespace spfft
