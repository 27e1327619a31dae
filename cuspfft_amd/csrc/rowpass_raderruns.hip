// rowpass_raderruns.hip -- k_raderruns: pass A of the split row transform when the run
// length N2 = p is an odd prime with a shape for p - 1 (c5: 40421 = 83 * 487, 486 = 2 * 243):
// each run's DFT of length p by Rader's algorithm instead of a Bluestein convolution of
// about twice the length.
//
// With g a primitive root of p and L = p - 1 = M1 M2 (gcd(M1, M2) = 1), as in k_rowrader:
//   a[q] = x[g^q],  b[q] = W_p^{g^-q}   (q < L),   E[g^-m] = x[0] + (a (*) b)[m],
//   E[0] = x[0] + sum_q a[q],
// the convolution as the prime-factor DFT_L = DFT_M1 (x) DFT_M2 (entry (n1, n2) holds
// a[(M2 n1 + M1 n2) mod L]), a product with B^ = DFT_L(b) / L and the inverse.  The
// column pass writes each run in that order (run position n2 + M2 n1 holds a at (n1, n2),
// position L holds x[0]).  A workgroup takes NG runs (rows of the tile's run list, the
// tile row r being run r % vrow_n1 of X0 row r / vrow_n1): phase A loads its columns
// (n2 of run s, CPT per thread), runs the length-M1 codelet and writes the M1 rows of
// each run to LDS (row k1 = 0 holds the column sums: a wave per run adds them up to
// sum_q a[q]); phase B is the fused kernel's LDS convolution over the NG M1 rows of
// length M2; phase C runs the
// inverse codelet with x[0] added to its input 0 (which adds it to every output) and
// scatters E to natural order in LDS; the runs' spectra E[k], k < p, go to the tile
// buffer T at pitch P (k_splitcols then takes them across the runs).
#include "fused_impl.h"

namespace spfft {

namespace {

constexpr int rr_gcd(int a, int b) { return b ? rr_gcd(b, a % b) : a; }

// L2 prefetch distance (workgroups; 0: none) and minimum waves per SIMD of k_raderruns.
// c5's runs (40421 = 83 * 487, 20211 rows): 4 waves per SIMD (128 VGPRs, no spills; 4
// workgroups of 33 KB LDS per CU) 3.87 ms against 4.42 ms at the compiler's 129 VGPRs
// (3 waves); with 5 waves (radices 9 x 9 x 3, 6 or 7 runs per workgroup) 3.95 - 4.17 ms;
// the prefetch 32 or 128 workgroups ahead 4.31 - 4.73 ms.
#ifndef SPFFT_RADERRUNS_AHEAD
#define SPFFT_RADERRUNS_AHEAD 0
#endif
#ifndef SPFFT_RADERRUNS_WAVES
#define SPFFT_RADERRUNS_WAVES 4
#endif
constexpr int kRrAhead = SPFFT_RADERRUNS_AHEAD;

// LDS entries: W_M2, the runs' M1 rows and then E of every run (NG p), the runs' sums,
// the prefetch's scratch line (32 entries)
constexpr size_t rr_lds_bytes(int m1, int ng, int r1, int r2, int r3) {
    const int m2 = r1 * r2 * r3;
    return (size_t)(m2 + ng * (m1 * m2 + 1) + ng + (kRrAhead ? 32 : 0)) * sizeof(float2);
}

template <int M1, int NG, int NT, int R1, int R2, int R3>
__global__ __launch_bounds__(NT, SPFFT_RADERRUNS_WAVES) void k_raderruns(RowPassArgs a) {
    constexpr int M2 = R1 * R2 * R3;
    constexpr int L = M1 * M2;  // p - 1
    constexpr int P = 0, LP = M2;  // unpadded rows (odd M2: strided accesses spread over banks)
    constexpr int NQ = NG * M2;    // columns of the workgroup's runs
    constexpr int CPT = (NQ + NT - 1) / NT;
    static_assert(rr_gcd(M1, M2) == 1, "the prime-factor split L = M1 x M2 needs coprime factors");
    static_assert(NT % 64 == 0, "whole waves");
    extern __shared__ float2 smem[];
    float2* s_tw = smem;                       // W_M2^i, i < M2
    float2* buf = s_tw + M2;                   // [run][k1][n2] rows, then E of each run
    float2* s_e0 = buf + NG * (L + 1);         // [run]: sum_q a[q]
    const int t = threadIdx.x;
    constexpr int p = L + 1;  // = a.rader_p
    constexpr int NW = NT / 64;
    const int r0 = blockIdx.x * NG;
    const int nrun = min(NG, a.tile_rows - r0);
    const bio::rsrc_t r_bh = bio::make_rsrc(a.bhat, (uint32_t)L * 8u);
    // the workgroup's runs: one descriptor from its first run to the end of its last one
    const size_t xb = x0_row(a, r0);
    const bio::rsrc_t r_x = bio::make_rsrc(a.X0 + xb, (uint32_t)(x0_row(a, r0 + nrun - 1) - xb + p) * 8u);

    // phase A
    float2 v[CPT][M1];
    int rb[CPT];  // the column's run offset in r_x (entries), or -1
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int q = t + c * NT;
        const int s = q / M2, n2 = q - s * M2;
        rb[c] = q < NQ && s < nrun ? (int)(x0_row(a, r0 + s) - xb) : -1;
#pragma unroll
        for (int n1 = 0; n1 < M1; n1++)
            v[c][n1] = bio::load2_stream(r_x, rb[c] >= 0 ? (rb[c] + n2 + M2 * n1) * 8 : bio::kOob, 0);
    }
    for (int i = t; i < M2; i += NT) s_tw[i] = a.tw2[i];
    if constexpr (kRrAhead > 0) {
        // after the workgroup's own loads: the runs of the workgroup kRrAhead later into L2
        prefetch_rows(a, r0 + kRrAhead * NG, NG, NW, reinterpret_cast<float*>(s_e0 + NG));
    }
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int q = t + c * NT;
        if (q >= NQ) continue;
        const int s = q / M2, n2 = q - s * M2;
        RegFft<M1>::run(v[c], false);
        float2* col = buf + s * M1 * LP + n2;
#pragma unroll
        for (int k = 0; k < M1; k++) col[k * LP] = v[c][k];
    }
    lds_barrier();
    // sum_q a[q] = sum over the n2 of row k1 = 0: a wave per run sums it (lanes, then a
    // butterfly over the lanes: every lane ends with the same bits).  Phase B's first
    // writes come after its first barrier, when these reads are done.
    const int lane = t & 63;
    for (int s = t >> 6; s < nrun; s += NW) {
        float2 sum = make_float2(0.f, 0.f);
        for (int n2 = lane; n2 < M2; n2 += 64) sum = cadd(sum, buf[s * M1 * LP + n2]);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            sum.x += __shfl_xor(sum.x, o);
            sum.y += __shfl_xor(sum.y, o);
        }
        if (lane == 0) s_e0[s] = sum;
    }

    // phase B: the convolution of the NG M1 rows with B^
    fused_conv<M2, NG * M1, M1, NT, R1, R2, R3, P, 0>(r_bh, buf, s_tw, 0, nullptr, nullptr);

    // phase C: the inverse codelet with x[0] added to its input 0
    int kk[CPT][M1];
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int q = t + c * NT;
        if (q >= NQ) continue;
        const int s = q / M2, n2 = q - s * M2;
        const float2 x0 = bio::load2(r_x, rb[c] >= 0 ? (rb[c] + L) * 8 : bio::kOob, 0);
        const float2* col = buf + s * M1 * LP + n2;
#pragma unroll
        for (int k = 0; k < M1; k++) v[c][k] = col[k * LP];
        v[c][0] = cadd(v[c][0], x0);
        RegFft<M1>::run(v[c], true);
#pragma unroll
        for (int n1 = 0; n1 < M1; n1++) kk[c][n1] = a.rader_k[n2 + M2 * n1];
    }
    lds_barrier();  // the rows are read: buf takes E of each run in natural order
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int q = t + c * NT;
        if (q >= NQ) continue;
        const int s = q / M2;
        float2* e = buf + s * p;
#pragma unroll
        for (int n1 = 0; n1 < M1; n1++) e[kk[c][n1]] = v[c][n1];
    }
    // E[0] = x[0] + sum_q a[q]
    if (t < nrun) buf[t * p] = cadd(s_e0[t], bio::load2(r_x, ((int)(x0_row(a, r0 + t) - xb) + L) * 8, 0));
    lds_barrier();
    // the runs' spectra into T (pitch P)
    for (int i = t; i < nrun * p; i += NT) {
        const int s = i / p, k = i - s * p;
        a.out[(size_t)(r0 + s) * a.P + k] = buf[i];
    }
}

}  // namespace

// Instantiated shapes X(M1, NG, NT, R1, R2, R3), p - 1 = M1 R1 R2 R3:
//   486 = 2 * (27 * 9)   (run length 487: c5's 40421 = 83 * 487)
#ifndef SPFFT_RADERRUNS_LIST
#define SPFFT_RADERRUNS_LIST(X) X(2, 8, 256, 27, 9, 1)
#endif

bool raderruns_shape(int L, int* m1, int* ng, int* nt, int* r) {
#define SPFFT_RR_SHAPE(m1_, ng_, nt_, r1, r2, r3)                    \
    if (L == m1_ * r1 * r2 * r3) {                                    \
        *m1 = m1_;                                                    \
        *ng = ng_;                                                    \
        *nt = nt_;                                                    \
        r[0] = r1;                                                    \
        r[1] = r2;                                                    \
        r[2] = r3;                                                    \
        return true;                                                  \
    }
    SPFFT_RADERRUNS_LIST(SPFFT_RR_SHAPE)
#undef SPFFT_RR_SHAPE
    return false;
}

size_t raderruns_lds_bytes(int L) {
#define SPFFT_RR_LDS(m1, ng, nt, r1, r2, r3) \
    if (L == m1 * r1 * r2 * r3) return rr_lds_bytes(m1, ng, r1, r2, r3);
    SPFFT_RADERRUNS_LIST(SPFFT_RR_LDS)
#undef SPFFT_RR_LDS
    return 0;
}

hipError_t launch_raderruns(const RowPassArgs& a, hipStream_t s, bool prepare_only) {
#define SPFFT_RR_CASE(m1, ng, nt, r1, r2, r3)                                                          \
    if (a.M1 == m1 && a.fused_rpw == ng && a.fused_nt == nt && a.f2.radix[0] == r1 &&                  \
        (a.f2.nstages > 1 ? a.f2.radix[1] : 1) == r2 && (a.f2.nstages > 2 ? a.f2.radix[2] : 1) == r3) { \
        constexpr size_t lds = rr_lds_bytes(m1, ng, r1, r2, r3);                                       \
        const void* k = reinterpret_cast<const void*>(&k_raderruns<m1, ng, nt, r1, r2, r3>);           \
        if (prepare_only) return hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
        if (a.tile_rows <= 0) return hipSuccess;                                                       \
        hipLaunchKernelGGL((k_raderruns<m1, ng, nt, r1, r2, r3>), dim3((a.tile_rows + ng - 1) / ng),   \
                           dim3(nt), lds, s, a);                                                       \
        return hipGetLastError();                                                                      \
    }
    SPFFT_RADERRUNS_LIST(SPFFT_RR_CASE)
#undef SPFFT_RR_CASE
    return hipErrorInvalidValue;
}

}  // namespace spfft
