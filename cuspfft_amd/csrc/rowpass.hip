// rowpass.hip -- column sparse-DFT pass and Bluestein row pass for gfx950.
//
// Replaces, per u-tile, the reference's memset + csc_build_bluestein_input
// + run_fft_mixed_23 (forward) + pointwise_mul_batched + run_fft_mixed_23
// (inverse) + scale + bluestein_finalize sequence
// (src/sparse_fft_bluestein.cu:2327-2370 and :2649-2696) with four kernels
// on a four-step split M = M1*M2 of the smooth Bluestein length:
//
//   k_colpass  pass 1.  A wave owns one column at a time and its 64 lanes own
//              64 consecutive output rows u, so the column's row indices are
//              wave-uniform loads and every lane sums W_R^{(u r) mod R} with the
//              exact integer phase (a scalar fastmod for the wave's first row,
//              then lane*r added and reduced with a float-reciprocal estimate
//              corrected in integers).  W_R comes from a two-level LDS table.
//              The chirp h[c] is applied, and a 64x64 LDS transpose makes the
//              X0 row writes coalesced (the reference writes with stride M).
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
// Only rows u <= R/2 go through the four kernels: the input is real, so each
// column DFT is Hermitian and X[R-u][v] = conj(X[u][(C - v) mod C]).  The
// Bluestein convolution yields the full length-C spectrum of row u, and
// k_ifft1 writes both row u and its mirror row R - u from it -- half of the
// reference's pass-1 and row-FFT work for the same output.
#include "rowpass_internal.h"

namespace spfft {

namespace {

// ---------------------------------------------------------------------------
// k_colpass: 64 columns x 64*U rows per workgroup.  Wave w takes columns
// c0 + w + 4i; lane l takes rows u0 + l + 64t (t < U).  For a nonzero row r:
//   ph_0 = (u0 r + l r) mod R     (scalar fastmod of u0 r, then l r added and
//                                  reduced with a float estimate fixed in integers)
//   ph_t = ph_{t-1} + (64 r mod R), one conditional subtraction
// W_R^ph comes from one LDS table of R entries (FULL) or from the two-level
// table hi[ph >> s] * lo[ph & (2^s - 1)].
// ---------------------------------------------------------------------------
template <bool PACKED, bool FULL, int U>
__global__ __launch_bounds__(kThreads) void k_colpass(RowPassArgs a) {
    extern __shared__ float2 smem[];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int c0 = blockIdx.x * 64;
    const int u0l = blockIdx.y * 64 * U;
    const uint32_t R = (uint32_t)a.R;
    const int nlo = 1 << a.hi_shift;
    float2* s_lo = smem;                       // FULL: the whole W_R table
    float2* s_hi = s_lo + nlo;
    if (FULL) {
        for (int i = threadIdx.x; i < (int)R; i += kThreads) s_lo[i] = a.wr_full[i];
    } else {
        for (int i = threadIdx.x; i < nlo; i += kThreads) s_lo[i] = a.wr_lo[i];
        for (int i = threadIdx.x; i < a.n_hi; i += kThreads) s_hi[i] = a.wr_hi[i];
    }
    __syncthreads();

    const uint32_t mask = (uint32_t)nlo - 1u;
    const int hs = a.hi_shift;
    const uint32_t u0 = (uint32_t)(a.u_base + u0l);
    for (int i = 0; i < 16; i++) {
        const int c = c0 + wid + 4 * i;
        if (c >= a.C) break;
        float2 acc[U];
#pragma unroll
        for (int t = 0; t < U; t++) acc[t] = make_float2(0.f, 0.f);
        const int beg = a.colstart[c], end = a.colstart[c + 1];
        for (int p = beg; p < end; p++) {
            uint32_t r = PACKED ? (uint32_t)((const uint16_t*)a.row_idx)[p]
                                : (uint32_t)((const int*)a.row_idx)[p];
            if (r >= R) r = fastmod_u32(r, a.r_magic, R);
            const uint32_t base = PACKED ? fastmod_u32(u0 * r, a.r_magic, R)
                                         : (uint32_t)(((uint64_t)u0 * r) % (uint64_t)R);
            const uint32_t d64 = fastmod_u32(r << 6, a.r_magic, R);
            // (u0 + lane) * r mod R = (base + lane * r) mod R with base + lane * r < 64 R
            const uint32_t x = base + __umul24((uint32_t)lane, r);
            const uint32_t q = (uint32_t)((float)x * a.inv_r);
            int ph = (int)(x - __umul24(q, R));
            if (ph < 0) ph += (int)R;
            else if (ph >= (int)R) ph -= (int)R;
#pragma unroll
            for (int t = 0; t < U; t++) {
                const float2 w = FULL ? s_lo[ph] : cmul(s_hi[(uint32_t)ph >> hs], s_lo[(uint32_t)ph & mask]);
                acc[t].x += w.x;
                acc[t].y += w.y;
                ph += (int)d64;
                if (ph >= (int)R) ph -= (int)R;
            }
        }
        const float2 hc = a.h[c];
#pragma unroll
        for (int t = 0; t < U; t++) {
            const int ul = u0l + lane + 64 * t;
            if (ul < a.tile_rows) a.X0[(size_t)ul * a.ldx + c] = cmul(acc[t], hc);
        }
    }
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
    if (a.wr_full) return (size_t)a.R * sizeof(float2);
    return ((size_t)(1 << a.hi_shift) + a.n_hi) * sizeof(float2);
}

template <bool PACKED, bool FULL>
static void launch_colpass_u(const RowPassArgs& a, hipStream_t s) {
    const int U = a.colpass_rows;
    const dim3 gp((a.C + 63) / 64, (a.tile_rows + 64 * U - 1) / (64 * U));
    const size_t lds = colpass_lds_bytes(a);
    switch (U) {
        case 1: hipLaunchKernelGGL((k_colpass<PACKED, FULL, 1>), gp, dim3(kThreads), lds, s, a); break;
        case 2: hipLaunchKernelGGL((k_colpass<PACKED, FULL, 2>), gp, dim3(kThreads), lds, s, a); break;
        default: hipLaunchKernelGGL((k_colpass<PACKED, FULL, 4>), gp, dim3(kThreads), lds, s, a); break;
    }
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

hipError_t launch_rowpass_tile(const RowPassArgs& a, bool packed, hipStream_t s, hipEvent_t* ev) {
    if (a.tile_rows <= 0) return hipSuccess;
    hipError_t e;
    if (ev && (e = hipEventRecord(ev[0], s)) != hipSuccess) return e;
    if (packed) {
        if (a.wr_full) launch_colpass_u<true, true>(a, s);
        else launch_colpass_u<true, false>(a, s);
    } else {
        if (a.wr_full) launch_colpass_u<false, true>(a, s);
        else launch_colpass_u<false, false>(a, s);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev && (e = hipEventRecord(ev[1], s)) != hipSuccess) return e;
    const dim3 g1((a.M2 + 63) / 64, (a.tile_rows + 3) / 4);
    if ((e = launch_fft1(a, g1, s, false)) != hipSuccess) return e;
    if (ev && (e = hipEventRecord(ev[2], s)) != hipSuccess) return e;
    if ((e = launch_conv(a, s)) != hipSuccess) return e;
    if (ev && (e = hipEventRecord(ev[3], s)) != hipSuccess) return e;
    if ((e = launch_fft1(a, g1, s, true)) != hipSuccess) return e;
    if (ev && (e = hipEventRecord(ev[4], s)) != hipSuccess) return e;
    return hipSuccess;
}

}  // namespace spfft
