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
#include "rowpass_internal.h"

namespace spfft {

namespace {

// ---------------------------------------------------------------------------
// k_colpass
// ---------------------------------------------------------------------------
template <bool PACKED>
__global__ __launch_bounds__(kThreads) void k_colpass(RowPassArgs a) {
    extern __shared__ float2 smem[];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int c0 = blockIdx.x * 64;
    const int u0l = blockIdx.y * 64;
    const int nlo = 1 << a.hi_shift;
    float2* s_lo = smem;
    float2* s_hi = s_lo + nlo;
    float2* s_x = s_hi + a.n_hi;  // [64 rows][kPadX]
    for (int i = threadIdx.x; i < nlo; i += kThreads) s_lo[i] = a.wr_lo[i];
    for (int i = threadIdx.x; i < a.n_hi; i += kThreads) s_hi[i] = a.wr_hi[i];
    __syncthreads();

    const uint32_t R = (uint32_t)a.R;
    const uint32_t mask = (uint32_t)nlo - 1u;
    const int hs = a.hi_shift;
    const uint32_t u0 = (uint32_t)(a.u_base + u0l);
    for (int i = 0; i < 16; i++) {
        const int cl = wid + 4 * i;
        const int c = c0 + cl;
        float2 acc = make_float2(0.f, 0.f);
        if (c < a.C) {
            const int beg = a.colstart[c], end = a.colstart[c + 1];
            for (int p = beg; p < end; p++) {
                uint32_t r = PACKED ? (uint32_t)((const uint16_t*)a.row_idx)[p]
                                    : (uint32_t)((const int*)a.row_idx)[p];
                if (r >= R) r = fastmod_u32(r, a.r_magic, R);
                const uint32_t base = PACKED ? fastmod_u32(u0 * r, a.r_magic, R)
                                             : (uint32_t)(((uint64_t)u0 * r) % (uint64_t)R);
                // (u0 + lane) * r mod R = (base + lane * r) mod R; x < 64 R
                const uint32_t x = base + __umul24((uint32_t)lane, r);
                const uint32_t q = (uint32_t)((float)x * a.inv_r);
                int ph = (int)(x - __umul24(q, R));
                if (ph < 0) ph += (int)R;
                else if (ph >= (int)R) ph -= (int)R;
                const float2 w = cmul(s_hi[(uint32_t)ph >> hs], s_lo[(uint32_t)ph & mask]);
                acc.x += w.x;
                acc.y += w.y;
            }
            acc = cmul(acc, a.h[c]);
        }
        s_x[lane * kPadX + cl] = acc;
    }
    __syncthreads();
    const int c = c0 + lane;
    if (c < a.C) {
        for (int i = 0; i < 16; i++) {
            const int ul = wid + 4 * i;
            if (u0l + ul < a.tile_rows) a.X0[(size_t)(u0l + ul) * a.ldx + c] = s_x[ul * kPadX + lane];
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
    return ((size_t)(1 << a.hi_shift) + a.n_hi + 64 * kPadX) * sizeof(float2);
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
    hipError_t e = launch_fft1_group_a(a, grid, s, inverse);
    if (e == hipErrorInvalidValue) e = launch_fft1_group_b(a, grid, s, inverse);
    if (e == hipErrorInvalidValue) e = launch_fft1_group_c(a, grid, s, inverse);
    return e;
}

hipError_t launch_rowpass_tile(const RowPassArgs& a, bool packed, hipStream_t s, hipEvent_t* ev) {
    if (a.tile_rows <= 0) return hipSuccess;
    hipError_t e;
    if (ev && (e = hipEventRecord(ev[0], s)) != hipSuccess) return e;
    const dim3 gp((a.C + 63) / 64, (a.tile_rows + 63) / 64);
    if (packed) hipLaunchKernelGGL((k_colpass<true>), gp, dim3(kThreads), colpass_lds_bytes(a), s, a);
    else hipLaunchKernelGGL((k_colpass<false>), gp, dim3(kThreads), colpass_lds_bytes(a), s, a);
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
