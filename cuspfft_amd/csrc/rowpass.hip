// rowpass.hip -- column sparse-DFT pass and Bluestein row pass for gfx950.
//
// Replaces, per u-tile, the reference's memset + csc_build_bluestein_input
// + run_fft_mixed_23 (forward) + pointwise_mul_batched + run_fft_mixed_23
// (inverse) + scale + bluestein_finalize sequence
// (src/sparse_fft_bluestein.cu:2327-2370 and :2649-2696) with three kernels
// built on a four-step split M = M1*M2 of the smooth Bluestein length:
//
//   k_colpass_fft1   pass 1 (exact integer phase (u*r) mod R, twiddles from a
//                    two-level LDS table) fused with the chirp pre-multiply,
//                    the zero padding c in [C, M) and the length-M1 sub-FFTs
//                    over n1 (n = M2*n1 + n2).  Writes T[u][k1][n2].
//   k_conv_rows      per (u, k1) row of T: outer twiddle W_M^{k1 n2}, FFT over
//                    n2 (length M2), product with B^/M, inverse FFT, conjugate
//                    twiddle.  One read and one write of T.
//   k_ifft1_finalize inverse length-M1 sub-FFTs over k1, chirp post-multiply
//                    and the half-spectrum write X[u][v], v <= C/2, at pitch P.
//
// All sub-FFTs run in LDS (spfft_device.h).  Every workgroup is 256 threads
// (4 wave64s); tables are staged into LDS at kernel start.
#include "spfft_device.h"
#include "spfft_kernels.h"

namespace spfft {

namespace {

__device__ __forceinline__ float2 wr_lookup(const float2* s_hi, const float2* s_lo, uint32_t ph,
                                            int hs, uint32_t mask) {
    return cmul(s_hi[ph >> hs], s_lo[ph & mask]);
}

template <bool PACKED, int UMAX>
__device__ __forceinline__ void k1_body(const RowPassArgs& a, int bx, int by, float2* smem) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int J = a.J;
    const int lj = 31 - __clz(J);
    const int j0 = bx * J;
    const int u0l = by * UMAX;
    const int nu = min(UMAX, a.tile_rows - u0l);
    const uint32_t u0 = (uint32_t)(a.u_base + u0l);
    const int L1 = a.M1;
    const int G = UMAX * J;
    const int nlo = 1 << a.hi_shift;
    float2* s_lo = smem;
    float2* s_hi = s_lo + nlo;
    float2* s_tw = s_hi + a.n_hi;
    float2* bufA = s_tw + L1;
    float2* bufB = bufA + G * L1;
    for (int i = tid; i < nlo; i += nt) s_lo[i] = a.wr_lo[i];
    for (int i = tid; i < a.n_hi; i += nt) s_hi[i] = a.wr_hi[i];
    for (int i = tid; i < L1; i += nt) s_tw[i] = a.tw1[i];
    __syncthreads();

    const uint32_t R = (uint32_t)a.R;
    const uint32_t mask = (uint32_t)nlo - 1u;
    const int hs = a.hi_shift;
    for (int q = tid; q < J * L1; q += nt) {
        const int jj = q & (J - 1);
        const int n1 = q >> lj;
        const int n2 = j0 + jj;
        const int c = n2 + a.M2 * n1;
        float2 acc[UMAX];
#pragma unroll
        for (int uu = 0; uu < UMAX; uu++) acc[uu] = make_float2(0.f, 0.f);
        if (n2 < a.M2 && c < a.C) {
            const int beg = a.colstart[c], end = a.colstart[c + 1];
            for (int p = beg; p < end; p++) {
                uint32_t r = PACKED ? (uint32_t)((const uint16_t*)a.row_idx)[p]
                                    : (uint32_t)((const int*)a.row_idx)[p];
                if (r >= R) r = fastmod_u32(r, a.r_magic, R);
                uint32_t ph = PACKED ? fastmod_u32(u0 * r, a.r_magic, R)
                                     : (uint32_t)(((uint64_t)u0 * (uint64_t)r) % (uint64_t)R);
#pragma unroll
                for (int uu = 0; uu < UMAX; uu++) {
                    const float2 w = wr_lookup(s_hi, s_lo, ph, hs, mask);
                    acc[uu].x += w.x;
                    acc[uu].y += w.y;
                    ph += r;
                    if (ph >= R) ph -= R;
                }
            }
            const float2 hc = a.h[c];
#pragma unroll
            for (int uu = 0; uu < UMAX; uu++) acc[uu] = cmul(acc[uu], hc);
        }
#pragma unroll
        for (int uu = 0; uu < UMAX; uu++) bufA[(uu * J + jj) * L1 + n1] = acc[uu];
    }
    __syncthreads();
    const float2* res = lds_fft(bufA, bufB, G, a.f1, s_tw, false);

    const int total = nu * L1 * J;
    const float inv_l1 = 1.f / (float)L1;
    for (int q = tid; q < total; q += nt) {
        const int jj = q & (J - 1);
        const int rest = q >> lj;
        const int uu = fdiv_small(rest, L1, inv_l1);
        const int k1 = rest - uu * L1;
        const int n2 = j0 + jj;
        if (n2 < a.M2)
            a.T[(size_t)(u0l + uu) * a.M + (size_t)k1 * a.M2 + n2] = res[(uu * J + jj) * L1 + k1];
    }
}

__device__ __forceinline__ void k2_body(const RowPassArgs& a, int bx, float2* smem) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int L2 = a.M2, M1 = a.M1, G = a.G2;
    float2* s_tw = smem;
    float2* bufA = s_tw + L2;
    float2* bufB = bufA + G * L2;
    int* s_k1 = reinterpret_cast<int*>(bufB + G * L2);  // k1 of each row, -1 past the tile
    for (int i = tid; i < L2; i += nt) s_tw[i] = a.tw2[i];
    const int row0 = bx * G;
    const int nrows = a.tile_rows * M1;
    for (int g = tid; g < G; g += nt) s_k1[g] = (row0 + g < nrows) ? (row0 + g) % M1 : -1;
    __syncthreads();
    const float inv_l2 = 1.f / (float)L2;
    for (int q = tid; q < G * L2; q += nt) {
        const int g = fdiv_small(q, L2, inv_l2);
        const int n2 = q - g * L2;
        const int k1 = s_k1[g];
        float2 x = make_float2(0.f, 0.f);
        if (k1 >= 0) x = cmul(a.T[(size_t)(row0 + g) * L2 + n2], a.twM[(size_t)k1 * L2 + n2]);
        bufA[q] = x;
    }
    __syncthreads();
    float2* f = lds_fft(bufA, bufB, G, a.f2, s_tw, false);
    float2* o = (f == bufA) ? bufB : bufA;
    for (int q = tid; q < G * L2; q += nt) {
        const int g = fdiv_small(q, L2, inv_l2);
        const int k2 = q - g * L2;
        const int k1 = s_k1[g];
        if (k1 >= 0) f[q] = cmul(f[q], a.bhat[(size_t)k1 * L2 + k2]);
    }
    __syncthreads();
    const float2* y = lds_fft(f, o, G, a.f2, s_tw, true);
    for (int q = tid; q < G * L2; q += nt) {
        const int g = fdiv_small(q, L2, inv_l2);
        const int n2 = q - g * L2;
        const int k1 = s_k1[g];
        if (k1 >= 0) a.T[(size_t)(row0 + g) * L2 + n2] = cmulc(y[q], a.twM[(size_t)k1 * L2 + n2]);
    }
}

template <int UMAX>
__device__ __forceinline__ void k3_body(const RowPassArgs& a, int bx, int by, float2* smem) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int J = a.J;
    const int lj = 31 - __clz(J);
    const int j0 = bx * J;
    const int u0l = by * UMAX;
    const int nu = min(UMAX, a.tile_rows - u0l);
    const int L1 = a.M1, M2 = a.M2;
    const int G = UMAX * J;
    float2* s_tw = smem;
    float2* bufA = s_tw + L1;
    float2* bufB = bufA + G * L1;
    for (int i = tid; i < L1; i += nt) s_tw[i] = a.tw1[i];
    const float inv_l1 = 1.f / (float)L1;
    for (int q = tid; q < UMAX * L1 * J; q += nt) {
        const int jj = q & (J - 1);
        const int rest = q >> lj;
        const int uu = fdiv_small(rest, L1, inv_l1);
        const int k1 = rest - uu * L1;
        const int n2 = j0 + jj;
        float2 x = make_float2(0.f, 0.f);
        if (n2 < M2 && uu < nu) x = a.T[(size_t)(u0l + uu) * a.M + (size_t)k1 * M2 + n2];
        bufA[(uu * J + jj) * L1 + k1] = x;
    }
    __syncthreads();
    const float2* y = lds_fft(bufA, bufB, G, a.f1, s_tw, true);
    // outputs v = n2 + M2*n1 < V; only n1 < ceil(V / M2) can qualify
    const int n1max = min(L1, (a.V + M2 - 1) / M2);
    const int total = nu * n1max * J;
    const float inv_n1 = 1.f / (float)n1max;
    for (int q = tid; q < total; q += nt) {
        const int jj = q & (J - 1);
        const int rest = q >> lj;
        const int uu = fdiv_small(rest, n1max, inv_n1);
        const int n1 = rest - uu * n1max;
        const int n2 = j0 + jj;
        const int v = n2 + M2 * n1;
        if (n2 < M2 && v < a.V) {
            const size_t u = (size_t)(a.u_base + u0l + uu);
            a.out[u * a.P + v] = cmul(y[(uu * J + jj) * L1 + n1], a.h[v]);
        }
    }
}

}  // namespace

template <bool PACKED, int UMAX>
__global__ __launch_bounds__(256) void k_colpass_fft1(RowPassArgs a) {
    extern __shared__ float2 smem[];
    k1_body<PACKED, UMAX>(a, blockIdx.x, blockIdx.y, smem);
}

__global__ __launch_bounds__(256) void k_conv_rows(RowPassArgs a) {
    extern __shared__ float2 smem[];
    k2_body(a, blockIdx.x, smem);
}

template <int UMAX>
__global__ __launch_bounds__(256) void k_ifft1_finalize(RowPassArgs a) {
    extern __shared__ float2 smem[];
    k3_body<UMAX>(a, blockIdx.x, blockIdx.y, smem);
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
size_t k1_lds_bytes(const RowPassArgs& a) {
    const size_t G = (size_t)a.U * a.J;
    return ((size_t)(1 << a.hi_shift) + a.n_hi + a.M1 + 2 * G * a.M1) * sizeof(float2);
}
size_t k2_lds_bytes(const RowPassArgs& a) {
    return ((size_t)a.M2 + 2 * (size_t)a.G2 * a.M2) * sizeof(float2) + (size_t)a.G2 * sizeof(int);
}
size_t k3_lds_bytes(const RowPassArgs& a) {
    const size_t G = (size_t)a.U * a.J;
    return ((size_t)a.M1 + 2 * G * a.M1) * sizeof(float2);
}

template <bool PACKED>
static hipError_t launch_k1_u(const RowPassArgs& a, dim3 grid, size_t lds, hipStream_t s) {
    switch (a.U) {
        case 1: hipLaunchKernelGGL((k_colpass_fft1<PACKED, 1>), grid, dim3(256), lds, s, a); break;
        case 2: hipLaunchKernelGGL((k_colpass_fft1<PACKED, 2>), grid, dim3(256), lds, s, a); break;
        case 4: hipLaunchKernelGGL((k_colpass_fft1<PACKED, 4>), grid, dim3(256), lds, s, a); break;
        case 8: hipLaunchKernelGGL((k_colpass_fft1<PACKED, 8>), grid, dim3(256), lds, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_rowpass_tile(const RowPassArgs& a, bool packed, hipStream_t s, hipEvent_t* ev) {
    if (a.tile_rows <= 0) return hipSuccess;
    hipError_t e;
    if (ev && (e = hipEventRecord(ev[0], s)) != hipSuccess) return e;
    const dim3 g13((a.M2 + a.J - 1) / a.J, (a.tile_rows + a.U - 1) / a.U);
    e = packed ? launch_k1_u<true>(a, g13, k1_lds_bytes(a), s)
               : launch_k1_u<false>(a, g13, k1_lds_bytes(a), s);
    if (e != hipSuccess) return e;
    if (ev && (e = hipEventRecord(ev[1], s)) != hipSuccess) return e;
    const dim3 g2(((size_t)a.tile_rows * a.M1 + a.G2 - 1) / a.G2);
    hipLaunchKernelGGL(k_conv_rows, g2, dim3(256), k2_lds_bytes(a), s, a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (ev && (e = hipEventRecord(ev[2], s)) != hipSuccess) return e;
    switch (a.U) {
        case 1: hipLaunchKernelGGL((k_ifft1_finalize<1>), g13, dim3(256), k3_lds_bytes(a), s, a); break;
        case 2: hipLaunchKernelGGL((k_ifft1_finalize<2>), g13, dim3(256), k3_lds_bytes(a), s, a); break;
        case 4: hipLaunchKernelGGL((k_ifft1_finalize<4>), g13, dim3(256), k3_lds_bytes(a), s, a); break;
        case 8: hipLaunchKernelGGL((k_ifft1_finalize<8>), g13, dim3(256), k3_lds_bytes(a), s, a); break;
        default: return hipErrorInvalidValue;
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (ev && (e = hipEventRecord(ev[3], s)) != hipSuccess) return e;
    return hipSuccess;
}

hipError_t set_rowpass_lds_limits(size_t max_bytes) {
    const void* fns[] = {
        (const void*)k_colpass_fft1<true, 1>, (const void*)k_colpass_fft1<true, 2>,
        (const void*)k_colpass_fft1<true, 4>, (const void*)k_colpass_fft1<true, 8>,
        (const void*)k_colpass_fft1<false, 1>, (const void*)k_colpass_fft1<false, 2>,
        (const void*)k_colpass_fft1<false, 4>, (const void*)k_colpass_fft1<false, 8>,
        (const void*)k_conv_rows,
        (const void*)k_ifft1_finalize<1>, (const void*)k_ifft1_finalize<2>,
        (const void*)k_ifft1_finalize<4>, (const void*)k_ifft1_finalize<8>,
    };
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace spfft
