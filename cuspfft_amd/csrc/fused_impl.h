// fused_impl.h -- k_rowfused: the whole Bluestein row pass of one output row u
// in one workgroup, with the length-M row held on chip (registers + LDS).
//
// The four-step split M = M1 * M2 (M2 = R1 * R2 * R3) of k_fft1 / k_conv /
// k_ifft1, without the tile buffer T in HBM:
//   phase A  each column n2 < M2 (CPT columns per thread) loads x[M2 n1 + n2]
//            (n1 < M1) from the pass-1 row and runs the length-M1 register
//            codelet; the M1 results stay in registers.
//   phase B  the M1 rows k1 (length M2) go through LDS in NB batches of
//            KB = M1 / NB rows: each thread swaps its columns of the batch with
//            the rows processed in the previous batch, then all threads run the
//            Stockham stages R1, R2, R3 forward (the first one times W_M^{k1 n2}
//            on its reads, the last one times B^[k1 + M1 k2] / M) and inverse
//            (the last one times W_M^{-k1 n2} on its writes), in place in LDS.
//   phase C  each column runs the inverse length-M1 codelet, applies the chirp
//            and writes row u and its mirror row R - u
//            (X[R-u][v] = conj(X[u][(C - v) mod C]) for a real input).
// HBM traffic per row: the pass-1 row in, the two half-spectrum rows out.
//
// Global memory is read and written with buffer instructions.  Table reads
// (always in range) put the per-thread part of the offset in a VGPR and the
// strided part (kbase * L + r * nb) in a scalar offset.  Row reads and writes
// that may fall outside the row are predicated without branches: their offset
// is replaced by kOob, past the end of the buffer, so the load returns 0 (the
// zero padding of x for n >= C) and the store is dropped (v > C/2).  Branches
// around each of the M1 accesses would keep the whole column live in registers.
// RPW > 1 puts RPW short rows in one workgroup.
#pragma once
#include "rowpass_internal.h"

namespace spfft {

namespace fused {

using rsrc_t = __amdgpu_buffer_rsrc_t;
// raw buffer: stride 0, num_records in bytes; word 3 = 32-bit data format (gfx9)
constexpr int kRsrcWord3 = 0x00020000;

__device__ __forceinline__ rsrc_t make_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, kRsrcWord3);
}
// An offset at or past num_records makes a load return 0 and drops a store;
// kOob is past every buffer made here (num_records <= 2^31 - 256).
constexpr int kOob = 0x7fffff00;

// The LLVM buffer intrinsics, bound by name: this compiler lowers the clang
// builtins __builtin_amdgcn_raw_buffer_load_b64 / _b128 to a single 32-bit load.
typedef float f32x2 __attribute__((ext_vector_type(2)));
__device__ f32x2 raw_load_v2f32(rsrc_t r, int voff, int soff, int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.v2f32");
__device__ void raw_store_v2f32(f32x2 v, rsrc_t r, int voff, int soff, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.store.v2f32");

__device__ __forceinline__ float2 load2(rsrc_t r, int voff, int soff) {
    const f32x2 d = raw_load_v2f32(r, voff, soff, 0);
    return make_float2(d[0], d[1]);
}
__device__ __forceinline__ void store2(float2 x, rsrc_t r, int voff, int soff) {
    const f32x2 d = {x.x, x.y};
    raw_store_v2f32(d, r, voff, soff, 0);
}

}  // namespace fused

// One in-place Stockham stage of radix R over the G rows (length L) in LDS,
// NT threads.  Row g belongs to k1 = kbase + g % KB.  Reads of the whole stage
// finish at a barrier before any write; the stage ends with a barrier.  TWIN
// multiplies the values read by W_M^{k1 n2} (first forward stage), TWOUT the
// values written by W_M^{-k1 n2} (last inverse stage), MUL by B^ (last forward).
template <int R, int L, int NS, bool INV, bool MUL, bool TWIN, bool TWOUT, int G, int KB, int NT>
__device__ __forceinline__ void fused_stage(fused::rsrc_t r_tw, fused::rsrc_t r_bh, float2* buf,
                                            const float2* s_tw, int kbase) {
    constexpr int nb = L / R;
    constexpr int total = G * nb;
    constexpr int MB = (total + NT - 1) / NT;
    constexpr int step = L / (NS * R);
    float2 v[MB][R];
#pragma unroll
    for (int b = 0; b < MB; b++) {
        const int t = threadIdx.x + b * NT;
        if (t < total) {
            const int g = t / nb;
            const int j = t - g * nb;
            const float2* src = buf + g * L + j;
#pragma unroll
            for (int r = 0; r < R; r++) v[b][r] = src[r * nb];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MB; b++) {
        const int t = threadIdx.x + b * NT;
        if (t < total) {
            const int g = t / nb;
            const int j = t - g * nb;
            const int jq = j / NS;
            const int k = j - jq * NS;
            const int rowoff = (g % KB) * L;  // (k1 - kbase) * L
            if (TWIN) {
#pragma unroll
                for (int r = 0; r < R; r++)
                    v[b][r] = cmul(v[b][r], fused::load2(r_tw, (rowoff + j) * 8, (kbase * L + r * nb) * 8));
            }
            if (NS > 1 && k != 0) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    const float2 w = s_tw[r * k * step];
                    v[b][r] = INV ? cmulc(v[b][r], w) : cmul(v[b][r], w);
                }
            }
            RegFft<R>::run(v[b], INV);
            const int base = jq * NS * R + k;
            if (MUL) {
#pragma unroll
                for (int r = 0; r < R; r++)
                    v[b][r] = cmul(v[b][r], fused::load2(r_bh, (rowoff + base) * 8, (kbase * L + r * NS) * 8));
            }
            if (TWOUT) {
#pragma unroll
                for (int r = 0; r < R; r++)
                    v[b][r] = cmulc(v[b][r], fused::load2(r_tw, (rowoff + base) * 8, (kbase * L + r * NS) * 8));
            }
            float2* dst = buf + g * L + base;
#pragma unroll
            for (int r = 0; r < R; r++) dst[r * NS] = v[b][r];
        }
    }
    __syncthreads();
}

// The stage list R, Rest... of one length-L FFT (radices of 1 are no stage).
template <int L, int NS, bool INV, int G, int KB, int NT, int R, int... Rest>
__device__ __forceinline__ void fused_stages(fused::rsrc_t r_tw, fused::rsrc_t r_bh, float2* buf,
                                             const float2* s_tw, int kbase) {
    constexpr bool last = sizeof...(Rest) == 0 || ((Rest == 1) && ...);
    if constexpr (R > 1)
        fused_stage<R, L, NS, INV, !INV && last, !INV && NS == 1, INV && last, G, KB, NT>(r_tw, r_bh, buf, s_tw,
                                                                                          kbase);
    if constexpr (!last) fused_stages<L, NS * R, INV, G, KB, NT, Rest...>(r_tw, r_bh, buf, s_tw, kbase);
}

template <int M1, int NB, int RPW, int NT, int R1, int R2, int R3>
__global__ __launch_bounds__(NT) void k_rowfused(RowPassArgs a) {
    constexpr int M2 = R1 * R2 * R3;
    constexpr int KB = M1 / NB;
    constexpr int G = RPW * KB;
    constexpr int CPT = (RPW * M2 + NT - 1) / NT;  // columns n2 per thread
    static_assert(M1 % NB == 0, "batches must divide M1");
    static_assert(RPW == 1 || CPT == 1, "several rows per workgroup: one column per thread");
    static_assert(R1 > 1 && (R2 > 1 || R3 == 1), "stage radices first");
    extern __shared__ float2 smem[];
    float2* s_tw = smem;      // W_M2^i, i < M2
    float2* buf = smem + M2;  // G rows of length M2: [row of workgroup][k1 of batch][n2]
    const int t = threadIdx.x;
    const int row0 = blockIdx.x * RPW;  // first tile row of this workgroup
    int n2[CPT], g[CPT];
    bool live[CPT];
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int q = t + c * NT;
        g[c] = RPW == 1 ? 0 : q / M2;
        n2[c] = q - g[c] * M2;
        live[c] = (RPW == 1 ? q < M2 : g[c] < RPW) && row0 + g[c] < a.tile_rows;
    }
    for (int i = t; i < M2; i += NT) s_tw[i] = a.tw2[i];
    const fused::rsrc_t r_tw = fused::make_rsrc(a.twM, (uint32_t)a.M * 8u);
    const fused::rsrc_t r_bh = fused::make_rsrc(a.bhat, (uint32_t)a.M * 8u);

    // phase A
    float2 v[CPT][M1];
    {
        const int nrows = min(RPW, a.tile_rows - row0);
        const fused::rsrc_t r_x =
            fused::make_rsrc(a.X0 + (size_t)row0 * a.ldx, (uint32_t)(nrows * a.ldx) * 8u);
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            if (live[c]) {
                const int rb = g[c] * a.ldx;
#pragma unroll
                for (int n1 = 0; n1 < M1; n1++) {
                    const int w = n2[c] + M2 * n1;
                    v[c][n1] = fused::load2(r_x, w < a.C ? (rb + w) * 8 : fused::kOob, 0);
                }
                RegFft<M1>::run(v[c], false);
            }
        }
    }

    // phase B
    cx::static_for<NB>([&](auto Bc) {
        constexpr int b = decltype(Bc)::value;
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            if (live[c]) {
                float2* col = buf + g[c] * KB * M2 + n2[c];
#pragma unroll
                for (int k = 0; k < KB; k++) {
                    if constexpr (b == 0) {
                        col[k * M2] = v[c][k];
                    } else {
                        const float2 done = col[k * M2];
                        col[k * M2] = v[c][b * KB + k];
                        v[c][(b - 1) * KB + k] = done;
                    }
                }
            }
        }
        __syncthreads();
        fused_stages<M2, 1, false, G, KB, NT, R1, R2, R3>(r_tw, r_bh, buf, s_tw, b * KB);
        fused_stages<M2, 1, true, G, KB, NT, R1, R2, R3>(r_tw, r_bh, buf, s_tw, b * KB);
    });

    // phase C.  Buffer descriptors must be uniform: one for the workgroup's output
    // rows u0 .. u0 + nrows - 1 and one for their mirror rows R - u (the lowest is mlo).
    const fused::rsrc_t r_h = fused::make_rsrc(a.h, (uint32_t)a.C * 8u);
    const int u0 = a.u_base + row0;
    const int nrows = min(RPW, a.tile_rows - row0);
    const int mlo = a.R - (u0 + nrows - 1);
    const fused::rsrc_t r_o = fused::make_rsrc(a.out + (size_t)u0 * a.P, (uint32_t)(nrows * a.P) * 8u);
    const fused::rsrc_t r_m = fused::make_rsrc(a.out + (size_t)mlo * a.P, (uint32_t)(nrows * a.P) * 8u);
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        if (live[c]) {
            const float2* col = buf + g[c] * KB * M2 + n2[c];
#pragma unroll
            for (int k = 0; k < KB; k++) v[c][(NB - 1) * KB + k] = col[k * M2];
            RegFft<M1>::run(v[c], true);
            const int u = u0 + g[c];
            const bool mirror = u > 0 && a.R - u > u;
            const int ob = g[c] * a.P;                  // row u in r_o
            const int mb = mirror ? (a.R - u - mlo) * a.P : 0;  // row R - u in r_m
#pragma unroll
            for (int n1 = 0; n1 < M1; n1++) {
                const int w = n2[c] + M2 * n1;
                const bool in = w < a.C;
                const float2 x = cmul(v[c][n1], fused::load2(r_h, in ? w * 8 : fused::kOob, 0));
                fused::store2(x, r_o, w < a.V ? (ob + w) * 8 : fused::kOob, 0);
                const int vm = w == 0 ? 0 : a.C - w;
                fused::store2(make_float2(x.x, -x.y), r_m, in && mirror && vm < a.V ? (mb + vm) * 8 : fused::kOob, 0);
            }
        }
    }
}

// Instantiated shapes: X(M1, NB, RPW, NT, R1, R2, R3), M = M1 * R1 * R2 * R3 (a
// radix of 1 is no stage), NT threads per workgroup; each fits its registers
// without spills.  Group A: the long rows (several LDS batches); group B: one batch.
//   34992 = 36 * 972   (3-smooth length of 2 * 16706 - 1)
//   33600 = 35 * 960   (7-smooth length of 2 * 16706 - 1)
//   17496 = 24 * 729, 16464 = 21 * 784, 5184 = 16 * 324 (3 rows), 4704 = 21 * 224 (4 rows)
#define SPFFT_FUSED_GROUP_A(X) \
    X(36, 2, 1, 512, 12, 9, 9) X(35, 5, 1, 512, 8, 12, 10)
#define SPFFT_FUSED_GROUP_B(X)                                                                  \
    X(24, 1, 1, 768, 27, 27, 1) X(21, 1, 1, 1024, 28, 28, 1) X(16, 1, 3, 1024, 18, 18, 1) \
    X(21, 1, 4, 1024, 8, 28, 1)
#define SPFFT_FUSED_LIST(X) SPFFT_FUSED_GROUP_A(X) SPFFT_FUSED_GROUP_B(X)

constexpr size_t fused_lds_bytes_of(int m1, int nb, int rpw, int m2) {
    return (size_t)m2 * sizeof(float2) * (1 + (size_t)rpw * (m1 / nb));
}

// Each returns hipErrorInvalidValue when the shape is not in its group.
hipError_t launch_fused_group_a(const RowPassArgs& a, hipStream_t s, bool prepare_only);
hipError_t launch_fused_group_b(const RowPassArgs& a, hipStream_t s, bool prepare_only);

#define SPFFT_FUSED_LAUNCH_CASE(m1, nb, rpw, nt, r1, r2, r3)                                       \
    if (a.M1 == m1 && a.fused_nb == nb && a.fused_rpw == rpw && a.fused_nt == nt &&                 \
        a.f2.radix[0] == r1 && (a.f2.nstages > 1 ? a.f2.radix[1] : 1) == r2 &&                      \
        (a.f2.nstages > 2 ? a.f2.radix[2] : 1) == r3) {                                             \
        constexpr size_t lds = fused_lds_bytes_of(m1, nb, rpw, r1 * r2 * r3);                       \
        auto* k = &k_rowfused<m1, nb, rpw, nt, r1, r2, r3>;                                         \
        if (prepare_only)                                                                           \
            return hipFuncSetAttribute(reinterpret_cast<const void*>(k),                            \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);       \
        hipLaunchKernelGGL(k, dim3((a.tile_rows + rpw - 1) / rpw), dim3(nt), lds, s, a);            \
        return hipGetLastError();                                                                   \
    }

}  // namespace spfft
