// fused_impl.h -- k_rowfused: the whole Bluestein row pass of one output row u
// in one workgroup, with the length-M row held on chip (registers + LDS).
//
// The four-step split M = M1 * M2 (M2 = R1 * R2 * R3) of k_fft1 / k_conv /
// k_ifft1, without the tile buffer T in HBM:
//   phase A  each column n2 < M2 (CPT columns per thread) loads x[M2 n1 + n2]
//            (n1 < M1) from the pass-1 row, runs the length-M1 register
//            codelet and multiplies by W_M^{k1 n2}; the M1 results stay in
//            registers.
//   phase B  the M1 rows k1 (length M2) go through LDS in NB batches of
//            KB = M1 / NB rows: each thread swaps its columns of the batch with
//            the rows processed in the previous batch, then all threads run the
//            Stockham stages R1, R2, R3 forward (the last one times
//            B^[k1 + M1 k2] / M) and R3, R2, R1 inverse, in place in LDS; the
//            last forward and first inverse stage share one pass.
//   phase C  each column multiplies by W_M^{-k1 n2}, runs the inverse
//            length-M1 codelet, applies the chirp and writes row u and its
//            mirror row R - u
//            (X[R-u][v] = conj(X[u][(C - v) mod C]) for a real input).
// HBM traffic per row: the pass-1 row in, the two half-spectrum rows out.
//
// Global memory is read and written with buffer instructions: the per-thread
// part of an offset in a VGPR, the strided part (M2 n1, kbase * L + r * nb) in
// a scalar offset.  On gfx950 the buffer bounds check takes the sum of both, so
// a row made a buffer of its valid entries gets the zero padding of x (reads
// past C return 0) and drops the writes past v = C/2 with no per-element code;
// other accesses that may leave the row take the offset kOob (past the end of
// every buffer) instead of a branch, since branches around each of the M1
// accesses would keep the whole column live in registers.  RPW > 1 puts RPW
// short rows in one workgroup (one descriptor for all, explicit bounds).
#pragma once
#include <algorithm>
#include <atomic>

#include "buffer_io.h"
#include "rowpass_internal.h"

namespace spfft {


// One in-place Stockham stage of radix R over the G rows (length L) in LDS,
// NT threads.  Row g belongs to k1 = kbase + g % KB.  Reads of the whole stage
// finish at a barrier before any write; the stage ends with a barrier.  (Doing
// the arithmetic before that barrier measured no faster and needs more
// registers.)
//   MODE kFwd / kInv: a forward / inverse stage after stages with product NS.
//   MODE kPivot: the last forward stage (NS = L / R), the product with B^, and
//     the first inverse stage, which runs the inverse radices in reverse order
//     so that it takes the same R values as the butterfly that produced them
//     (j + r L/R): one LDS round trip for both.
enum { kFwd = 0, kInv = 1, kPivot = 2 };
// the W_M^e table: 256 low entries and (M + 255) / 256 high ones, M <= 65536
constexpr int kFusedWmEntries = 256 + 256;

// Workgroup barrier for LDS data only: the LDS accesses before it complete and are seen
// by every wave after it.  Unlike __syncthreads it does not wait for global memory
// operations, so loads in flight (the row prefetch of the Rader kernel) stay in flight.
// The row kernels exchange no data through global memory within a workgroup.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Persistent launches (resident workgroups looping over the row groups, the tables
// loaded once) for the direct transform, whose rows are short: on p14 (16384^2) the
// row kernel went 0.98 -> 0.84 ms.  The Bluestein rows are long enough that the
// static share of rows to workgroups costs more at the tail than the launches save
// (c2: 0.378 -> 0.393 ms), so they keep one workgroup per row group.  So do the
// direct shapes whose loop spills more than 2 VGPRs at the 128-VGPR cap of 1024
// threads (the compiler's report): (M1, RPW, radices) below.
constexpr bool fused_persistent(int m1, int nb, int rpw, int r1, int r2, int r3, bool direct) {
    constexpr int no[11][5] = {{21, 4, 8, 28, 1},  {18, 9, 12, 9, 1},  {18, 5, 16, 12, 1},
                               {16, 4, 27, 9, 1},  {18, 4, 27, 9, 1},  {18, 4, 16, 16, 1},
                               {16, 2, 27, 18, 1}, {18, 2, 27, 18, 1}, {18, 2, 8, 8, 8},
                               {16, 1, 36, 27, 1}, {18, 1, 32, 32, 1}};
    for (const auto& s : no)
        if (m1 == s[0] && rpw == s[1] && r1 == s[2] && r2 == s[3] && r3 == s[4]) return false;
    return direct && nb == 1;
}

// LDS row layout.  Power-of-two stage radices access a row with lane strides of
// R entries (8 R bytes), which puts all lanes of a wave on a few LDS banks.  When
// P = 1 the row is stored padded, entry i at i + floor(i / 32) (row pitch
// L + L / 32), which spreads those strides over the banks.  Padding is used only
// where it costs no instructions: every LDS access of the stages is a
// per-thread base plus a compile-time offset, and x + y keeps that form in the
// padded layout iff (x mod 32) + (y mod 32) < 32 for all its bases x and offsets
// y -- checked at compile time for every stage of the shape (fused_pad_ok).
// P = 2 pads by one entry every 8 (entry i at i + floor(i / 8)), for radix-8 stages of
// rows whose length is a multiple of 8 but not of 32 (the Rader kernel); accesses that
// do not keep the base + offset form in that layout compute their index per element.
template <int P>
__host__ __device__ constexpr int lds_idx(int i) { return P == 2 ? i + (i >> 3) : P ? i + (i >> 5) : i; }
template <int L, int P>
constexpr int kRowPitch = P == 2 ? L + L / 8 : P ? L + L / 32 : L;

// For padding period m: does x + y keep the per-thread-base + compile-time-offset
// form ((x mod m) + (y mod m) < m for every base x and offset y) for the reads
// (write = false) or the writes of one stage?  Bases and offsets as in pad_stage_ok.
constexpr bool stage_split_ok(int L, int R, int NS, bool pivot, int m, bool write) {
    const int nb = L / R;
    int base = 0, off = 0;
    for (int j = 0; j < nb; j++) {
        const int b = !write ? j : pivot ? j * R : (j / NS) * NS * R + j % NS;
        base = b % m > base ? b % m : base;
    }
    for (int r = 0; r < R; r++) {
        const int o = !write ? r * nb : pivot ? r : r * NS;
        off = o % m > off ? o % m : off;
    }
    return base + off < m;
}

// (x mod 32) + (y mod 32) < 32 for the bases x and offsets y of one stage:
// reads j + r (L / R); writes ((j / NS) NS R + j % NS) + r NS, or (pivot: the
// inverse half) j R + r.
constexpr bool pad_stage_ok(int L, int R, int NS, bool pivot) {
    const int nb = L / R;
    int rd = 0, wr_base = 0, wr_off = 0, pv_base = 0;
    for (int r = 0; r < R; r++) {
        rd = (r * nb) % 32 > rd ? (r * nb) % 32 : rd;
        wr_off = (r * NS) % 32 > wr_off ? (r * NS) % 32 : wr_off;
    }
    int jmax = 0;
    for (int j = 0; j < nb; j++) {
        jmax = j % 32 > jmax ? j % 32 : jmax;
        const int b = (j / NS) * NS * R + j % NS;
        wr_base = b % 32 > wr_base ? b % 32 : wr_base;
        pv_base = (j * R) % 32 > pv_base ? (j * R) % 32 : pv_base;
    }
    const int pv_off = R - 1 < 31 ? R - 1 : 31;
    return jmax + rd < 32 && wr_base + wr_off < 32 && (!pivot || pv_base + pv_off % 32 < 32);
}
// every stage of the Bluestein convolution (or of the forward transform) of M2 = R1 R2 R3
constexpr bool fused_pad_ok(int R1, int R2, int R3) {
    const int L = R1 * R2 * R3;
    if (L % 32) return false;
    bool ok = pad_stage_ok(L, R1, 1, false);
    if (R2 > 1) ok = ok && pad_stage_ok(L, R2, R1, R3 == 1);
    if (R3 > 1) ok = ok && pad_stage_ok(L, R3, R1 * R2, true);
    if (R3 > 1) ok = ok && pad_stage_ok(L, R2, R3, false);  // inverse stages
    if (R2 > 1) ok = ok && pad_stage_ok(L, R1, R2 * R3, false);
    if (R2 == 1) ok = ok && pad_stage_ok(L, R1, 1, true);
    return ok;
}

constexpr size_t fused_lds_bytes_pitch(int m1, int nb, int rpw, int m2, size_t pitch) {
    // W_M2 table, the G = rpw * m1 / nb rows of a batch, the W_M table
    return ((size_t)m2 + pitch * (size_t)rpw * (m1 / nb) + kFusedWmEntries) * sizeof(float2);
}
// padded rows when the stages allow it and the padded rows still fit in 160 KiB
constexpr bool fused_pad(int m1, int nb, int rpw, int r1, int r2, int r3) {
    const int m2 = r1 * r2 * r3;
    return fused_pad_ok(r1, r2, r3) &&
           fused_lds_bytes_pitch(m1, nb, rpw, m2, (size_t)m2 + m2 / 32) <= 160 * 1024;
}
// r-major twiddle tables.  In a stage with NS R = L (the last forward one, and the
// last inverse one), k = j runs across the lanes and the twiddle W_L^{r k} read
// from W_L^i is a stride-r LDS access.  Stored r-major, entry (r - 1) NS + k, the
// lanes read consecutive entries at a compile-time offset per r.  Table A: the
// last forward stage (radix ra); table B: the last inverse stage (radix r1, the
// Bluestein convolution only; shared with A when ra == r1).  Used when they fit.
constexpr int fused_tp_radix(int r2, int r3) { return r3 > 1 ? r3 : r2 > 1 ? r2 : 0; }
constexpr int fused_tp_a(int r1, int r2, int r3) {
    const int ra = fused_tp_radix(r2, r3);
    return ra ? (ra - 1) * (r1 * r2 * r3 / ra) : 0;
}
constexpr int fused_tp_b(int r1, int r2, int r3) {
    const int ra = fused_tp_radix(r2, r3);
    return ra && ra != r1 ? (r1 - 1) * r2 * r3 : 0;
}
constexpr size_t fused_lds_bytes_base(int m1, int nb, int rpw, int r1, int r2, int r3) {
    const int m2 = r1 * r2 * r3;
    return fused_lds_bytes_pitch(m1, nb, rpw, m2,
                                 fused_pad(m1, nb, rpw, r1, r2, r3) ? (size_t)m2 + m2 / 32 : (size_t)m2);
}
constexpr bool fused_tp(int m1, int nb, int rpw, int r1, int r2, int r3) {
    return fused_tp_a(r1, r2, r3) > 0 &&
           fused_lds_bytes_base(m1, nb, rpw, r1, r2, r3) +
                   (size_t)(fused_tp_a(r1, r2, r3) + fused_tp_b(r1, r2, r3)) * sizeof(float2) <=
               160 * 1024;
}
constexpr size_t fused_lds_bytes_of(int m1, int nb, int rpw, int r1, int r2, int r3) {
    return fused_lds_bytes_base(m1, nb, rpw, r1, r2, r3) +
           (fused_tp(m1, nb, rpw, r1, r2, r3)
                ? (size_t)(fused_tp_a(r1, r2, r3) + fused_tp_b(r1, r2, r3)) * sizeof(float2)
                : 0);
}

// L2 prefetch of the pass-1 rows of the row group 128 workgroups later (it runs on the
// same XCD: workgroups go to the 8 XCDs in turn), by direct-to-LDS loads of one 4-byte
// word per 128-byte line into a 256-byte scratch past the shape's LDS that all waves
// share (never read).  For one-batch shapes launched one workgroup per row group, when
// the scratch fits without lowering the number of resident workgroups per CU.  Measured:
// c2's row kernel (17496 = 24 * 27 * 27) 0.383 -> 0.368 ms (0.358 at a distance of 128
// workgroups instead of 256, 0.362 at 64); the two-batch c3 Bluestein
// shape (34992) 1.73 -> 1.86 ms and the persistent direct p14 kernel 0.84 -> 0.89 ms, so
// those have none: their loads of B^ (or next row group) soon after wait for it.
#ifndef SPFFT_FUSED_AHEAD
#define SPFFT_FUSED_AHEAD 128
#endif
constexpr int kFusedAhead = SPFFT_FUSED_AHEAD;
constexpr size_t kPrefetchScratch = 64 * sizeof(float);
constexpr bool fused_pf(int m1, int nb, int rpw, int r1, int r2, int r3) {
    return nb == 1 && fused_lds_bytes_of(m1, nb, rpw, r1, r2, r3) + kPrefetchScratch <= 160 * 1024 &&
           (160 * 1024) / (fused_lds_bytes_of(m1, nb, rpw, r1, r2, r3) + kPrefetchScratch) ==
               (160 * 1024) / fused_lds_bytes_of(m1, nb, rpw, r1, r2, r3);
}
constexpr size_t fused_launch_lds(int m1, int nb, int rpw, int r1, int r2, int r3) {
    return fused_lds_bytes_of(m1, nb, rpw, r1, r2, r3) + (fused_pf(m1, nb, rpw, r1, r2, r3) ? kPrefetchScratch : 0);
}
// X0 offset (entries) of tile row r: r ldx, or (vrow_n1 > 0: the split transform's runs)
// run r % vrow_n1 of X0 row r / vrow_n1
__device__ __forceinline__ size_t x0_row(const RowPassArgs& a, int r) {
    if (a.vrow_n1 == 0) return (size_t)r * a.ldx;
    const int q = (int)__umulhi((uint32_t)r, a.vrow_magic);  // r / vrow_n1
    return (size_t)q * a.vrow_ld + (size_t)(r - q * a.vrow_n1) * a.C;
}

// rows [row0, row0 + nrows) of the tile's pass-1 buffer into L2; nw waves share the lines
__device__ __forceinline__ void prefetch_rows(const RowPassArgs& a, int row0, int nrows, int nw, float* scr) {
    if (row0 >= a.tile_rows) return;
    nrows = min(nrows, a.tile_rows - row0);
    const size_t base = x0_row(a, row0);
    const char* p = reinterpret_cast<const char*>(a.X0 + base);
    const int lines = (int)(((x0_row(a, row0 + nrows - 1) - base) * 8 + (size_t)a.C * 8 + 127) >> 7);
    const int lane = threadIdx.x & 63;
    for (int i = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); i * 64 < lines; i += nw) {
        const int line = i * 64 + lane;
        if (line < lines)
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(p + (size_t)line * 128),
                                             (__attribute__((address_space(3))) void*)scr, 4, 0, 0);
    }
}


// W_M^e, e = k1 n2 < M, from the two-level LDS table hi[e >> 8] * lo[e & 255].
__device__ __forceinline__ float2 wm_pow(const float2* s_wm, int e) {
    return cmul(s_wm[256 + (e >> 8)], s_wm[e & 255]);
}

// v[k1] times W_M^{k1 n2} (INV: its conjugate) for the thread's column n2, k1 < M1:
// with k1 = A i + j (j < A), W_M^{k1 n2} = W_M^{A i n2} W_M^{j n2}, so a column
// takes A + B - 2 table lookups and one product per k1 instead of a lookup per k1.
// (Applied in registers here rather than on the first and last LDS stage: 5%
// fewer VALU and 19% fewer LDS instructions, 7% faster on the c2 and c3 rows.)
template <int M1, bool INV>
__device__ __forceinline__ void fused_twiddle(float2* v, int n2, const float2* s_wm) {
    constexpr int A = M1 >= 49 ? 8 : M1 >= 25 ? 6 : M1 >= 9 ? 4 : 2;
    constexpr int B = (M1 + A - 1) / A;
    float2 lo[A], hi[B];
    lo[0] = hi[0] = make_float2(1.f, 0.f);
#pragma unroll
    for (int j = 1; j < A; j++) lo[j] = wm_pow(s_wm, j * n2);
#pragma unroll
    for (int i = 1; i < B; i++) hi[i] = wm_pow(s_wm, A * i * n2);
#pragma unroll
    for (int k1 = 1; k1 < M1; k1++) {
        const int i = k1 / A, j = k1 % A;
        const float2 w = i == 0 ? lo[j] : j == 0 ? hi[i] : cmul(hi[i], lo[j]);
        v[k1] = INV ? cmulc(v[k1], w) : cmul(v[k1], w);
    }
}

template <int R, int L, int NS, int MODE, int G, int KB, int NT, int P, int TPR = 0>
__device__ __forceinline__ void fused_stage(bio::rsrc_t r_bh, float2* buf, const float2* s_tw, int kbase,
                                            const float2* s_tp = nullptr) {
    constexpr int nb = L / R;
    constexpr int LP = kRowPitch<L, P>;
    constexpr int total = G * nb;
    constexpr int MB = (total + NT - 1) / NT;
    constexpr int step = L / (NS * R);
    constexpr bool INV = MODE == kInv;
    static_assert(MODE != kPivot || NS == nb, "the pivot stage is the last forward stage");
    static_assert(!TPR || step == 1, "the r-major table serves the stages with NS R = L");
    // (P = 1 shapes satisfy both at every stage: fused_pad_ok)
    constexpr bool RSPLIT = P != 2 || stage_split_ok(L, R, NS, MODE == kPivot, 8, false);
    constexpr bool WSPLIT = P != 2 || stage_split_ok(L, R, NS, MODE == kPivot, 8, true);
    float2 v[MB][R];
#pragma unroll
    for (int b = 0; b < MB; b++) {
        const int t = threadIdx.x + b * NT;
        if (t < total) {
            const int g = t / nb;
            const int j = t - g * nb;
            const float2* src = buf + g * LP + (RSPLIT ? lds_idx<P>(j) : 0);
#pragma unroll
            for (int r = 0; r < R; r++) v[b][r] = src[RSPLIT ? lds_idx<P>(r * nb) : lds_idx<P>(j + r * nb)];
        }
    }
    lds_barrier();
#pragma unroll
    for (int b = 0; b < MB; b++) {
        const int t = threadIdx.x + b * NT;
        if (t < total) {
            const int g = t / nb;
            const int j = t - g * nb;
            const int jq = j / NS;
            const int k = j - jq * NS;
            const int rowoff = (g % KB) * L;  // (k1 - kbase) * L
            if (NS > 1 && k != 0) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    const float2 w = TPR ? s_tp[(r - 1) * NS + k] : s_tw[r * k * step];
                    v[b][r] = INV ? cmulc(v[b][r], w) : cmul(v[b][r], w);
                }
            }
            RegFft<R>::run(v[b], INV);
            int base = jq * NS * R + k;
            int stride = NS;
            if (MODE == kPivot) {
                // B^ at the forward outputs j + r NS, then the inverse radix-R stage
                // with no earlier stages: outputs j R + r
#pragma unroll
                for (int r = 0; r < R; r++)
                    v[b][r] = cmul(v[b][r], bio::load2(r_bh, (rowoff + base) * 8, (kbase * L + r * NS) * 8));
                RegFft<R>::run(v[b], true);
                base = j * R;
                stride = 1;
            }
            float2* dst = buf + g * LP + (WSPLIT ? lds_idx<P>(base) : 0);
#pragma unroll
            for (int r = 0; r < R; r++)
                dst[WSPLIT ? lds_idx<P>(r * stride) : lds_idx<P>(base + r * stride)] = v[b][r];
        }
    }
    lds_barrier();
}

// The convolution of the batch's rows: forward stages R1 .. Rn, B^, inverse
// stages Rn .. R1 (radices of 1 are no stage), the last forward and first
// inverse stage merged (kPivot).
// (TP: the r-major tables s_ta / s_tb of the pivot and the last inverse stage)
template <int L, int G, int KB, int NT, int R1, int R2, int R3, int P, int TP>
__device__ __forceinline__ void fused_conv(bio::rsrc_t r_bh, float2* buf, const float2* s_tw, int kbase,
                                           const float2* s_ta, const float2* s_tb) {
    if constexpr (R3 > 1) {
        fused_stage<R1, L, 1, kFwd, G, KB, NT, P>(r_bh, buf, s_tw, kbase);
        fused_stage<R2, L, R1, kFwd, G, KB, NT, P>(r_bh, buf, s_tw, kbase);
        fused_stage<R3, L, R1 * R2, kPivot, G, KB, NT, P, TP>(r_bh, buf, s_tw, kbase, s_ta);
        fused_stage<R2, L, R3, kInv, G, KB, NT, P>(r_bh, buf, s_tw, kbase);
        fused_stage<R1, L, R3 * R2, kInv, G, KB, NT, P, TP>(r_bh, buf, s_tw, kbase, s_tb);
    } else if constexpr (R2 > 1) {
        fused_stage<R1, L, 1, kFwd, G, KB, NT, P>(r_bh, buf, s_tw, kbase);
        fused_stage<R2, L, R1, kPivot, G, KB, NT, P, TP>(r_bh, buf, s_tw, kbase, s_ta);
        fused_stage<R1, L, R2, kInv, G, KB, NT, P, TP>(r_bh, buf, s_tw, kbase, s_tb);
    } else {
        fused_stage<R1, L, 1, kPivot, G, KB, NT, P>(r_bh, buf, s_tw, kbase);
    }
}

// The forward stages only (the direct transform of a smooth C: no Bluestein).
template <int L, int G, int KB, int NT, int R1, int R2, int R3, int P, int TP>
__device__ __forceinline__ void fused_forward(bio::rsrc_t r_bh, float2* buf, const float2* s_tw,
                                              const float2* s_ta) {
    fused_stage<R1, L, 1, kFwd, G, KB, NT, P>(r_bh, buf, s_tw, 0);
    if constexpr (R3 > 1) {
        fused_stage<R2, L, R1, kFwd, G, KB, NT, P>(r_bh, buf, s_tw, 0);
        fused_stage<R3, L, R1 * R2, kFwd, G, KB, NT, P, TP>(r_bh, buf, s_tw, 0, s_ta);
    } else if constexpr (R2 > 1) {
        fused_stage<R2, L, R1, kFwd, G, KB, NT, P, TP>(r_bh, buf, s_tw, 0, s_ta);
    }
}

// DIRECT: the row transform is one length-C FFT (M = C = M1 M2, a shape with the
// whole row in LDS), with no Bluestein chirps: phase A as below on all M1 inputs,
// phase B the forward stages only, then each thread reads its LDS column n2 back
// -- entries X[k1 + M1 n2], k1 < M1 -- and writes those of row u (v <= C/2) and of
// its mirror row R - u (X[R-u][C - w] = conj(X[u][w])).  The column pass then runs
// with h = 1.
template <int M1, int NB, int RPW, int NT, int R1, int R2, int R3, bool DIRECT, bool PF>
__device__ __forceinline__ void rowfused_body(const RowPassArgs& a, const int bx, const int gx) {
    constexpr int M2 = R1 * R2 * R3;
    constexpr int KB = M1 / NB;
    constexpr int G = RPW * KB;
    constexpr int CPT = (RPW * M2 + NT - 1) / NT;  // columns n2 per thread
    static_assert(M1 % NB == 0, "batches must divide M1");
    static_assert(RPW == 1 || CPT == 1, "several rows per workgroup: one column per thread");
    static_assert(R1 > 1 && (R2 > 1 || R3 == 1), "stage radices first");
    static_assert(!DIRECT || NB == 1, "the direct transform holds the whole row in LDS");
    // Bluestein: M >= 2C - 1, so the entries w = n2 + M2 n1 < C have n1 < NZ: only
    // those inputs can be non-zero and only those outputs are written (the codelets
    // are pruned to them at compile time; this file is built without signed
    // zeros, so x + 0 folds to x).  DIRECT: every entry.
    constexpr int NZ = DIRECT ? M1 : (M1 + 1) / 2;
    constexpr int P = fused_pad(M1, NB, RPW, R1, R2, R3) ? 1 : 0;  // padded LDS rows
    constexpr int LP = kRowPitch<M2, P>;
    constexpr int TP = fused_tp(M1, NB, RPW, R1, R2, R3) ? 1 : 0;   // r-major twiddle tables
    constexpr int RA = fused_tp_radix(R2, R3), NA = TP ? fused_tp_a(R1, R2, R3) : 0;
    constexpr int NBT = TP ? fused_tp_b(R1, R2, R3) : 0;
    extern __shared__ float2 smem[];
    float2* s_tw = smem;                    // W_M2^i, i < M2
    float2* s_wm = s_tw + M2;               // W_M^e: 256 low + (M + 255) / 256 high entries
    float2* s_ta = s_wm + kFusedWmEntries;  // TP: W_M2^{r k} at (r - 1) (M2 / RA) + k
    float2* s_tb = NBT ? s_ta + NA : s_ta;  // TP: the same for radix R1 (when R1 != RA)
    float2* buf = s_ta + NA + NBT;          // G rows of length M2: [row of workgroup][k1 of batch][n2]
    const int t = threadIdx.x;
    const bio::rsrc_t r_bh = bio::make_rsrc(a.bhat, (uint32_t)a.M * 8u);
    // the twiddle tables, once per workgroup
    auto fill_tables = [&]() {
        for (int i = t; i < M2; i += NT) s_tw[i] = a.tw2[i];
        for (int i = t; i < 256 + (a.M + 255) / 256; i += NT) s_wm[i] = a.wm2[i];
        if constexpr (TP) {
            constexpr int nsa = NA ? M2 / RA : 1, nsb = NBT ? M2 / R1 : 1;
            for (int i = t; i < NA; i += NT) s_ta[i] = a.tw2[(i / nsa + 1) * (i % nsa)];
            for (int i = t; i < NBT; i += NT) s_tb[i] = a.tw2[(i / nsb + 1) * (i % nsb)];
        }
    };

    // One row group (the RPW rows from tile row row0); fill: the first one of the
    // workgroup, which fills the tables once its loads are out.
    auto row_group = [&](const int row0, const bool fill) {
    // the thread's columns n2 (row g of the workgroup's RPW rows)
    int n2[CPT], g[CPT];
    bool live[CPT];
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int q = t + c * NT;
        g[c] = RPW == 1 ? 0 : q / M2;
        n2[c] = q - g[c] * M2;
        live[c] = (RPW == 1 ? q < M2 : g[c] < RPW) && row0 + g[c] < a.tile_rows;
    }

    // phase A: the loads go out first, then the tables are filled (first row group);
    // the codelet's twiddles read W_M from LDS after a barrier, which also ends the
    // previous row group's reads of buf
    float2 v[CPT][M1];
    {
        const int nrows = min(RPW, a.tile_rows - row0);
        // RPW == 1: the row as a buffer of C entries, so reads past it return 0 (the
        // zero padding) with the stride M2 n1 as the scalar offset
        const size_t xb = x0_row(a, row0);
        const bio::rsrc_t r_x = bio::make_rsrc(
            a.X0 + xb, (uint32_t)(RPW == 1 ? a.C : x0_row(a, row0 + nrows - 1) - xb + a.C) * 8u);
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            if (live[c]) {
                const int rb = (int)(x0_row(a, row0 + g[c]) - xb);
#pragma unroll
                for (int n1 = 0; n1 < M1; n1++) {
                    if (n1 >= NZ) {
                        v[c][n1] = make_float2(0.f, 0.f);
                    } else if constexpr (RPW == 1) {
                        v[c][n1] = bio::load2_stream(r_x, n2[c] * 8, M2 * n1 * 8);
                    } else {
                        const int w = n2[c] + M2 * n1;
                        v[c][n1] = bio::load2_stream(r_x, w < a.C ? (rb + w) * 8 : bio::kOob, 0);
                    }
                }
            }
        }
        if (fill) fill_tables();
        lds_barrier();
        if constexpr (PF && !fused_persistent(M1, NB, RPW, R1, R2, R3, DIRECT)) {
            // after the row group's own loads; the barriers below are LDS-only, so it stays
            // in flight until phase B's loads of B^ wait for it
            prefetch_rows(a, row0 + kFusedAhead * RPW, RPW, NT / 64,
                          reinterpret_cast<float*>(reinterpret_cast<char*>(smem) +
                                                   fused_lds_bytes_of(M1, NB, RPW, R1, R2, R3)));
        }
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            if (live[c]) {
                RegFft<M1>::run(v[c], false);
                fused_twiddle<M1, false>(v[c], n2[c], s_wm);
            }
        }
    }

    // phase B
    cx::static_for<NB>([&](auto Bc) {
        constexpr int b = decltype(Bc)::value;
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            if (live[c]) {
                float2* col = buf + g[c] * KB * LP + lds_idx<P>(n2[c]);
#pragma unroll
                for (int k = 0; k < KB; k++) {
                    if constexpr (b == 0) {
                        col[k * LP] = v[c][k];
                    } else {
                        const float2 done = col[k * LP];
                        col[k * LP] = v[c][b * KB + k];
                        v[c][(b - 1) * KB + k] = done;
                    }
                }
            }
        }
        lds_barrier();
        if constexpr (DIRECT)
            fused_forward<M2, G, KB, NT, R1, R2, R3, P, TP>(r_bh, buf, s_tw, s_ta);
        else
            fused_conv<M2, G, KB, NT, R1, R2, R3, P, TP>(r_bh, buf, s_tw, b * KB, s_ta, s_tb);
    });

    // phase C
    const bio::rsrc_t r_h = bio::make_rsrc(a.h, (uint32_t)a.C * 8u);  // reads past C return 0
    const int u0 = a.u_base + row0;
    if constexpr (DIRECT) {
        // lanes take consecutive n2: conflict-free LDS reads; a lane's M1 outputs are
        // M1 consecutive entries of the row (the L2 merges the lanes' partial lines)
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            if (!live[c]) continue;
            const float2* col = buf + g[c] * KB * LP + lds_idx<P>(n2[c]);
            const int u = u0 + g[c];
            const bool mirror = u > 0 && a.R - u > u;
            float2* o = a.out + (size_t)u * a.P;
            float2* om = a.out_m + (size_t)(mirror ? a.R - u : 0) * a.P;
#pragma unroll
            for (int k1 = 0; k1 < M1; k1++) {
                const float2 x = col[k1 * LP];
                const int w = k1 + M1 * n2[c];
                if (w < a.V) o[w] = x;
                const int vm = w == 0 ? 0 : a.C - w;
                if (mirror && vm < a.V) om[vm] = make_float2(x.x, -x.y);
            }
        }
    } else if constexpr (RPW == 1) {
        // Row u as a buffer of C/2 + 1 entries: writes of v > C/2 are dropped.  Mirror
        // row R - u: X[R-u][C - w] = conj(X[u][w]) for 0 < w < C, at offset
        // M - 1 - w = (M2 (M1 - 1 - n1) + (M2 - 1 - n2)) from a base M - 1 - C entries
        // before the row, a buffer ending at entry C/2 of the row: writes of
        // C - w > C/2 are dropped, those of w >= C take kOob, w = 0 is written apart.
        const int u = u0;
        const bool mirror = u > 0 && a.R - u > u;
        const int dm = a.M - 1 - a.C;
        const bio::rsrc_t r_o = bio::make_rsrc(a.out + (size_t)u * a.P, (uint32_t)a.V * 8u);
        const bio::rsrc_t r_m = bio::make_rsrc(mirror ? a.out_m + (size_t)(a.R - u) * a.P - dm : a.out,
                                                   mirror ? (uint32_t)(dm + a.V) * 8u : 0u);
        const bio::rsrc_t r_m0 = bio::make_rsrc(mirror ? a.out_m + (size_t)(a.R - u) * a.P : a.out,
                                                    mirror ? 8u : 0u);
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            if (live[c]) {
                const float2* col = buf + lds_idx<P>(n2[c]);
#pragma unroll
                for (int k = 0; k < KB; k++) v[c][(NB - 1) * KB + k] = col[k * LP];
                fused_twiddle<M1, true>(v[c], n2[c], s_wm);
                RegFft<M1>::run(v[c], true);
                const int n1c = (a.C - n2[c] + M2 - 1) / M2;  // columns w = n2 + M2 n1 < C
                const int vm0 = (M2 - 1 - n2[c]) * 8;
#pragma unroll
                for (int n1 = 0; n1 < NZ; n1++) {
                    const float2 x = cmul(v[c][n1], bio::load2(r_h, n2[c] * 8, M2 * n1 * 8));
                    bio::store2_stream(x, r_o, n2[c] * 8, M2 * n1 * 8);
                    bio::store2_stream(make_float2(x.x, -x.y), r_m, n1 < n1c ? vm0 : bio::kOob,
                                  M2 * (M1 - 1 - n1) * 8);
                    if (n1 == 0 && n2[c] == 0) bio::store2_stream(make_float2(x.x, -x.y), r_m0, 0, 0);
                }
            }
        }
    } else {
        // one descriptor for the workgroup's output rows u0 .. u0 + nrows - 1 and one for
        // their mirror rows R - u (the lowest is mlo); explicit bounds
        const int nrows = min(RPW, a.tile_rows - row0);
        const int mlo = a.R - (u0 + nrows - 1);
        const bio::rsrc_t r_o = bio::make_rsrc(a.out + (size_t)u0 * a.P, (uint32_t)(nrows * a.P) * 8u);
        const bio::rsrc_t r_m = bio::make_rsrc(a.out_m + (size_t)mlo * a.P, (uint32_t)(nrows * a.P) * 8u);
        if (live[0]) {
            const float2* col = buf + g[0] * KB * LP + lds_idx<P>(n2[0]);
#pragma unroll
            for (int k = 0; k < KB; k++) v[0][(NB - 1) * KB + k] = col[k * LP];
            fused_twiddle<M1, true>(v[0], n2[0], s_wm);
            RegFft<M1>::run(v[0], true);
            const int u = u0 + g[0];
            const bool mirror = u > 0 && a.R - u > u;
            const int ob = g[0] * a.P;                          // row u in r_o
            const int mb = mirror ? (a.R - u - mlo) * a.P : 0;  // row R - u in r_m
#pragma unroll
            for (int n1 = 0; n1 < NZ; n1++) {
                const int w = n2[0] + M2 * n1;
                const bool in = w < a.C;
                const float2 x = cmul(v[0][n1], bio::load2(r_h, n2[0] * 8, M2 * n1 * 8));
                bio::store2_stream(x, r_o, w < a.V ? (ob + w) * 8 : bio::kOob, 0);
                const int vm = w == 0 ? 0 : a.C - w;
                bio::store2_stream(make_float2(x.x, -x.y), r_m, in && mirror && vm < a.V ? (mb + vm) * 8 : bio::kOob, 0);
            }
        }
    }
    };  // row_group

    if constexpr (fused_persistent(M1, NB, RPW, R1, R2, R3, DIRECT)) {
        // persistent launch: the resident workgroups take the row groups blockIdx.x,
        // blockIdx.x + gridDim.x, ... and keep the tables in LDS between them
        const int nblk = (a.tile_rows + RPW - 1) / RPW;
        for (int blk = bx; blk < nblk; blk += gx) row_group(blk * RPW, blk == bx);
    } else {
        row_group(bx * RPW, true);
    }
    // the prefetch writes this workgroup's LDS: it lands before the workgroup ends
    if constexpr (PF) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

template <int M1, int NB, int RPW, int NT, int R1, int R2, int R3, bool DIRECT>
__global__ __launch_bounds__(NT) void k_rowfused(RowPassArgs a) {
    rowfused_body<M1, NB, RPW, NT, R1, R2, R3, DIRECT, fused_pf(M1, NB, RPW, R1, R2, R3)>(a, blockIdx.x, gridDim.x);
}

// Batched form: work item (matrix descriptor, row group) per workgroup, one row group
// each (no persistent loop).
template <int M1, int NB, int RPW, int NT, int R1, int R2, int R3, bool DIRECT>
__global__ __launch_bounds__(NT) void k_rowfused_batch(const RowPassArgs* __restrict__ descs,
                                                       const int2* __restrict__ work) {
    const int2 w = work[blockIdx.x];
    rowfused_body<M1, NB, RPW, NT, R1, R2, R3, DIRECT, false>(descs[w.x], w.y, 0x40000000);
}

// Instantiated shapes: X(M1, NB, RPW, NT, R1, R2, R3), M = M1 * R1 * R2 * R3 (a
// radix of 1 is no stage), NT threads per workgroup, at most 160 KiB of LDS.
// Group A: the c3 lengths and the power-of-two convolution lengths of its
// columns; group B: the c1/c2 lengths; groups C-H: a ladder over the 3-smooth
// lengths from 1024 to 32768 (one workgroup holds RPW rows of M <= 20000 in one
// LDS batch, longer rows in NB batches).
//   34992 = 36 * 972   (3-smooth length of 2 * 16706 - 1)
//   33600 = 35 * 960   (7-smooth length of 2 * 16706 - 1)
//   36864 = 36 * 1024, 35840 = 35 * 1024 (power-of-two convolution lengths)
//   17496 = 24 * 729, 16464 = 21 * 784, 5184 = 16 * 324 (3 rows), 4704 = 21 * 224 (4 rows)
//   (Direct shapes for c1's width 2310 = 35 * (6 * 11) = 22 * (7 * 15), 4 rows, with the
//   radix-11 butterfly, measured no faster than Bluestein at 5184 on c1: row kernel 0.043
//   and 0.037 ms against 0.040; c1's 1156 short rows are latency-bound.  Not kept.)
#define SPFFT_FUSED_GROUP_A(X)                                                              \
    X(36, 2, 1, 512, 9, 12, 9) X(35, 5, 1, 512, 8, 12, 10) X(36, 3, 1, 512, 16, 8, 8)     \
    X(35, 5, 1, 512, 16, 8, 8)
#define SPFFT_FUSED_GROUP_B(X)                                                                  \
    X(24, 1, 1, 768, 27, 27, 1) X(21, 1, 1, 1024, 28, 28, 1) X(16, 1, 3, 1024, 18, 18, 1) \
    X(21, 1, 4, 1024, 8, 28, 1)
//   1024 .. 1944
#define SPFFT_FUSED_GROUP_C(X)                                                                     \
    X(16, 1, 16, 1024, 8, 8, 1) X(16, 1, 14, 1024, 9, 8, 1) X(16, 1, 12, 1024, 9, 9, 1)          \
    X(18, 1, 12, 1024, 9, 9, 1) X(16, 1, 10, 1024, 12, 8, 1) X(16, 1, 9, 1024, 12, 9, 1)         \
    X(18, 1, 9, 1024, 12, 9, 1)
//   2048 .. 3888
#define SPFFT_FUSED_GROUP_D(X)                                                                     \
    X(16, 1, 8, 1024, 16, 8, 1) X(16, 1, 7, 1024, 12, 12, 1) X(16, 1, 6, 1024, 18, 9, 1)         \
    X(18, 1, 6, 1024, 18, 9, 1) X(16, 1, 5, 1024, 16, 12, 1) X(18, 1, 5, 1024, 16, 12, 1)        \
    X(16, 1, 4, 1024, 27, 9, 1)
//   4096 .. 7776
#define SPFFT_FUSED_GROUP_E(X)                                                                     \
    X(16, 1, 4, 1024, 16, 16, 1) X(18, 1, 4, 1024, 27, 9, 1) X(18, 1, 4, 1024, 16, 16, 1)        \
    X(18, 1, 3, 1024, 18, 18, 1) X(24, 1, 3, 768, 16, 16, 1) X(27, 1, 2, 512, 16, 16, 1)         \
    X(16, 1, 2, 1024, 27, 18, 1)
//   8192 .. 13122
#define SPFFT_FUSED_GROUP_F(X)                                                                     \
    X(16, 1, 2, 1024, 8, 8, 8) X(18, 1, 2, 1024, 27, 18, 1) X(18, 1, 2, 1024, 8, 8, 8)           \
    X(16, 1, 1, 704, 27, 24, 1) X(18, 1, 1, 704, 27, 24, 1) X(16, 1, 1, 768, 32, 24, 1)          \
    X(18, 1, 1, 768, 27, 27, 1)
//   13824 .. 23328
#define SPFFT_FUSED_GROUP_G(X)                                                                     \
    X(18, 1, 1, 768, 32, 24, 1) X(16, 1, 1, 1024, 36, 27, 1) X(16, 1, 1, 1024, 32, 32, 1)        \
    X(18, 1, 1, 1024, 32, 32, 1) X(24, 2, 1, 896, 8, 12, 9) X(24, 2, 1, 512, 9, 12, 9)
//   24576 .. 32768
#define SPFFT_FUSED_GROUP_H(X)                                                                     \
    X(24, 2, 1, 512, 16, 8, 8) X(27, 3, 1, 512, 9, 12, 9) X(27, 3, 1, 512, 16, 8, 8)             \
    X(32, 2, 1, 512, 9, 12, 9) X(32, 4, 1, 512, 16, 8, 8)
#define SPFFT_FUSED_LIST(X)                                                                        \
    SPFFT_FUSED_GROUP_A(X) SPFFT_FUSED_GROUP_B(X) SPFFT_FUSED_GROUP_C(X) SPFFT_FUSED_GROUP_D(X)     \
    SPFFT_FUSED_GROUP_E(X) SPFFT_FUSED_GROUP_F(X) SPFFT_FUSED_GROUP_G(X) SPFFT_FUSED_GROUP_H(X)


// Each returns hipErrorInvalidValue when the shape is not in its group.
// (batch: work items of k_rowfused_batch instead of one matrix's launch)
struct FusedBatch {
    const RowPassArgs* descs;
    const int2* work;
    int n;
};
hipError_t launch_fused_group_a(const RowPassArgs& a, hipStream_t s, bool prepare_only,
                                const FusedBatch* batch = nullptr);
hipError_t launch_fused_group_b(const RowPassArgs& a, hipStream_t s, bool prepare_only,
                                const FusedBatch* batch = nullptr);
hipError_t launch_fused_group_c(const RowPassArgs& a, hipStream_t s, bool prepare_only,
                                const FusedBatch* batch = nullptr);
hipError_t launch_fused_group_d(const RowPassArgs& a, hipStream_t s, bool prepare_only,
                                const FusedBatch* batch = nullptr);
hipError_t launch_fused_group_e(const RowPassArgs& a, hipStream_t s, bool prepare_only,
                                const FusedBatch* batch = nullptr);
hipError_t launch_fused_group_f(const RowPassArgs& a, hipStream_t s, bool prepare_only,
                                const FusedBatch* batch = nullptr);
hipError_t launch_fused_group_g(const RowPassArgs& a, hipStream_t s, bool prepare_only,
                                const FusedBatch* batch = nullptr);
hipError_t launch_fused_group_h(const RowPassArgs& a, hipStream_t s, bool prepare_only,
                                const FusedBatch* batch = nullptr);

// (a.direct: the direct-transform instantiation, single-batch shapes only; a.fused_cus > 0:
// persistent launch, as many workgroups as can be resident on that many CUs, each looping
// over row groups)
#define SPFFT_FUSED_LAUNCH_CASE(m1, nb, rpw, nt, r1, r2, r3)                                       \
    if (a.M1 == m1 && a.fused_nb == nb && a.fused_rpw == rpw && a.fused_nt == nt &&                 \
        a.f2.radix[0] == r1 && (a.f2.nstages > 1 ? a.f2.radix[1] : 1) == r2 &&                      \
        (a.f2.nstages > 2 ? a.f2.radix[2] : 1) == r3) {                                             \
        constexpr size_t lds = fused_lds_bytes_of(m1, nb, rpw, r1, r2, r3);                         \
        constexpr size_t lds_k = fused_launch_lds(m1, nb, rpw, r1, r2, r3);                         \
        auto* k = &k_rowfused<m1, nb, rpw, nt, r1, r2, r3, false>;                                  \
        if (a.direct) {                                                                             \
            if constexpr (nb == 1) k = &k_rowfused<m1, nb, rpw, nt, r1, r2, r3, nb == 1>;           \
            else return hipErrorInvalidValue;                                                       \
        }                                                                                           \
        if (prepare_only) {                                                                         \
            auto* kb = &k_rowfused_batch<m1, nb, rpw, nt, r1, r2, r3, false>;                       \
            if (a.direct) {                                                                         \
                if constexpr (nb == 1) kb = &k_rowfused_batch<m1, nb, rpw, nt, r1, r2, r3, nb == 1>;  \
            }                                                                                       \
            hipError_t e0 = hipFuncSetAttribute(reinterpret_cast<const void*>(kb),                  \
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
            if (e0 != hipSuccess) return e0;                                                        \
            return hipFuncSetAttribute(reinterpret_cast<const void*>(k),                            \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_k);     \
        }                                                                                           \
        if (batch) {                                                                                \
            auto* kb = &k_rowfused_batch<m1, nb, rpw, nt, r1, r2, r3, false>;                       \
            if (a.direct) {                                                                         \
                if constexpr (nb == 1) kb = &k_rowfused_batch<m1, nb, rpw, nt, r1, r2, r3, nb == 1>;  \
            }                                                                                       \
            hipLaunchKernelGGL(kb, dim3(batch->n), dim3(nt), lds, s, batch->descs, batch->work);     \
            return hipGetLastError();                                                               \
        }                                                                                           \
        const int nblk = (a.tile_rows + rpw - 1) / rpw;                                             \
        int grid = nblk;                                                                            \
        if (a.fused_cus > 0 && fused_persistent(m1, nb, rpw, r1, r2, r3, a.direct != 0)) {         \
            /* resident workgroups per CU, per form (any host thread may fill it: the same value) */ \
            static std::atomic<int> occ[2] = {{0}, {0}};                                            \
            int o = occ[a.direct ? 1 : 0].load(std::memory_order_relaxed);                          \
            if (!o) {                                                                               \
                int n = 0;                                                                          \
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, nt, lds_k) != hipSuccess)   \
                    n = 1;                                                                          \
                o = std::max(n, 1);                                                                 \
                occ[a.direct ? 1 : 0].store(o, std::memory_order_relaxed);                          \
            }                                                                                       \
            grid = std::min(nblk, a.fused_cus * o);                                                 \
        }                                                                                           \
        hipLaunchKernelGGL(k, dim3(grid), dim3(nt), lds_k, s, a);                                   \
        return hipGetLastError();                                                                   \
    }

}  // namespace spfft
