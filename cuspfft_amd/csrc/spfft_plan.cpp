// spfft_plan.cpp -- host planner and C ABI (include/spfft_c.h).
//
// Planning mirrors the reference drivers' set-up (src/sparse_fft_bluestein.cu:
// 2256-2320 for --csc-mixed-radix, :2557-2642 for --csc-cufft-smooth):
// smooth Bluestein length M, chirp h, kernel b and B^ = DFT_M(b).  Here every
// table is produced in FP64 on the host with exact integer reduction of the
// angle (k^2 mod 2C, k*n mod M) and rounded once to FP32; B^ comes from an
// FP64 FFT and has the 1/M scale folded in.  The device work (CSC build, pass
// 1, row pass) is enqueued on the caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <memory>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/spfft_c.h"
#include "spfft_kernels.h"
#include "spfft_types.h"
#include "spfft_plan_internal.h"

using spfft::FftStages;
using spfft::RowPassArgs;

namespace spfft_internal {

thread_local std::string g_last_error;

spfft_status fail(spfft_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}

spfft_status hip_fail(hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? SPFFT_ERR_ALLOC : SPFFT_ERR_HIP,
                std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

// The caller's options (NULL: the defaults), rejected when the caller's struct has
// another size than this library's.
spfft_status read_opts(const spfft_opts* in, spfft_opts* out) {
    if (!in) {
        spfft_opts_default(out);
        return SPFFT_OK;
    }
    if (in->struct_size != (int)sizeof(spfft_opts))
        return fail(SPFFT_ERR_INVALID_ARG, "spfft_opts.struct_size is " + std::to_string(in->struct_size) +
                                               ", this library's spfft_opts has " +
                                               std::to_string(sizeof(spfft_opts)) +
                                               " bytes (initialise it with spfft_opts_default)");
    *out = *in;
    return SPFFT_OK;
}

spfft_status check_info(const spfft_plan_info* info) {
    if (!info) return fail(SPFFT_ERR_INVALID_ARG, "NULL info");
    if (info->struct_size != (int)sizeof(spfft_plan_info))
        return fail(SPFFT_ERR_INVALID_ARG, "spfft_plan_info.struct_size is " + std::to_string(info->struct_size) +
                                               ", this library's spfft_plan_info has " +
                                               std::to_string(sizeof(spfft_plan_info)) + " bytes");
    return SPFFT_OK;
}

// Checks the caller's result struct and clears it (struct_size kept).
spfft_status check_result(spfft_result* res) {
    if (!res) return fail(SPFFT_ERR_INVALID_ARG, "NULL result");
    if (res->struct_size != (int)sizeof(spfft_result))
        return fail(SPFFT_ERR_INVALID_ARG, "spfft_result.struct_size is " + std::to_string(res->struct_size) +
                                               ", this library's spfft_result has " +
                                               std::to_string(sizeof(spfft_result)) + " bytes");
    std::memset(res, 0, sizeof(*res));
    res->struct_size = (int)sizeof(spfft_result);
    return SPFFT_OK;
}

}  // namespace spfft_internal

using namespace spfft_internal;

namespace {

const double kPi = 3.14159265358979323846264338327950288;
typedef std::complex<double> cd;

// exp(sign * 2 pi i * k / n), k reduced exactly mod n first.
cd unit_root(long long k, long long n, int sign) {
    long long kk = k % n;
    if (kk < 0) kk += n;
    const double a = 2.0 * kPi * (double)kk / (double)n;
    return cd(std::cos(a), sign * std::sin(a));
}

float2 to_f2(cd z) { return make_float2((float)z.real(), (float)z.imag()); }

// Round a float to the nearest BF16 value (ties to even), returned as a float.
float bf16_round(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    b += 0x7FFFu + ((b >> 16) & 1u);
    b &= 0xFFFF0000u;
    float r;
    std::memcpy(&r, &b, 4);
    return r;
}

int next_smooth(int n, const int* primes, int np) {
    if (n < 1) n = 1;
    for (int m = n;; m++) {
        int x = m;
        for (int i = 0; i < np; i++)
            while (x % primes[i] == 0) x /= primes[i];
        if (x == 1) return m;
    }
}

bool is_7_smooth(int n) {
    if (n < 1) return false;
    for (int p : {2, 3, 5, 7})
        while (n % p == 0) n /= p;
    return n == 1;
}

// Radix sequence for a 7-smooth length: 8s, then one 4 or 2, then 7s, 5s, 3s.
bool plan_stages(int L, FftStages& f) {
    f.L = L;
    f.nstages = 0;
    int n = L, c2 = 0;
    while (n % 2 == 0) { n /= 2; c2++; }
    std::vector<int> r;
    while (c2 >= 3) { r.push_back(8); c2 -= 3; }
    if (c2 == 2) r.push_back(4);
    if (c2 == 1) r.push_back(2);
    for (int p : {7, 5, 3})
        while (n % p == 0) { r.push_back(p); n /= p; }
    if (n != 1 || (int)r.size() > spfft::kMaxStages) return false;
    for (size_t i = 0; i < r.size(); i++) f.radix[i] = r[i];
    f.nstages = (int)r.size();
    return true;
}

// FP64 Stockham FFT (same autosort form as the device kernels), any n: radices
// 8, 4, 2, 7, 5, 3, then every remaining prime factor as a direct DFT stage.
void fft64(std::vector<cd>& x, int sign) {
    const int n = (int)x.size();
    if (n <= 1) return;
    std::vector<int> rad;
    {
        int m = n;
        for (int p : {8, 4, 2, 7, 5, 3})
            while (m % p == 0) { rad.push_back(p); m /= p; }
        for (int p = 11; m > 1; p += 2) {
            if ((long long)p * p > m) p = m;
            while (m % p == 0) { rad.push_back(p); m /= p; }
        }
    }
    std::vector<cd> y(n), tw(n);
    for (int i = 0; i < n; i++) tw[i] = unit_root(i, n, sign);
    int Ns = 1;
    for (int R : rad) {
        const int nb = n / R, step = n / (Ns * R);
        std::vector<cd> F((size_t)R * R), v(R), o(R);
        for (int a = 0; a < R; a++)
            for (int b = 0; b < R; b++) F[(size_t)a * R + b] = unit_root((long long)a * b, R, sign);
        for (int j = 0; j < nb; j++) {
            const int k = j % Ns;
            for (int r = 0; r < R; r++) v[r] = x[j + r * nb] * tw[(long long)r * k * step % n];
            for (int p = 0; p < R; p++) {
                cd acc = 0;
                for (int q = 0; q < R; q++) acc += F[(size_t)p * R + q] * v[q];
                o[p] = acc;
            }
            const int d = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; r++) y[d + r * Ns] = o[r];
        }
        x.swap(y);
        Ns *= R;
    }
}

bool is_odd_prime(long long n) {
    if (n < 3 || n % 2 == 0) return false;
    for (long long d = 3; d * d <= n; d += 2)
        if (n % d == 0) return false;
    return true;
}

long long pow_mod(long long b, long long e, long long m) {
    long long r = 1 % m;
    b %= m;
    for (; e > 0; e >>= 1, b = b * b % m)
        if (e & 1) r = r * b % m;
    return r;
}

// a^-1 mod m for gcd(a, m) = 1 (extended Euclid); 0 for m = 1.
long long inv_mod(long long a, long long m) {
    long long r0 = m, r1 = ((a % m) + m) % m, s0 = 0, s1 = 1;
    while (r1 != 0) {
        const long long q = r0 / r1, r2 = r0 - q * r1, s2 = s0 - q * s1;
        r0 = r1; r1 = r2; s0 = s1; s1 = s2;
    }
    return m == 1 ? 0 : ((s0 % m) + m) % m;
}

// The smallest primitive root g of the prime p: g^((p-1)/q) != 1 for every prime q | p - 1.
long long primitive_root(long long p) {
    std::vector<long long> qs;
    long long n = p - 1;
    for (long long q = 2; q * q <= n; q++)
        if (n % q == 0) {
            qs.push_back(q);
            while (n % q == 0) n /= q;
        }
    if (n > 1) qs.push_back(n);
    for (long long g = 2;; g++) {
        bool ok = true;
        for (long long q : qs) ok = ok && pow_mod(g, (p - 1) / q, p) != 1;
        if (ok) return g;
    }
}

// Rader tables for a prime p with L = p - 1 = m1 m2 (gcd(m1, m2) = 1) in the prime-factor
// order of the kernels (rowpass_rader.hip, rowpass_raderruns.hip): with g a primitive
// root of p, entry (n1, n2) of a convolution operand (position n2 + m2 n1) is a[q], q =
// (m2 n1 + m1 n2) mod L; gpos[n2 + m2 n1] = g^q mod p (the input index it holds),
// kt[n2 + m2 n1] = g^-q mod p (the output index it gives); b[q] = W_p^{g^-q}, and
// frequency (k1, k2) of B^ = DFT_L(b) / L, at bh[k1 m2 + k2], is (c1 k1 + c2 k2) mod L
// with c1 = m2 (m2^-1 mod m1), c2 = m1 (m1^-1 mod m2).
void rader_tables(long long pr, long long m1, long long m2, std::vector<int>& gpos, std::vector<int>& kt,
                  std::vector<float2>& bh) {
    const long long L = pr - 1, g = primitive_root(pr), ginv = pow_mod(g, pr - 2, pr);
    std::vector<int> gpow((size_t)L), gneg((size_t)L);
    std::vector<cd> b((size_t)L);
    long long gm = 1, gq = 1;
    for (long long q = 0; q < L; q++) {
        gneg[q] = (int)gm;
        gpow[q] = (int)gq;
        b[q] = unit_root(gm, pr, -1);
        gm = gm * ginv % pr;
        gq = gq * g % pr;
    }
    gpos.assign((size_t)L, 0);
    kt.assign((size_t)L, 0);
    for (long long n1 = 0; n1 < m1; n1++)
        for (long long n2 = 0; n2 < m2; n2++) {
            const size_t pos = (size_t)(n2 + m2 * n1);
            const long long q = (m2 * n1 + m1 * n2) % L;
            kt[pos] = gneg[q];
            gpos[pos] = gpow[q];
        }
    fft64(b, -1);
    const long long c1 = m2 * inv_mod(m2, m1) % L, c2 = m1 * inv_mod(m1, m2) % L;
    const double inv_l = 1.0 / (double)L;
    bh.assign((size_t)L, make_float2(0.f, 0.f));
    for (long long k1 = 0; k1 < m1; k1++)
        for (long long k2 = 0; k2 < m2; k2++)
            bh[(size_t)(k1 * m2 + k2)] = to_f2(b[(size_t)((c1 * k1 + c2 * k2) % L)] * inv_l);
}

size_t g_lds_limit = 0;
int g_cu_count = 0;  // compute units of the device (persistent fused launches)
constexpr size_t kGfx950Lds = 163840;
std::mutex g_lds_mutex;

spfft_status ensure_lds_limit(size_t& budget) {
    std::lock_guard<std::mutex> lk(g_lds_mutex);
    if (g_lds_limit == 0) {
        int dev = 0, v = 0;
        SPFFT_HIP(hipGetDevice(&dev), "hipGetDevice");
        SPFFT_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev),
                  "hipDeviceGetAttribute");
        g_lds_limit = (size_t)v;
        SPFFT_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
        g_cu_count = v;
    }
    budget = std::min<size_t>(g_lds_limit, 65536);
    return SPFFT_OK;
}

}  // namespace

namespace {

// k_conv_fast plan for a length-L FFT: one stage if L itself is an
// instantiated radix, else the most balanced instantiated pair r1 * r2 = L.
bool fast_conv_plan(int L, FftStages& f) {
    f.L = L;
    if (spfft::conv_fast_supported(L, 1)) {
        f.nstages = 1;
        f.radix[0] = L;
        f.ns_before[0] = 1;
        return true;
    }
    for (int r1 = 8; r1 <= 64; r1++) {
        if (L % r1 || L / r1 > r1 || !spfft::conv_fast_supported(r1, L / r1)) continue;
        // smallest r1 >= r2 is the most balanced pair
        f.nstages = 2;
        f.radix[0] = r1;
        f.radix[1] = L / r1;
        f.ns_before[0] = 1;
        f.ns_before[1] = r1;
        return true;
    }
    return false;
}

spfft_status choose_split(spfft_plan_s& p, int m1_req, size_t budget, bool mfma) {
    const int M = p.M;
    RowPassArgs& a = p.base;
    if (M == 1) {  // cols == 1: nothing to convolve beyond the 1-point product
        p.M1 = 1;
        p.M2 = 1;
        a.conv_kind = 0;
        a.f2.L = 1;
        a.f2.nstages = 0;
        return SPFFT_OK;
    }
    // score: fast convolution kernel first, then fewer stages, then M1 near 40
    int best = -1;
    double best_score = 1e18;
    for (int d = 1; d <= 64; d++) {
        if (M % d || !(mfma ? spfft::mfma_supports_m1(d) : spfft::rowpass_supports_m1(d))) continue;
        if (m1_req > 0 && d != m1_req) continue;
        const int m2 = M / d;
        FftStages f;
        double score;
        // LDS of one row: fast kernel (table + row), generic (table + 2 rows)
        if (fast_conv_plan(m2, f) && (size_t)m2 * 16 + 4 <= budget) score = f.nstages * 10.0;
        else if (plan_stages(m2, f) && (size_t)m2 * 24 + 4 <= budget) score = 1000.0 + f.nstages;
        else continue;
        score += std::fabs(std::log((double)d / (mfma ? 32.0 : 40.0)));
        if (score < best_score) {
            best_score = score;
            best = d;
        }
    }
    if (best < 0) {
        if (m1_req > 0)
            return fail(SPFFT_ERR_UNSUPPORTED, "m1 = " + std::to_string(m1_req) +
                                                   " gives no supported split of M = " + std::to_string(M));
        return fail(SPFFT_ERR_UNSUPPORTED, "Bluestein length " + std::to_string(M) +
                                               " has no supported M1 x M2 split");
    }
    p.M1 = best;
    p.M2 = M / best;
    if (fast_conv_plan(p.M2, a.f2)) {
        a.conv_kind = 1;
    } else {
        a.conv_kind = 0;
        plan_stages(p.M2, a.f2);
    }
    return SPFFT_OK;
}

spfft_status choose_tiling(spfft_plan_s& p, size_t budget) {
    RowPassArgs& a = p.base;
    const int req = p.opts.conv_rows;
    if (req < 0 || req > 64) return fail(SPFFT_ERR_INVALID_ARG, "conv_rows must be in [0, 64]");
    int best = 0;
    const int gmax = a.conv_kind == 1 ? 64 : 16;
    for (int G = 1; G <= gmax; G++) {
        a.G2 = G;
        bool ok = spfft::conv_lds_bytes(a) <= budget;
        for (int s = 0; ok && a.conv_kind == 1 && s < a.f2.nstages; s++) {
            const int R = a.f2.radix[s];
            ok = (long)G * (a.M2 / R) <= 256L * spfft::conv_max_butterflies(R);
        }
        if (!ok) break;
        best = G;
        if (req && G == req) break;
    }
    if (best == 0 || (req && best != req))
        return fail(SPFFT_ERR_UNSUPPORTED, "convolution rows do not fit in LDS / registers");
    a.G2 = best;
    return SPFFT_OK;
}

bool check_host_coo(int rows, int cols, int nnz, const int* r, const int* c) {
    for (int i = 0; i < nnz; i++)
        if (r[i] < 0 || r[i] >= rows || c[i] < 0 || c[i] >= cols) return false;
    return true;
}

}  // namespace

namespace spfft_internal {

spfft_status ensure_csc_capacity(spfft_plan_s& p, int nnz) {
    if (nnz <= p.nnz_cap) return SPFFT_OK;
    const int cap = std::max(nnz, 1);
    const size_t scan_n = std::max(spfft::csc_scan_tmp_elems(p.C),
                                   spfft::csc_scan_tmp_elems((int)spfft::csc_sort_hist_elems(cap)));
    const size_t ib = p.packed ? 2 : 4;
    SPFFT_HIP(p.scan_tmp.alloc(scan_n * sizeof(int2)), "hipMalloc(scan)");
    SPFFT_HIP(p.hist.alloc(spfft::csc_sort_hist_elems(cap) * sizeof(int)), "hipMalloc(hist)");
    SPFFT_HIP(p.hist_off.alloc((spfft::csc_sort_hist_elems(cap) + 1) * sizeof(int)), "hipMalloc(hist)");
    SPFFT_HIP(p.row_idx.alloc((size_t)cap * ib), "hipMalloc(row_idx)");
    if (spfft::csc_sort_passes(p.C) > 1) {
        SPFFT_HIP(p.keys0.alloc((size_t)cap * 4), "hipMalloc(keys)");
        SPFFT_HIP(p.keys1.alloc((size_t)cap * 4), "hipMalloc(keys)");
        SPFFT_HIP(p.vals0.alloc((size_t)cap * 4), "hipMalloc(vals)");
        SPFFT_HIP(p.vals1.alloc((size_t)cap * 4), "hipMalloc(vals)");
    }
    p.nnz_cap = cap;
    return SPFFT_OK;
}

spfft_status do_csc_build(spfft_plan_s& p, const int* d_row, const int* d_col, int nnz, hipStream_t s) {
    if (nnz < 0) return fail(SPFFT_ERR_INVALID_ARG, "nnz must be >= 0");
    if (nnz > 0 && (!d_row || !d_col)) return fail(SPFFT_ERR_INVALID_ARG, "NULL COO pointer");
    spfft_status st = ensure_csc_capacity(p, nnz);
    if (st != SPFFT_OK) return st;
    spfft::CscBuildArgs a{};
    a.R = p.R;
    a.C = p.C;
    a.nnz = nnz;
    a.packed = p.packed;
    a.row = d_row;
    a.col = d_col;
    a.counts = p.counts.as<int>();
    a.colstart = p.colstart.as<int>();
    a.sparse_cols = p.sparse_cols.p;
    a.col_ptr = p.col_ptr.as<int>();
    a.row_idx = p.row_idx.p;
    a.totals = p.totals.as<int>();
    a.err = p.err.as<int>();
    a.scan_tmp = p.scan_tmp.as<int2>();
    a.hist = p.hist.as<int>();
    a.hist_off = p.hist_off.as<int>();
    a.keys[0] = p.keys0.as<int>();
    a.keys[1] = p.keys1.as<int>();
    a.vals[0] = p.vals0.as<int>();
    a.vals[1] = p.vals1.as<int>();
    hipEvent_t* ev = nullptr;
    if (p.profiling) {
        const size_t at = p.ev_used;
        ev = p.take_events(2);
        if (!ev) return fail(SPFFT_ERR_HIP, "hipEventCreate failed");
        p.ev_groups.push_back({0, at});
        SPFFT_HIP(hipEventRecord(ev[0], s), "hipEventRecord");
    }
    SPFFT_HIP(spfft::launch_csc_build(a, s), "csc build launch");
    // column-pass rows per lane (the option, or from the number of rows): the
    // extra rows cost one complex product per nonzero instead of a phase
    // reduction and a table lookup; eight rows per lane (8-column tiles of 512
    // rows) once the tile has at least four such row blocks, else four
    p.colpass_rows = p.base.colpass_rows;
    if (p.colpass_rows == 0) p.colpass_rows = auto_colpass_rows(p, nnz);
    SPFFT_HIP(spfft::launch_colsched(a.colstart, p.base.colperm, p.C, p.colpass_rows, p.col_sched.as<int>(),
                                     p.col_sched_off.as<int>(), s),
              "column schedule launch");
    if (ev) SPFFT_HIP(hipEventRecord(ev[1], s), "hipEventRecord");
    p.last_nnz = nnz;
    p.have_csc = true;
    return SPFFT_OK;
}

size_t device_lds_limit() { return g_lds_limit; }

// Column-pass kernel of the automatic plan for nnz nonzeros: rows per lane (a lane's
// further rows follow a three-term recurrence, two FMAs per component, instead of a phase
// reduction and a table lookup): 16 rows per lane (8-column tiles of 1024 rows) when a
// tile has at least 2048 rows and the columns average at least kAutoRows16Degree
// nonzeros, else 8 (tiles of 512 rows).  Measured (one MI355X, whole transforms):
// * uniform matrices of degree 12-29 at n = 6000-16001 and the banded 8219^2 one (degree
//   29.5): 16 rows fastest, 1-9% ahead of 8 and of the matrix-core kernel (which 16 rows
//   replaced here; it stays a candidate for the autotuner);
// * degree 7-9 (the power-law 16706^2 and 40421^2): 16 rows lose 19-22% to 8;
// * tiles below 2048 rows (n = 1000-3500, degree 6-26, and the 2310^2 matrix): 8 rows
//   3-10% ahead of 4 (the earlier choice there) and of 16.
int auto_colpass_rows(const spfft_plan_s& p, int64_t nnz) {
    const int rows = std::min(p.opts.u_tile, p.R / 2 + 1);
    return rows >= 2048 && nnz >= (int64_t)kAutoRows16Degree * p.C ? 16 : 8;
}

// The one-tile row-pass arguments of plan p for the batched form: p.base with the
// column-pass choices that a CSC build and row pass of this plan would resolve (rows per
// lane, row pairing, whole W_R table); the CSC, X0 and output pointers are the caller's.
void batch_descriptor(const spfft_plan_s& p, int64_t nnz, spfft::RowPassArgs& a) {
    a = p.base;
    a.P = p.P;
    int U = p.base.colpass_rows;
    if (U == 0) U = auto_colpass_rows(p, nnz);
    a.colpass_rows = U;
    a.pair_half = pair_eligible(p) && U >= 2 ? 1 : 0;
    if (a.wr_full && spfft::colpass_lds_bytes(a) > g_lds_limit) a.wr_full = nullptr;
    a.u_base = 0;
    a.tile_rows = p.R / 2 + 1;
    a.sub_rows = 0;
}

// Half-period row pairing in the column pass (SPFFT_PAIR_AUTO): R even, one tile
// holds every computed row [0, R/2], no sampled row limit, and not one row per
// lane (a paired lane owns U/2 rows and their U/2 partners).
bool pair_eligible(const spfft_plan_s& p) {
    return p.opts.colpass_pair == SPFFT_PAIR_AUTO && p.R % 2 == 0 && p.opts.row_limit == 0 &&
           p.opts.u_tile >= p.R / 2 + 1 && p.base.colpass_rows != 1;
}

// The plan's row-pass arguments for output pitch `pitch` (the output rows and the
// tile range are set per tile); *rows: the computed rows u < *rows.
static spfft_status rowpass_setup(spfft_plan_s& p, int64_t pitch, float2* T, float2* X0, RowPassArgs& a,
                                  int* rows_out) {
    if (!p.have_csc) return fail(SPFFT_ERR_INVALID_ARG, "no CSC built yet");
    if (pitch < p.V) return fail(SPFFT_ERR_INVALID_ARG, "output pitch smaller than cols/2+1");
    a = p.base;
    a.P = (int)pitch;
    a.colpass_rows = p.colpass_rows;  // resolved with the column schedule at the CSC build
    *rows_out = p.opts.row_limit > 0 ? std::min(p.R / 2 + 1, p.opts.row_limit) : p.R / 2 + 1;
    // rows u and R/2 - u from one phase lookup: R even, one tile of all rows, and
    // at least two output rows per lane (half of them partner rows)
    a.pair_half = pair_eligible(p) && a.colpass_rows >= 2 ? 1 : 0;
    // the whole W_R table only while it fits in LDS next to the row tile
    if (a.wr_full && spfft::colpass_lds_bytes(a) > g_lds_limit) a.wr_full = nullptr;

    a.colstart = p.colstart.as<int>();
    a.col_sched = p.col_sched.as<int>();
    a.col_sched_off = p.col_sched_off.as<int>();
    a.row_idx = p.row_idx.p;
    if (T) a.T = T;
    if (X0) a.X0 = X0;
    if ((!a.T && !a.fused) || !a.X0) return fail(SPFFT_ERR_INVALID_ARG, "plan has no row-pass workspace");
    return SPFFT_OK;
}

// one row-pass tile (rows u0 .. u0 + n - 1 and their mirror rows), with the
// plan's profiling events when enabled
static spfft_status rowpass_tile(spfft_plan_s& p, RowPassArgs& a, int u0, int n, hipStream_t s) {
    a.u_base = u0;
    a.tile_rows = n;
    a.sub_rows = p.base.fused ? 0 : p.t_rows;
    hipEvent_t* ev = nullptr;
    if (p.profiling) {
        const size_t at = p.ev_used;
        ev = p.take_events(5);
        if (!ev) return fail(SPFFT_ERR_HIP, "hipEventCreate failed");
        p.ev_groups.push_back({1, at});
    }
    if (!a.split_n1) {
        SPFFT_HIP(spfft::launch_rowpass_tile(a, p.packed, s, ev), "row pass launch");
        return SPFFT_OK;
    }
    // split C = N1 N2: the column pass into X0 (each row N1 runs of N2 entries), pass A
    // (the width-N2 plan's row kernel over the tile's n N1 runs as rows u N1 + n1, full
    // spectra into T at a pitch of split_p2, no mirror rows: R = 0), pass B
    // (k_splitcols).  Events: column pass, pass A (as the "fft1" stage), nothing, pass B
    // (as "ifft1_finalize").
    RowPassArgs b = p.split_rr ? p.split_rr_args : p.split_sub->base;
    b.X0 = a.X0;
    b.vrow_n1 = a.split_n1;  // tile row u N1 + n1 is run n1 of X0 row u
    b.vrow_magic = (uint32_t)(((1ULL << 32) + a.split_n1 - 1) / a.split_n1);
    b.vrow_ld = a.ldx;
    b.out = b.out_m = a.T;
    b.P = a.split_p2;
    b.V = a.split_n2;
    b.R = 0;
    b.u_base = 0;
    b.tile_rows = n * a.split_n1;
    b.sub_rows = 0;
    if (ev) SPFFT_HIP(hipEventRecord(ev[0], s), "hipEventRecord");
    SPFFT_HIP(spfft::launch_colpass_tile(a, p.packed, s), "column pass launch");
    if (ev) SPFFT_HIP(hipEventRecord(ev[1], s), "hipEventRecord");
    SPFFT_HIP(p.split_rr ? spfft::launch_raderruns(b, s, false) : spfft::launch_fused(b, s, false),
              "split row kernel launch");
    if (ev) SPFFT_HIP(hipEventRecord(ev[2], s), "hipEventRecord");
    if (ev) SPFFT_HIP(hipEventRecord(ev[3], s), "hipEventRecord");
    SPFFT_HIP(spfft::launch_split_cols(a, s), "split column kernel launch");
    if (ev) SPFFT_HIP(hipEventRecord(ev[4], s), "hipEventRecord");
    return SPFFT_OK;
}

spfft_status do_rowpass(spfft_plan_s& p, void* d_out, int64_t pitch, hipStream_t s, float2* T,
                        float2* X0) {
    if (!d_out) return fail(SPFFT_ERR_INVALID_ARG, "NULL output pointer");
    RowPassArgs a;
    int rows = 0;
    spfft_status st = rowpass_setup(p, pitch, T, X0, a, &rows);
    if (st != SPFFT_OK) return st;
    a.out = a.out_m = static_cast<float2*>(d_out);
    // Rows u in [0, R/2] are computed; each also writes its mirror row R - u
    // (X[R-u][v] = conj(X[u][(C - v) mod C]) for a real input).
    for (int u0 = 0; u0 < rows; u0 += p.opts.u_tile)
        if ((st = rowpass_tile(p, a, u0, std::min(p.opts.u_tile, rows - u0), s)) != SPFFT_OK) return st;
    return SPFFT_OK;
}

// p + e elements as a device address (e may leave the allocation: a kernel only
// adds offsets that land inside it)
static float2* elem_shift(float2* p, int64_t e) {
    return reinterpret_cast<float2*>(reinterpret_cast<uintptr_t>(p) + (uintptr_t)(e * (int64_t)sizeof(float2)));
}

// Streamed output, after the reference's double-buffered streaming drivers
// (src/sparse_fft_bluestein.cu:912-1145): tile k of computed rows u0 .. u0+n-1
// writes chunk k % 2 on the caller's stream -- rows u at chunk row u - u0, their
// mirror rows R - u (u in [1, (R-1)/2]) at chunk row n + (u0 + n - 1 - u), so both
// groups are contiguous -- and the copy stream moves them to h_out while the next
// tile computes.  A chunk is rewritten only after its copies are done.
spfft_status do_rowpass_to_host(spfft_plan_s& p, void* h_out, int64_t pitch, hipStream_t s) {
    if (!h_out) return fail(SPFFT_ERR_INVALID_ARG, "NULL host output pointer");
    if (pitch < p.V) return fail(SPFFT_ERR_INVALID_ARG, "output pitch smaller than cols/2+1");
    RowPassArgs a;
    int rows = 0;
    spfft_status st = rowpass_setup(p, p.P, nullptr, nullptr, a, &rows);
    if (st != SPFFT_OK) return st;
    const int tile = p.opts.u_tile;
    const size_t chunk_bytes = (size_t)2 * std::min(tile, rows) * p.P * sizeof(float2);
    for (DevBuf& c : p.chunk)
        if (c.bytes < chunk_bytes) SPFFT_HIP(c.alloc(chunk_bytes), "hipMalloc(output chunk)");
    if (!p.copy_stream) SPFFT_HIP(hipStreamCreateWithFlags(&p.copy_stream, hipStreamNonBlocking), "hipStreamCreate");
    for (hipEvent_t* e : {&p.chunk_ready[0], &p.chunk_ready[1], &p.chunk_free[0], &p.chunk_free[1], &p.copies_done})
        if (!*e) SPFFT_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming), "hipEventCreate");
    float2* host = static_cast<float2*>(h_out);
    const size_t hp = (size_t)pitch * sizeof(float2), dp = (size_t)p.P * sizeof(float2);
    const size_t width = (size_t)p.V * sizeof(float2);
    int k = 0;
    for (int u0 = 0; u0 < rows; u0 += tile, k++) {
        const int n = std::min(tile, rows - u0), b = k & 1;
        float2* ch = p.chunk[b].as<float2>();
        if (k >= 2) SPFFT_HIP(hipStreamWaitEvent(s, p.chunk_free[b], 0), "hipStreamWaitEvent");
        a.out = elem_shift(ch, -(int64_t)u0 * p.P);
        a.out_m = elem_shift(ch, (int64_t)(2 * n + u0 - 1 - p.R) * p.P);
        if ((st = rowpass_tile(p, a, u0, n, s)) != SPFFT_OK) return st;
        SPFFT_HIP(hipEventRecord(p.chunk_ready[b], s), "hipEventRecord");
        SPFFT_HIP(hipStreamWaitEvent(p.copy_stream, p.chunk_ready[b], 0), "hipStreamWaitEvent");
        SPFFT_HIP(hipMemcpy2DAsync(host + (size_t)u0 * pitch, hp, ch, dp, width, n, hipMemcpyDeviceToHost,
                                   p.copy_stream),
                  "hipMemcpy2DAsync(rows)");
        const int lo = std::max(u0, 1), hi = std::min(u0 + n - 1, (p.R - 1) / 2);
        if (lo <= hi)
            SPFFT_HIP(hipMemcpy2DAsync(host + (size_t)(p.R - hi) * pitch, hp, ch + (size_t)(2 * n + u0 - 1 - hi) * p.P,
                                       dp, width, hi - lo + 1, hipMemcpyDeviceToHost, p.copy_stream),
                      "hipMemcpy2DAsync(mirror rows)");
        SPFFT_HIP(hipEventRecord(p.chunk_free[b], p.copy_stream), "hipEventRecord");
    }
    // work queued on the caller's stream after this call sees the copies done
    SPFFT_HIP(hipEventRecord(p.copies_done, p.copy_stream), "hipEventRecord");
    SPFFT_HIP(hipStreamWaitEvent(s, p.copies_done, 0), "hipStreamWaitEvent");
    return SPFFT_OK;
}

}  // namespace spfft_internal

extern "C" {

const char* spfft_last_error(void) { return g_last_error.c_str(); }
const char* spfft_version(void) { return "spfft-mi355x 0.1 (gfx950)"; }

void spfft_opts_default(spfft_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->struct_size = (int)sizeof(spfft_opts);
    o->length_rule = SPFFT_SMOOTH_3;
    o->u_tile = 0;  // automatic
    o->precision = SPFFT_PREC_FP32;
}

void spfft_reference_opts(int rule, int u_tile, spfft_opts* o) {
    if (!o) return;
    spfft_opts_default(o);
    o->length_rule = rule;
    o->u_tile = u_tile;
    o->row_transform = SPFFT_TRANSFORM_BLUESTEIN;
}

int spfft_next_3_smooth(int n) {
    static const int p[] = {2, 3};
    return next_smooth(n, p, 2);
}

int spfft_next_7_smooth(int n) {
    static const int p[] = {2, 3, 5, 7};
    return next_smooth(n, p, 4);
}

int spfft_fused_lengths(int* out, int cap) {
    std::vector<int> ms;
    spfft::FusedShape f;
    for (int i = 0; spfft::fused_shape(i, &f); i++) {
        const int m = f.m1 * f.r[0] * f.r[1] * f.r[2];
        if (is_7_smooth(m)) ms.push_back(m);  // Bluestein lengths (direct-only shapes: spfft_direct_lengths)
    }
    std::sort(ms.begin(), ms.end());
    ms.erase(std::unique(ms.begin(), ms.end()), ms.end());
    for (int i = 0; out && i < (int)ms.size() && i < cap; i++) out[i] = ms[i];
    return (int)ms.size();
}

int spfft_direct_lengths(int* out, int cap) {
    std::vector<int> ms;
    spfft::FusedShape f;
    for (int i = 0; spfft::fused_shape(i, &f); i++)
        if (f.nb == 1) ms.push_back(f.m1 * f.r[0] * f.r[1] * f.r[2]);
    std::sort(ms.begin(), ms.end());
    ms.erase(std::unique(ms.begin(), ms.end()), ms.end());
    for (int i = 0; out && i < (int)ms.size() && i < cap; i++) out[i] = ms[i];
    return (int)ms.size();
}

int spfft_rader_widths(int* out, int cap) {
    std::vector<int> ws;
    spfft::RaderShape r;
    for (int i = 0; spfft::rader_shape(i, &r); i++) {
        const int p = r.m1 * r.r[0] * r.r[1] * r.r[2] + 1;
        if (is_odd_prime(p)) ws.push_back(r.ns * p);
    }
    std::sort(ws.begin(), ws.end());
    ws.erase(std::unique(ws.begin(), ws.end()), ws.end());
    for (int i = 0; out && i < (int)ws.size() && i < cap; i++) out[i] = ws[i];
    return (int)ws.size();
}

int spfft_factor_3_smooth(int n, int* out, int cap) {
    int c3 = 0, k = 0;
    if (n < 1) return -1;
    while (n % 3 == 0) { c3++; n /= 3; }
    for (int i = 0; i < c3 / 2; i++) { if (out && k < cap) out[k] = 9; k++; }
    if (c3 % 2) { if (out && k < cap) out[k] = 3; k++; }
    while (n % 2 == 0) { if (out && k < cap) out[k] = 2; k++; n /= 2; }
    return n == 1 ? k : -1;
}

}  // extern "C"

namespace {

// Shape planning only (no device allocation): M, the M1 x M2 split, the
// convolution kernel and its rows per workgroup.
// Fused row-kernel shape for M, or -1.  lds_limit: the device's LDS per workgroup.
int find_fused(int M, int m1_req, size_t lds_limit, spfft::FusedShape* out, bool one_batch = false,
               int nt_req = 0) {
    spfft::FusedShape f;
    for (int i = 0; spfft::fused_shape(i, &f); i++) {
        if (f.m1 * f.r[0] * f.r[1] * f.r[2] != M || (m1_req && f.m1 != m1_req) || f.lds_bytes > lds_limit ||
            (one_batch && f.nb != 1) || (nt_req && f.nt != nt_req))
            continue;
        *out = f;
        return i;
    }
    return -1;
}

// Rader shape for the convolution length L = p - 1 with ns convolutions per row (2: width
// 2p, 1: width p), as a fused shape (one batch, rpw = the ns row groups of k_rowrader),
// or -1.
int find_rader(int L, int ns, int m1_req, size_t lds_limit, bool mfma, spfft::FusedShape* out, int nt_req = 0) {
    spfft::RaderShape r;
    for (int i = 0; spfft::rader_shape(i, &r); i++) {
        if (r.m1 * r.r[0] * r.r[1] * r.r[2] != L || r.ns != ns || (m1_req && r.m1 != m1_req) ||
            r.lds_bytes > lds_limit || r.mfma != (mfma ? 1 : 0) || (nt_req && r.nt != nt_req))
            continue;
        *out = spfft::FusedShape{r.m1, 1, r.ns, r.nt, {r.r[0], r.r[1], r.r[2]}, r.lds_bytes};
        return i;
    }
    return -1;
}

// The width-n2 row pass of a split plan, on chip: the direct transform when n2 has a
// one-batch fused shape, else Bluestein at the smallest fused length M >= 2 n2 - 1 within
// 2.5 n2 (Rader is not used there: its row kernel writes half spectra).  *len: its length.
bool split_sub_opts(int n2, size_t lds_limit, spfft_opts* so, int* len) {
    spfft_opts_default(so);
    spfft::FusedShape f;
    if (find_fused(n2, 0, lds_limit, &f, true) >= 0) {
        so->row_transform = SPFFT_TRANSFORM_DIRECT;
        *len = n2;
        return true;
    }
    int best = 0;
    for (int i = 0; spfft::fused_shape(i, &f); i++) {
        const int m = f.m1 * f.r[0] * f.r[1] * f.r[2];
        if (f.lds_bytes <= lds_limit && is_7_smooth(m) && m >= 2 * n2 - 1 && (best == 0 || m < best)) best = m;
    }
    if (best == 0 || best > 2.5 * n2 + 64) return false;
    so->row_transform = SPFFT_TRANSFORM_BLUESTEIN;
    so->fft_len = best;
    *len = best;
    return true;
}

// Does an odd prime run length n2 have a k_raderruns shape (Rader over n2 - 1)?
bool split_rr_ok(int n2, size_t lds_limit) {
    int m1, ng, nt, r[3];
    return is_odd_prime(n2) && spfft::raderruns_shape(n2 - 1, &m1, &ng, &nt, r) &&
           spfft::raderruns_lds_bytes(n2 - 1) <= lds_limit;
}

// The split C = n1 * n2 with the smallest n1 (only n1 = n1_req when that is > 0) that
// has a column kernel and leaves an on-chip row pass of width n2, or 0.  *rr: that pass
// is k_raderruns (an odd prime n2 with a shape: a length-(n2 - 1) convolution instead of
// Bluestein's >= 2 n2 - 1), else the width-n2 plan of *so.
int find_split(int cols, int n1_req, int tile_rows, size_t lds_limit, spfft_opts* so, int* n2_out, int* len,
               bool* rr) {
    for (int n1 = 2; n1 <= 127 && n1 <= cols / 2; n1++) {
        if (n1_req > 0 && n1 != n1_req) continue;
        bool direct = false;
        if (cols % n1 || !spfft::split_supports_n1(n1, &direct)) continue;
        if ((long long)tile_rows * n1 > (1 << 22)) continue;  // rows of the width-n2 plan
        const int n2 = cols / n1;
        *rr = split_rr_ok(n2, lds_limit);
        if (*rr) {
            *n2_out = n2;
            *len = n2 - 1;
            return n1;
        }
        if (split_sub_opts(n2, lds_limit, so, len)) {
            *n2_out = n2;
            return n1;
        }
    }
    return 0;
}

spfft_status plan_shape(int rows, int cols, const spfft_opts* opts_in, size_t budget, size_t lds_limit,
                        spfft_plan_s* p) {
    spfft_opts o;
    spfft_status st0 = read_opts(opts_in, &o);
    if (st0 != SPFFT_OK) return st0;
    if (rows <= 0 || cols <= 0) return fail(SPFFT_ERR_INVALID_ARG, "rows and cols must be positive");
    if (o.u_tile < 0) return fail(SPFFT_ERR_INVALID_ARG, "CSC tile size must be positive");
    if (o.precision != SPFFT_PREC_FP32 && o.precision != SPFFT_PREC_BF16X3)
        return fail(SPFFT_ERR_UNSUPPORTED, "unknown precision form");
    const bool mfma = o.precision == SPFFT_PREC_BF16X3;
    if (mfma && o.m1 && !spfft::mfma_supports_m1(o.m1))
        return fail(SPFFT_ERR_UNSUPPORTED, "the BF16x3 form needs m1 = 16 or 32");
    if (rows > (1 << 22)) return fail(SPFFT_ERR_UNSUPPORTED, "rows > 2^22 not supported");
    if (cols > (1 << 20)) return fail(SPFFT_ERR_UNSUPPORTED, "cols > 2^20 not supported");
    int M;
    if (o.fft_len > 0) {
        if (o.fft_len < 2 * cols - 1 || !is_7_smooth(o.fft_len))
            return fail(SPFFT_ERR_INVALID_ARG, "fft_len must be 7-smooth and >= 2*cols-1");
        M = o.fft_len;
    } else if (o.length_rule == SPFFT_SMOOTH_3 || o.length_rule == SPFFT_SMOOTH_7) {
        auto next = [&](int n) {
            return o.length_rule == SPFFT_SMOOTH_3 ? spfft_next_3_smooth(n) : spfft_next_7_smooth(n);
        };
        M = next(2 * cols - 1);
        if (mfma) {
            // smallest smooth M that splits as 32 * m2 or 16 * m2 with a fast m2 kernel
            // (a generic convolution kernel if no such M lies within 25% of 2*cols-1)
            auto ok = [&](int m, bool fast) {
                FftStages f;
                for (int m1 : {32, 16}) {
                    if (m % m1 || (o.m1 && o.m1 != m1)) continue;
                    const int m2 = m / m1;
                    if (fast ? (fast_conv_plan(m2, f) && (size_t)m2 * 16 + 4 <= budget)
                             : (plan_stages(m2, f) && (size_t)m2 * 24 + 4 <= budget))
                        return true;
                }
                return false;
            };
            const int m0 = M, limit = (int)std::min<long>(1L << 24, (long)(1.25 * (2.0 * cols - 1)) + 64);
            while (M <= limit && !ok(M, true)) M = next(M + 1);
            if (M > limit) {
                M = m0;
                while (M < (1 << 24) && !ok(M, false)) M = next(M + 1);
            }
        }
    } else {
        return fail(SPFFT_ERR_INVALID_ARG, "length_rule must be 3 or 7");
    }
    if (o.row_persist != 0 && o.row_persist != 1) return fail(SPFFT_ERR_INVALID_ARG, "row_persist must be 0 or 1");
    // the direct length-cols FFT: a single-batch fused shape of length cols
    if (o.row_transform < SPFFT_TRANSFORM_AUTO || o.row_transform > SPFFT_TRANSFORM_SPLIT)
        return fail(SPFFT_ERR_INVALID_ARG, "row_transform must be 0, 1, 2, 3 or 4");
    spfft::FusedShape fd{};
    const bool chip_ok = !mfma && o.fft_len == 0 && o.conv_rows == 0 && o.row_kernel != SPFFT_ROW_FOUR_STEP;
    if (o.row_threads < 0 || o.row_threads % 64) return fail(SPFFT_ERR_INVALID_ARG, "row_threads must be 0 or a multiple of 64");
    const bool direct_ok = chip_ok && find_fused(cols, o.m1, lds_limit, &fd, true, o.row_threads) >= 0;
    if (o.row_transform == SPFFT_TRANSFORM_DIRECT && !direct_ok)
        return fail(SPFFT_ERR_UNSUPPORTED, "no direct transform for cols = " + std::to_string(cols) +
                                               " with these options");
    const bool direct =
        direct_ok && (o.row_transform == SPFFT_TRANSFORM_AUTO || o.row_transform == SPFFT_TRANSFORM_DIRECT);
    // Rader: cols = 2p or cols = p, p an odd prime, with a shape for the convolution
    // length p - 1 (the precision form has Rader shapes of its own: k_rowrader_mfma)
    const bool rader_chip_ok = o.fft_len == 0 && o.conv_rows == 0 && o.row_kernel != SPFFT_ROW_FOUR_STEP;
    const int rader_ns = cols % 2 == 0 ? 2 : 1;
    const int rader_pr = cols / rader_ns;
    const bool rader_ok = rader_chip_ok && is_odd_prime(rader_pr) &&
                          find_rader(rader_pr - 1, rader_ns, o.m1, lds_limit, mfma, &fd, o.row_threads) >= 0;
    if (o.row_transform == SPFFT_TRANSFORM_RADER && !rader_ok)
        return fail(SPFFT_ERR_UNSUPPORTED, "no Rader transform for cols = " + std::to_string(cols) +
                                               " with these options");
    const bool rader =
        !direct && rader_ok && (o.row_transform == SPFFT_TRANSFORM_AUTO || o.row_transform == SPFFT_TRANSFORM_RADER);
    if (direct) {
        find_fused(cols, o.m1, lds_limit, &fd, true, o.row_threads);
        M = cols;
    }
    if (rader) M = rader_pr - 1;
    p->R = rows;
    p->C = cols;
    p->V = cols / 2 + 1;
    p->P = (p->V + 3) / 4 * 4;
    p->M = M;
    p->opts = o;
    p->packed = rows <= 65536 && cols <= 65536;
    if (o.row_kernel < SPFFT_ROW_AUTO || o.row_kernel > SPFFT_ROW_FUSED)
        return fail(SPFFT_ERR_INVALID_ARG, "row_kernel must be 0, 1 or 2");
    spfft::FusedShape fs = fd;
    const bool fused_ok = direct || rader ||
                          (!mfma && o.conv_rows == 0 && o.row_kernel != SPFFT_ROW_FOUR_STEP &&
                           o.row_transform != SPFFT_TRANSFORM_SPLIT &&
                           find_fused(M, o.m1, lds_limit, &fs, false, o.row_threads) >= 0);
    // split (C = N1 N2): when asked for, or (AUTO) when no on-chip row pass exists for C
    int split_n1 = 0, split_n2 = 0;
    {
        // (m1: the split's N1 when SPLIT is asked for)
        const bool split_opts_ok = !mfma && o.fft_len == 0 && o.conv_rows == 0 &&
                                   (o.m1 == 0 || o.row_transform == SPFFT_TRANSFORM_SPLIT) && o.t_rows == 0 &&
                                   o.row_threads == 0 && o.row_kernel != SPFFT_ROW_FOUR_STEP;
        const bool want = o.row_transform == SPFFT_TRANSFORM_SPLIT ||
                          (o.row_transform == SPFFT_TRANSFORM_AUTO && !fused_ok);
        const int tile_rows = std::min(o.u_tile > 0 ? o.u_tile : rows / 2 + 1, rows / 2 + 1);
        if (split_opts_ok && want)
            split_n1 = find_split(cols, o.row_transform == SPFFT_TRANSFORM_SPLIT ? o.m1 : 0, tile_rows, lds_limit,
                                  &p->split_opts, &split_n2, &p->split_len,
                                  &p->split_rr);
        if (o.row_transform == SPFFT_TRANSFORM_SPLIT && !split_n1)
            return fail(SPFFT_ERR_UNSUPPORTED, "no split transform for cols = " + std::to_string(cols) +
                                                   " with these options");
    }
    if (o.row_kernel == SPFFT_ROW_FUSED && !fused_ok && !split_n1)
        return fail(SPFFT_ERR_UNSUPPORTED, "no fused row-kernel shape for M = " + std::to_string(M) +
                                               (mfma ? " in the BF16x3 form" : ""));
    RowPassArgs& a = p->base;
    spfft_status st = SPFFT_OK;
    p->opts.u_tile = o.u_tile;
    if (split_n1) {
        // the row pass is the width-N2 plan's kernel plus k_splitcols: no M of its own
        // (the tables sized by M, M1, M2 are one entry)
        M = 1;
        p->M = p->M1 = p->M2 = 1;
        a.fused = 0;
        a.conv_kind = 0;
        a.f2.L = 1;
        a.f2.nstages = 0;
        a.split_n1 = split_n1;
        a.split_n2 = split_n2;
    } else if (fused_ok) {
        p->M1 = fs.m1;
        p->M2 = fs.r[0] * fs.r[1] * fs.r[2];
        a.fused = 1;
        a.direct = direct ? 1 : 0;
        a.rader = rader ? 1 : 0;
        a.rader_p = rader ? rader_pr : 0;
        a.fused_nb = fs.nb;
        a.fused_rpw = fs.rpw;
        a.fused_nt = fs.nt;
        a.conv_kind = 1;
        a.f2.L = p->M2;
        a.f2.nstages = 0;
        for (int i = 0, ns = 1; i < 3; i++) {
            if (fs.r[i] == 1) continue;
            a.f2.radix[a.f2.nstages] = fs.r[i];
            a.f2.ns_before[a.f2.nstages++] = ns;
            ns *= fs.r[i];
        }
    } else {
        a.fused = 0;
        st = choose_split(*p, o.m1, budget, mfma);
    }
    if (st != SPFFT_OK && !split_n1 && o.fft_len == 0 && o.m1 == 0) {
        // a rule-derived M without a supported split (e.g. 5^7): the next smooth length
        auto next = [&](int n) {
            return o.length_rule == SPFFT_SMOOTH_3 ? spfft_next_3_smooth(n) : spfft_next_7_smooth(n);
        };
        for (int tries = 0; st != SPFFT_OK && tries < 64 && M < (1 << 24); tries++) {
            M = next(M + 1);
            p->M = M;
            st = choose_split(*p, 0, budget, mfma);
        }
    }
    if (st != SPFFT_OK) return st;
    // u_tile 0: one tile of every computed row
    if (p->opts.u_tile == 0) p->opts.u_tile = rows / 2 + 1;
    // one tile is at most 65535 workgroup rows of the length-M1 kernels' grid
    p->opts.u_tile = std::min(p->opts.u_tile, 65535);
    // four-step kernels: T holds t_rows rows at a time (0: the whole tile)
    if (o.t_rows < 0) return fail(SPFFT_ERR_INVALID_ARG, "t_rows must be >= 0");
    p->t_rows = 0;
    if (!a.fused && !split_n1) {
        const int tile = std::min(p->opts.u_tile, rows / 2 + 1);
        p->t_rows = o.t_rows > 0 ? std::min(o.t_rows, tile) : tile;
    }
    a.R = rows;
    a.C = cols;
    a.V = p->V;
    a.P = p->P;
    a.M = M;
    a.M1 = p->M1;
    a.M2 = p->M2;
    a.ldx = (cols + 63) / 64 * 64;
    a.split_p2 = split_n1 ? (split_n2 + 15) / 16 * 16 : 0;
    int hs = 8;
    while ((1LL << (2 * hs)) < rows) hs++;
    a.hi_shift = hs;
    a.n_hi = (rows + (1 << hs) - 1) >> hs;
    a.r_magic = ~0ULL / (uint64_t)rows + 1;
    // just below 1/R: floor(x * inv_r) is floor(x / R) or one less for x < 2^28
    a.inv_r = (float)((1.0 / (double)rows) * (1.0 - 0x1p-20));
    a.precision = mfma ? 1 : 0;
    if (o.colpass_rows < 0 || (o.colpass_rows > 16 && o.colpass_rows != spfft::kColpassMfma) ||
        (o.colpass_rows & (o.colpass_rows - 1)))
        return fail(SPFFT_ERR_INVALID_ARG, "colpass_rows must be 0, 1, 2, 4, 8, 16 or 32 (matrix cores)");
    a.colpass_rows = o.colpass_rows ? o.colpass_rows : 0;  // 0: set from the average degree at CSC build
    if (o.colpass_table < 0 || o.colpass_table > 2)
        return fail(SPFFT_ERR_INVALID_ARG, "colpass_table must be 0, 1 or 2");
    if (o.colpass_pair != SPFFT_PAIR_AUTO && o.colpass_pair != SPFFT_PAIR_OFF)
        return fail(SPFFT_ERR_INVALID_ARG, "colpass_pair must be 0 or 1");
    const bool full_fits = (size_t)rows * 8 <= budget;
    if (o.colpass_table == 2 && !full_fits)
        return fail(SPFFT_ERR_UNSUPPORTED, "W_R table does not fit in LDS");
    // automatic: the whole table only for short columns.  Next to the column pass's row
    // tile (37-74 KB) a whole table of more rows leaves one workgroup per CU, and the
    // two-level table's extra product costs less than the lost occupancy (R = 6000: column
    // pass 0.104 ms whole, 0.041 ms two-level; R = 8000: 0.184 / 0.067 ms)
    p->colpass_full = o.colpass_table == 2 || (o.colpass_table == 0 && full_fits && rows <= kAutoFullTableRows);
    if (a.fused || split_n1) {
        a.G2 = 0;
        return SPFFT_OK;
    }
    return choose_tiling(*p, budget);
}

void fill_info(const spfft_plan_s* p, spfft_plan_info* info) {
    std::memset(info, 0, sizeof(*info));
    info->struct_size = (int)sizeof(spfft_plan_info);
    info->rows = p->R;
    info->cols = p->C;
    info->out_cols = p->V;
    info->pitch = p->P;
    info->fft_len = p->M;
    info->m1 = p->M1;
    info->m2 = p->M2;
    info->packed = p->packed ? 1 : 0;
    info->u_tile = p->opts.u_tile;
    info->conv_rows = p->base.G2;
    info->conv_fast = p->base.conv_kind;
    info->conv_stages = p->base.f2.nstages;
    for (int i = 0; i < 8; i++) info->conv_radix[i] = i < p->base.f2.nstages ? p->base.f2.radix[i] : 0;
    info->precision = p->base.precision;
    info->colpass_rows = p->base.colpass_rows;
    info->colpass_full_table = p->colpass_full ? 1 : 0;
    info->workspace_bytes = p->table_bytes;
    info->row_kernel = p->base.fused ? SPFFT_ROW_FUSED : SPFFT_ROW_FOUR_STEP;
    info->colpass_paired = pair_eligible(*p) ? 1 : 0;
    info->direct = p->base.direct;
    info->rader = p->base.rader;
    info->row_transform = p->base.direct ? SPFFT_TRANSFORM_DIRECT
                          : p->base.rader ? SPFFT_TRANSFORM_RADER
                          : p->base.split_n1 ? SPFFT_TRANSFORM_SPLIT
                                          : SPFFT_TRANSFORM_BLUESTEIN;
    info->t_rows = p->t_rows;
    info->row_threads = p->base.fused ? p->base.fused_nt : 0;
    if (p->base.split_n1) {
        info->fft_len = p->split_len;
        info->m1 = p->base.split_n1;
        info->m2 = p->base.split_n2;
        info->row_kernel = SPFFT_ROW_FUSED;
        info->row_threads = p->split_sub ? p->split_sub->base.fused_nt : 0;
        int m1, ng, nt, r[3];
        if (p->split_rr && spfft::raderruns_shape(p->base.split_n2 - 1, &m1, &ng, &nt, r)) info->row_threads = nt;
    }
}

}  // namespace

extern "C" {

spfft_status spfft_plan_describe(int rows, int cols, const spfft_opts* opts, spfft_plan_info* info) {
    spfft_status st0 = check_info(info);
    if (st0 != SPFFT_OK) return st0;
    spfft_plan_s p;
    // the planner's view of a gfx950 device: 160 KiB of LDS per workgroup
    spfft_status st = plan_shape(rows, cols, opts, 65536, kGfx950Lds, &p);
    if (st != SPFFT_OK) return st;
    fill_info(&p, info);
    return SPFFT_OK;
}

spfft_status spfft_plan_create(int rows, int cols, const spfft_opts* opts_in, spfft_plan* out) {
    if (!out) return fail(SPFFT_ERR_INVALID_ARG, "NULL plan pointer");
    *out = nullptr;
    return plan_create_ex(rows, cols, opts_in, true, out);
}

}  // extern "C"

namespace spfft_internal {

void workspace_bytes(const spfft_plan_s& p, size_t* t_bytes, size_t* x0_bytes) {
    const size_t tile_rows = (size_t)std::min(p.opts.u_tile, p.R / 2 + 1);
    *t_bytes = p.base.split_n1 ? tile_rows * (size_t)p.base.split_n1 * p.base.split_p2 * sizeof(float2)
               : p.base.fused ? 0
                              : (size_t)p.t_rows * (size_t)p.M * sizeof(float2);
    *x0_bytes = tile_rows * (size_t)p.base.ldx * sizeof(float2);
}

spfft_status plan_create_ex(int rows, int cols, const spfft_opts* opts_in, bool own_workspace,
                            spfft_plan_s** out) {
    *out = nullptr;
    spfft_plan_s* p = new (std::nothrow) spfft_plan_s();
    if (!p) return fail(SPFFT_ERR_ALLOC, "host allocation failed");
    std::unique_ptr<spfft_plan_s> guard(p);
    size_t budget = 0;
    spfft_status st = ensure_lds_limit(budget);
    if (st != SPFFT_OK) return st;
    {
        static std::once_flag once;
        static hipError_t prep = hipSuccess;
        std::call_once(once, [] { prep = spfft::colpass_prepare(g_lds_limit); });
        SPFFT_HIP(prep, "column pass LDS limit");
    }
    st = plan_shape(rows, cols, opts_in, budget, g_lds_limit, p);
    if (st != SPFFT_OK) return st;
    RowPassArgs& a = p->base;
    // fused row kernel: persistent launch over the device's CUs unless turned off
    a.fused_cus = a.fused && p->opts.row_persist == 0 ? g_cu_count : 0;
    const int M = p->M;
    const int hs = a.hi_shift;
    const spfft_opts& o = p->opts;

    // ---- tables (FP64 -> FP32) ----
    std::vector<float2> h(cols), lo(1 << hs), hi(a.n_hi), t2(p->M2), tm((size_t)M), bh((size_t)M);
    for (int k = 0; k < cols; k++) {
        const long long k2 = (long long)k * k % (2LL * cols);
        // exp(-i pi k^2 / C); the direct and Rader transforms have no chirp (h = 1)
        h[k] = a.direct || a.rader ? make_float2(1.f, 0.f) : to_f2(unit_root(k2, 2LL * cols, -1));
    }
    if (a.split_n1) {
        // split C = N1 N2: the width-N2 plan (its row kernel takes the runs of N2 entries
        // of each X0 row as its rows: X0 position n1 N2 + n2 holds column n1 + N1 n2); the column pass
        // applies that plan's input chirp, h[n1 N2 + n2] = exp(-i pi n2^2 / N2) (1 for
        // its direct transform); W_C and, for an odd prime N1 > 31, the DFT_N1 table of
        // k_splitcols
        const int n1s = a.split_n1, n2s = a.split_n2;
        std::vector<int> perm((size_t)cols);
        if (p->split_rr) {
            // Rader over each run (k_raderruns): run position j = n2 + M2 n1 < L holds
            // x[g^q] of the run (column i1 + N1 g^q), position L holds x[0] (rader_tables)
            int m1, ng, nt, r[3];
            spfft::raderruns_shape(n2s - 1, &m1, &ng, &nt, r);
            const int L = n2s - 1, m2 = r[0] * r[1] * r[2];
            std::vector<int> gpos, kt;
            std::vector<float2> bhr, t2r((size_t)m2);
            rader_tables(n2s, m1, m2, gpos, kt, bhr);
            for (int i1 = 0; i1 < n1s; i1++) {
                for (int j = 0; j < L; j++) perm[(size_t)i1 * n2s + j] = i1 + n1s * gpos[j];
                perm[(size_t)i1 * n2s + L] = i1;
            }
            for (int k = 0; k < cols; k++) h[k] = make_float2(1.f, 0.f);
            for (int i = 0; i < m2; i++) t2r[i] = to_f2(unit_root(i, m2, -1));
            SPFFT_HIP(upload(p->rader_k, kt), "upload Rader indices");
            SPFFT_HIP(upload(p->rr_bhat, bhr), "upload B^");
            SPFFT_HIP(upload(p->rr_tw2, t2r), "upload twiddles");
            spfft::RowPassArgs& b = p->split_rr_args;
            b = spfft::RowPassArgs{};
            b.C = n2s;
            b.M = L;
            b.M1 = m1;
            b.M2 = m2;
            b.rader = 1;
            b.rader_p = n2s;
            b.fused = 1;
            b.fused_rpw = ng;
            b.fused_nt = nt;
            b.f2.L = m2;
            b.f2.nstages = 0;
            for (int i = 0, ns = 1; i < 3; i++) {
                if (r[i] == 1) continue;
                b.f2.radix[b.f2.nstages] = r[i];
                b.f2.ns_before[b.f2.nstages++] = ns;
                ns *= r[i];
            }
            b.rader_k = p->rader_k.as<int>();
            b.bhat = p->rr_bhat.as<float2>();
            b.tw2 = p->rr_tw2.as<float2>();
            SPFFT_HIP(spfft::launch_raderruns(b, nullptr, true), "Rader run kernel LDS limit");
            p->split_len = L;
        } else {
            const int tile_rows = std::min(p->opts.u_tile, rows / 2 + 1);
            spfft_plan_s* sub = nullptr;
            st = plan_create_ex(tile_rows * n1s, n2s, &p->split_opts, false, &sub);
            if (st != SPFFT_OK) return st;
            p->split_sub.reset(sub);
            if (!sub->base.fused || sub->base.rader)
                return fail(SPFFT_ERR_UNSUPPORTED, "split transform: the width-" + std::to_string(n2s) +
                                                       " row pass is not on chip");
            p->split_len = sub->M;
            for (int i1 = 0; i1 < n1s; i1++)
                for (int i2 = 0; i2 < n2s; i2++) {
                    const size_t pos = (size_t)i1 * n2s + i2;
                    perm[pos] = i1 + n1s * i2;
                    const long long q = (long long)i2 * i2 % (2LL * n2s);
                    h[pos] = sub->base.direct ? make_float2(1.f, 0.f) : to_f2(unit_root(q, 2LL * n2s, -1));
                }
        }
        std::vector<float2> w((size_t)256 + (cols + 255) / 256);
        for (int e = 0; e < 256; e++) w[e] = to_f2(unit_root(e, cols, -1));
        for (int i = 0; i < (cols + 255) / 256; i++) w[256 + i] = to_f2(unit_root(256LL * i, cols, -1));
        SPFFT_HIP(upload(p->colperm, perm), "upload split column order");
        SPFFT_HIP(upload(p->split_w, w), "upload twiddles");
        a.colperm = p->colperm.as<int>();
        a.split_w = p->split_w.as<float2>();
        bool dir = false;
        spfft::split_supports_n1(n1s, &dir);
        if (dir) {
            const int H = (n1s - 1) / 2;
            std::vector<float> t((size_t)2 * H * H);
            for (int pp = 1; pp <= H; pp++)
                for (int q = 1; q <= H; q++) {
                    const cd z = unit_root((long long)pp * q, n1s, +1);  // cos + i sin
                    t[2 * ((size_t)(pp - 1) * H + q - 1)] = (float)z.real();
                    t[2 * ((size_t)(pp - 1) * H + q - 1) + 1] = (float)z.imag();
                }
            SPFFT_HIP(upload(p->split_dft, t), "upload DFT table");
            a.split_dft = p->split_dft.as<float2>();
        }
    }
    for (int l = 0; l < (1 << hs); l++) lo[l] = to_f2(unit_root(l, rows, -1));
    for (int j = 0; j < a.n_hi; j++) hi[j] = to_f2(unit_root((long long)j << hs, rows, -1));
    for (int i = 0; i < p->M2; i++) t2[i] = to_f2(unit_root(i, p->M2, -1));
    for (int k1 = 0; k1 < p->M1; k1++)
        for (int n2 = 0; n2 < p->M2; n2++)
            tm[(size_t)k1 * p->M2 + n2] = to_f2(unit_root((long long)k1 * n2, M, -1));
    if (a.rader) {
        // width 2p: X0 position s L + n2 + M2 n1 holds the column of a_s at (n1, n2),
        // (p s + 2 g^q) mod 2p; positions 2L, 2L + 1 hold columns 0 and p.  Width p:
        // position n2 + M2 n1 holds column g^q, position L column 0 (rader_tables)
        const long long pr = a.rader_p, m1 = p->M1, m2 = p->M2;
        if (std::gcd(m1, m2) != 1) return fail(SPFFT_ERR_UNSUPPORTED, "Rader shape with non-coprime M1, M2");
        std::vector<int> gpos, kt, perm((size_t)cols);
        rader_tables(pr, m1, m2, gpos, kt, bh);
        if (cols == 2 * pr) {
            for (int pos = 0; pos < M; pos++) {
                perm[pos] = 2 * gpos[pos];
                perm[(size_t)M + pos] = (int)((pr + 2LL * gpos[pos]) % (2 * pr));
            }
            perm[2 * (size_t)M] = 0;
            perm[2 * (size_t)M + 1] = (int)pr;
        } else {
            for (int pos = 0; pos < M; pos++) perm[pos] = gpos[pos];
            perm[(size_t)M] = 0;
        }
        SPFFT_HIP(upload(p->rader_k, kt), "upload Rader indices");
        SPFFT_HIP(upload(p->colperm, perm), "upload Rader column order");
        a.rader_k = p->rader_k.as<int>();
        a.colperm = p->colperm.as<int>();
    } else if (!a.direct && !a.split_n1) {
        // b[0] = 1, b[k] = b[M-k] = exp(+i pi k^2 / C) for 1 <= k < C
        std::vector<cd> b((size_t)M, cd(0, 0));
        b[0] = cd(1, 0);
        for (int k = 1; k < cols; k++) {
            const long long k2 = (long long)k * k % (2LL * cols);
            const cd z = unit_root(k2, 2LL * cols, +1);
            b[k] = z;
            b[M - k] = z;
        }
        fft64(b, -1);
        const double inv_m = 1.0 / (double)M;
        for (int k1 = 0; k1 < p->M1; k1++)
            for (int k2 = 0; k2 < p->M2; k2++)
                bh[(size_t)k1 * p->M2 + k2] = to_f2(b[(size_t)k1 + (size_t)p->M1 * k2] * inv_m);
    }
    SPFFT_HIP(upload(p->h, h), "upload chirp");
    SPFFT_HIP(upload(p->wr_lo, lo), "upload twiddles");
    SPFFT_HIP(upload(p->wr_hi, hi), "upload twiddles");
    if (p->colpass_full) {
        std::vector<float2> full(rows);
        for (int k = 0; k < rows; k++) full[k] = to_f2(unit_root(k, rows, -1));
        SPFFT_HIP(upload(p->wr_full, full), "upload twiddles");
    }
    SPFFT_HIP(upload(p->tw2, t2), "upload twiddles");
    SPFFT_HIP(upload(p->twM, tm), "upload twiddles");
    if (a.fused) {
        // two-level W_M table of the fused row kernel: low 256 entries, then W_M^{256 i}
        std::vector<float2> w2(256 + (M + 255) / 256);
        for (int e = 0; e < 256; e++) w2[e] = to_f2(unit_root(e, M, -1));
        for (int i = 0; i < (M + 255) / 256; i++) w2[256 + i] = to_f2(unit_root(256LL * i, M, -1));
        SPFFT_HIP(upload(p->wm2, w2), "upload twiddles");
    }
    SPFFT_HIP(upload(p->bhat, bh), "upload B^");
    if (a.precision == 1 && a.rader) {
        // k_rowrader_mfma: each lane's A-operand fragments of the 32 x 32 matrices
        // C[k][n] = cos(2 pi k n / M1), S[k][n] = sin(2 pi k n / M1) (k, n < M1; zero
        // beyond), split into BF16 hi + lo parts from FP64: [C/S][hi/lo][k-step][lane][j],
        // lane l holding row k = l % 32, column n = 16 k-step + 8 (l / 32) + j
        const int L = p->M1;
        std::vector<uint16_t> fr((size_t)2 * 2 * 2 * 64 * 8);
        for (int d = 0; d < 2; d++)
            for (int ks = 0; ks < 2; ks++)
                for (int l = 0; l < 64; l++)
                    for (int j = 0; j < 8; j++) {
                        const int k = l % 32, n = 16 * ks + 8 * (l / 32) + j;
                        double v = 0.0;
                        if (k < L && n < L) {
                            const cd w = unit_root((long long)k * n, L, +1);  // cos + i sin
                            v = d == 0 ? w.real() : w.imag();
                        }
                        const float hi = bf16_round((float)v);
                        const float lo = bf16_round((float)(v - (double)hi));
                        uint32_t bh, bl;
                        std::memcpy(&bh, &hi, 4);
                        std::memcpy(&bl, &lo, 4);
                        fr[((((size_t)d * 2 + 0) * 2 + ks) * 64 + l) * 8 + j] = (uint16_t)(bh >> 16);
                        fr[((((size_t)d * 2 + 1) * 2 + ks) * 64 + l) * 8 + j] = (uint16_t)(bl >> 16);
                    }
        SPFFT_HIP(upload(p->mfma_frag, fr), "upload DFT fragments");
    } else if (a.precision == 1) {
        // cos and sin of 2 pi i / M1 split into BF16 hi + lo parts from FP64
        const int L = p->M1;
        std::vector<float> t((size_t)4 * L);
        for (int i = 0; i < L; i++) {
            const cd w = unit_root(i, L, +1);  // cos + i sin
            const double parts[2] = {w.real(), w.imag()};
            for (int k = 0; k < 2; k++) {
                const float hi = bf16_round((float)parts[k]);
                const float lo = bf16_round((float)(parts[k] - (double)hi));
                t[(size_t)(2 * k) * L + i] = hi;
                t[(size_t)(2 * k + 1) * L + i] = lo;
            }
        }
        SPFFT_HIP(upload(p->mfma_dft, t), "upload DFT matrix");
    }
    const int tile_rows = std::min(o.u_tile, rows / 2 + 1);
    if (own_workspace) {
        if (a.split_n1)
            SPFFT_HIP(p->T.alloc((size_t)tile_rows * a.split_n1 * a.split_p2 * sizeof(float2)),
                      "hipMalloc(tile buffer)");
        else if (!a.fused)
            SPFFT_HIP(p->T.alloc((size_t)p->t_rows * M * sizeof(float2)), "hipMalloc(tile buffer)");
        SPFFT_HIP(p->X0.alloc((size_t)tile_rows * a.ldx * sizeof(float2)), "hipMalloc(pass-1 buffer)");
    }
    SPFFT_HIP(p->counts.alloc((size_t)(cols + 1) * 4), "hipMalloc(counts)");
    SPFFT_HIP(p->colstart.alloc((size_t)(cols + 1) * 4), "hipMalloc(colstart)");
    SPFFT_HIP(p->sparse_cols.alloc((size_t)cols * (p->packed ? 2 : 4)), "hipMalloc(sparse_cols)");
    SPFFT_HIP(p->col_ptr.alloc((size_t)(cols + 1) * 4), "hipMalloc(col_ptr)");
    SPFFT_HIP(p->totals.alloc(2 * 4), "hipMalloc(totals)");
    SPFFT_HIP(p->err.alloc(4), "hipMalloc(err)");
    {
        size_t off_elems = 0;
        const size_t sched_elems = spfft::colsched_elems(cols, &off_elems);
        SPFFT_HIP(p->col_sched.alloc(sched_elems * 4), "hipMalloc(column schedule)");
        SPFFT_HIP(p->col_sched_off.alloc(off_elems * 4), "hipMalloc(column schedule)");
    }
    a.h = p->h.as<float2>();
    a.wr_hi = p->wr_hi.as<float2>();
    a.wr_lo = p->wr_lo.as<float2>();
    a.wr_full = p->colpass_full ? p->wr_full.as<float2>() : nullptr;
    a.tw2 = p->tw2.as<float2>();
    a.twM = p->twM.as<float2>();
    a.wm2 = a.fused ? p->wm2.as<float2>() : nullptr;
    a.bhat = p->bhat.as<float2>();
    a.mfma_dft = a.precision == 1 && !a.rader ? p->mfma_dft.as<float>() : nullptr;
    a.mfma_frag = a.precision == 1 && a.rader ? p->mfma_frag.p : nullptr;
    a.T = p->T.as<float2>();
    a.X0 = p->X0.as<float2>();
    if (a.fused)
        SPFFT_HIP(a.rader ? spfft::launch_rader(a, nullptr, true) : spfft::launch_fused(a, nullptr, true),
                  "fused row kernel LDS limit");
    p->table_bytes = p->h.bytes + p->wr_lo.bytes + p->wr_hi.bytes + p->wr_full.bytes + p->tw2.bytes + p->twM.bytes +
                     p->wm2.bytes + p->rader_k.bytes + p->colperm.bytes +
                     p->bhat.bytes + p->mfma_dft.bytes + p->mfma_frag.bytes + p->T.bytes + p->X0.bytes +
                     p->split_w.bytes + p->split_dft.bytes + p->rr_bhat.bytes + p->rr_tw2.bytes +
                     (p->split_sub ? p->split_sub->table_bytes : 0);
    *out = guard.release();
    return SPFFT_OK;
}

}  // namespace spfft_internal

extern "C" {

spfft_status spfft_plan_destroy(spfft_plan p) {
    delete p;
    return SPFFT_OK;
}

spfft_status spfft_plan_get_info(spfft_plan p, spfft_plan_info* info) {
    if (!p) return fail(SPFFT_ERR_INVALID_ARG, "NULL argument");
    spfft_status st0 = check_info(info);
    if (st0 != SPFFT_OK) return st0;
    fill_info(p, info);
    return SPFFT_OK;
}

spfft_status spfft_csc_build(spfft_plan p, const int* d_row, const int* d_col, int nnz, void* stream) {
    if (!p) return fail(SPFFT_ERR_INVALID_ARG, "NULL plan");
    return do_csc_build(*p, d_row, d_col, nnz, (hipStream_t)stream);
}

spfft_status spfft_check_input(spfft_plan p, void* stream) {
    if (!p) return fail(SPFFT_ERR_INVALID_ARG, "NULL plan");
    if (!p->have_csc) return fail(SPFFT_ERR_INVALID_ARG, "no CSC built yet");
    SPFFT_HIP(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
    int err = 0;
    SPFFT_HIP(hipMemcpy(&err, p->err.p, 4, hipMemcpyDeviceToHost), "hipMemcpy(err)");
    if (err) return fail(SPFFT_ERR_BAD_INPUT, "COO index outside [0, rows) x [0, cols)");
    return SPFFT_OK;
}

spfft_status spfft_csc_get(spfft_plan p, void* stream, spfft_csc_view* v) {
    if (!v) return fail(SPFFT_ERR_INVALID_ARG, "NULL view");
    spfft_status st = spfft_check_input(p, stream);
    if (st != SPFFT_OK) return st;
    int tot[2] = {0, 0};
    SPFFT_HIP(hipMemcpy(tot, p->totals.p, 8, hipMemcpyDeviceToHost), "hipMemcpy(totals)");
    v->nnz = tot[0];
    v->n_sparse_cols = tot[1];
    v->packed = p->packed ? 1 : 0;
    v->d_sparse_cols = p->sparse_cols.p;
    v->d_col_ptr = p->col_ptr.as<int>();
    v->d_row_idx = p->row_idx.p;
    v->d_col_start = p->colstart.as<int>();
    return SPFFT_OK;
}

spfft_status spfft_rowpass(spfft_plan p, void* d_out, int64_t pitch, void* stream) {
    if (!p) return fail(SPFFT_ERR_INVALID_ARG, "NULL plan");
    return do_rowpass(*p, d_out, pitch, (hipStream_t)stream, nullptr, nullptr);
}

spfft_status spfft_execute(spfft_plan p, const int* d_row, const int* d_col, int nnz, void* d_out,
                           int64_t pitch, void* stream) {
    if (!p) return fail(SPFFT_ERR_INVALID_ARG, "NULL plan");
    spfft_status st = do_csc_build(*p, d_row, d_col, nnz, (hipStream_t)stream);
    if (st != SPFFT_OK) return st;
    return do_rowpass(*p, d_out, pitch, (hipStream_t)stream, nullptr, nullptr);
}

spfft_status spfft_rowpass_to_host(spfft_plan p, void* h_out, int64_t pitch, void* stream) {
    if (!p) return fail(SPFFT_ERR_INVALID_ARG, "NULL plan");
    return do_rowpass_to_host(*p, h_out, pitch, (hipStream_t)stream);
}

spfft_status spfft_execute_to_host(spfft_plan p, const int* d_row, const int* d_col, int nnz, void* h_out,
                                   int64_t pitch, void* stream) {
    if (!p) return fail(SPFFT_ERR_INVALID_ARG, "NULL plan");
    if (!h_out) return fail(SPFFT_ERR_INVALID_ARG, "NULL host output pointer");
    if (pitch < p->V) return fail(SPFFT_ERR_INVALID_ARG, "output pitch smaller than cols/2+1");
    spfft_status st = do_csc_build(*p, d_row, d_col, nnz, (hipStream_t)stream);
    if (st != SPFFT_OK) return st;
    return do_rowpass_to_host(*p, h_out, pitch, (hipStream_t)stream);
}

spfft_status spfft_profile_enable(spfft_plan p, int on) {
    if (!p) return fail(SPFFT_ERR_INVALID_ARG, "NULL plan");
    p->profiling = on != 0;
    p->ev_used = 0;
    p->ev_groups.clear();
    return SPFFT_OK;
}

spfft_status spfft_profile_read(spfft_plan p, void* stream, spfft_profile* out) {
    if (!p || !out) return fail(SPFFT_ERR_INVALID_ARG, "NULL argument");
    std::memset(out, 0, sizeof(*out));
    SPFFT_HIP(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
    for (const auto& g : p->ev_groups) {
        hipEvent_t* ev = &p->ev_pool[g.second];
        float ms = 0.f;
        if (g.first == 0) {
            SPFFT_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]), "hipEventElapsedTime");
            out->ms[0] += ms;
            out->launches[0] += 1;
        } else {
            for (int k = 0; k < 4; k++) {
                SPFFT_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]), "hipEventElapsedTime");
                out->ms[1 + k] += ms;
                out->launches[1 + k] += 1;
            }
        }
    }
    p->ev_used = 0;
    p->ev_groups.clear();
    return SPFFT_OK;
}

spfft_status spfft_free_output(void* d) {
    if (d) SPFFT_HIP(hipFree(d), "hipFree");
    return SPFFT_OK;
}

// Host COO in; the output in device memory (h_out NULL: res->d_output, the caller
// frees it) or streamed to host memory h_out at pitch h_pitch.
static spfft_status run_host_coo(int rows, int cols, int nnz, const int* h_row, const int* h_col,
                                 const spfft_opts* opts, void* h_out, int64_t h_pitch, spfft_result* res) {
    spfft_status st = check_result(res);
    if (st != SPFFT_OK) return st;
    spfft_opts o;
    if ((st = read_opts(opts, &o)) != SPFFT_OK) return st;
    if (o.u_tile < 0) return fail(SPFFT_ERR_INVALID_ARG, "CSC tile size must be positive");
    if (rows <= 0 || cols <= 0 || nnz < 0) return fail(SPFFT_ERR_INVALID_ARG, "bad matrix shape");
    if (nnz > 0 && (!h_row || !h_col)) return fail(SPFFT_ERR_INVALID_ARG, "NULL COO pointer");
    if (!check_host_coo(rows, cols, nnz, h_row, h_col))
        return fail(SPFFT_ERR_BAD_INPUT, "COO index outside [0, rows) x [0, cols)");
    size_t free0 = 0, total = 0;
    SPFFT_HIP(hipMemGetInfo(&free0, &total), "hipMemGetInfo");
    spfft_plan p = nullptr;
    st = spfft_plan_create(rows, cols, &o, &p);
    if (st != SPFFT_OK) return st;
    std::unique_ptr<spfft_plan_s> pg(p);
    DevBuf drow, dcol, dout;
    SPFFT_HIP(drow.alloc((size_t)nnz * 4), "hipMalloc(COO)");
    SPFFT_HIP(dcol.alloc((size_t)nnz * 4), "hipMalloc(COO)");
    if (nnz > 0) {
        SPFFT_HIP(hipMemcpy(drow.p, h_row, (size_t)nnz * 4, hipMemcpyHostToDevice), "hipMemcpy(COO)");
        SPFFT_HIP(hipMemcpy(dcol.p, h_col, (size_t)nnz * 4, hipMemcpyHostToDevice), "hipMemcpy(COO)");
    }
    if (!h_out) SPFFT_HIP(dout.alloc((size_t)rows * p->P * sizeof(float2)), "hipMalloc(output)");
    st = ensure_csc_capacity(*p, nnz);
    if (st != SPFFT_OK) return st;
    SPFFT_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
    hipEvent_t e0, e1, e2;
    SPFFT_HIP(hipEventCreate(&e0), "hipEventCreate");
    SPFFT_HIP(hipEventCreate(&e1), "hipEventCreate");
    SPFFT_HIP(hipEventCreate(&e2), "hipEventCreate");
    struct EvGuard {
        hipEvent_t a, b, c;
        ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); (void)hipEventDestroy(c); }
    } evg{e0, e1, e2};
    SPFFT_HIP(hipEventRecord(e0, 0), "hipEventRecord");
    st = do_csc_build(*p, drow.as<int>(), dcol.as<int>(), nnz, 0);
    if (st != SPFFT_OK) return st;
    SPFFT_HIP(hipEventRecord(e1, 0), "hipEventRecord");
    st = h_out ? do_rowpass_to_host(*p, h_out, h_pitch, 0) : do_rowpass(*p, dout.p, p->P, 0, nullptr, nullptr);
    if (st != SPFFT_OK) return st;
    SPFFT_HIP(hipEventRecord(e2, 0), "hipEventRecord");
    SPFFT_HIP(hipEventSynchronize(e2), "hipEventSynchronize");
    float ms_loop = 0.f, ms_all = 0.f;
    SPFFT_HIP(hipEventElapsedTime(&ms_loop, e1, e2), "hipEventElapsedTime");
    SPFFT_HIP(hipEventElapsedTime(&ms_all, e0, e2), "hipEventElapsedTime");
    st = spfft_check_input(p, 0);
    if (st != SPFFT_OK) return st;
    size_t free1 = 0;
    SPFFT_HIP(hipMemGetInfo(&free1, &total), "hipMemGetInfo");
    res->d_output = dout.p;  // NULL when streamed
    dout.p = nullptr;        // ownership passes to the caller
    res->output_cols = h_out ? (int)h_pitch : p->P;
    res->ms = ms_loop;
    res->ms_e2e = ms_all;
    res->mem_bytes = free1 < free0 ? free0 - free1 : 0;
    res->fft_len = p->M;
    res->row_transform = p->base.direct ? SPFFT_TRANSFORM_DIRECT
                         : p->base.rader ? SPFFT_TRANSFORM_RADER
                                         : SPFFT_TRANSFORM_BLUESTEIN;
    return SPFFT_OK;
}

spfft_status spfft_run_host_coo(int rows, int cols, int nnz, const int* h_row, const int* h_col,
                                const spfft_opts* opts, spfft_result* res) {
    return run_host_coo(rows, cols, nnz, h_row, h_col, opts, nullptr, 0, res);
}

spfft_status spfft_run_host_coo_to_host(int rows, int cols, int nnz, const int* h_row, const int* h_col,
                                        const spfft_opts* opts, void* h_out, int64_t pitch,
                                        spfft_result* res) {
    if (!h_out) {
        if (check_result(res) == SPFFT_OK) {}
        return fail(SPFFT_ERR_INVALID_ARG, "NULL host output pointer");
    }
    return run_host_coo(rows, cols, nnz, h_row, h_col, opts, h_out, pitch, res);
}

spfft_status spfft_csc_bluestein_mixed_radix(int rows, int cols, int nnz, const int* h_row,
                                             const int* h_col, int u_tile, spfft_result* res) {
    if (u_tile <= 0) return fail(SPFFT_ERR_INVALID_ARG, "CSC tile size must be positive");
    spfft_opts o;
    spfft_reference_opts(SPFFT_SMOOTH_3, u_tile, &o);
    return spfft_run_host_coo(rows, cols, nnz, h_row, h_col, &o, res);
}

spfft_status spfft_csc_bluestein_cufft_smooth(int rows, int cols, int nnz, const int* h_row,
                                              const int* h_col, int u_tile, spfft_result* res) {
    if (u_tile <= 0) return fail(SPFFT_ERR_INVALID_ARG, "CSC tile size must be positive");
    spfft_opts o;
    spfft_reference_opts(SPFFT_SMOOTH_7, u_tile, &o);
    return spfft_run_host_coo(rows, cols, nnz, h_row, h_col, &o, res);
}

}  // extern "C"
