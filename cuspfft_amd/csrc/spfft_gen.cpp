// spfft_gen.cpp -- seeded synthetic binary matrices and the Matrix Market
// reader (host only, no device calls).
//
// Every matrix is returned in the canonical COO form the reference's read_mtx
// produces (src/mtx_reader.cu:60-74): duplicates removed, sorted by (row, col).
// All generators emit symmetric matrices with a zero diagonal.
//
// RNG: xoshiro256** (Blackman & Vigna) seeded by four SplitMix64 outputs of the
// 64-bit seed; uniform doubles are (x >> 11) * 2^-53.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_set>
#include <utility>
#include <vector>

#include "../../include/spfft_c.h"

namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& z) {
        uint64_t x = (z += 0x9E3779B97F4A7C15ULL);
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
        return x ^ (x >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (int i = 0; i < 4; i++) s[i] = splitmix(z);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
};

typedef std::vector<std::pair<int, int>> Pairs;

// Mirror i<j pairs, sort, dedup -> host COO owned by malloc.
spfft_status emit(int n, const Pairs& upper, spfft_host_coo* out) {
    Pairs all;
    all.reserve(upper.size() * 2);
    for (const auto& e : upper) {
        all.push_back(e);
        all.push_back({e.second, e.first});
    }
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    out->rows = n;
    out->cols = n;
    out->nnz = (int)all.size();
    out->row_idx = (int*)std::malloc(sizeof(int) * std::max<size_t>(all.size(), 1));
    out->col_idx = (int*)std::malloc(sizeof(int) * std::max<size_t>(all.size(), 1));
    if (!out->row_idx || !out->col_idx) return SPFFT_ERR_ALLOC;
    for (size_t i = 0; i < all.size(); i++) {
        out->row_idx[i] = all[i].first;
        out->col_idx[i] = all[i].second;
    }
    return SPFFT_OK;
}

// Uniform G(n, m): m distinct pairs i<j.
void gen_uniform(int n, long long m, Rng& rng, Pairs& e) {
    const long long maxp = (long long)n * (n - 1) / 2;
    if (m > maxp) m = maxp;
    std::unordered_set<uint64_t> seen;
    seen.reserve((size_t)m * 2);
    while ((long long)e.size() < m) {
        int i = (int)rng.below(n), j = (int)rng.below(n);
        if (i == j) continue;
        if (i > j) std::swap(i, j);
        if (seen.insert((uint64_t)i * n + j).second) e.push_back({i, j});
    }
}

// Banded: pair (i, i+d), 1 <= d <= cutoff, kept with probability
// p0 * exp(-d / scale), p0 chosen so that the expected stored nnz of the band
// is (1 - frac) * target; then frac * target / 2 distinct uniform pairs with
// |i - j| > cutoff.
void gen_banded(int n, double target, double scale, int cutoff, double frac, Rng& rng, Pairs& e) {
    double mass = 0;
    for (int d = 1; d <= cutoff && d < n; d++) mass += (double)(n - d) * std::exp(-d / scale);
    const double band_pairs = 0.5 * target * (1.0 - frac);
    const double p0 = mass > 0 ? std::min(1.0, band_pairs / mass) : 0.0;
    std::vector<double> pd(cutoff + 1);
    for (int d = 1; d <= cutoff; d++) pd[d] = p0 * std::exp(-d / scale);
    for (int i = 0; i < n; i++)
        for (int d = 1; d <= cutoff && i + d < n; d++)
            if (rng.uniform() < pd[d]) e.push_back({i, i + d});
    const long long off = std::llround(0.5 * target * frac);
    long long far_pairs = 0;
    if (n - 1 > cutoff) far_pairs = (long long)(n - 1 - cutoff) * (n - cutoff) / 2;
    const long long want = std::min(off, far_pairs);
    std::unordered_set<uint64_t> seen;
    long long got = 0;
    while (got < want) {
        int i = (int)rng.below(n), j = (int)rng.below(n);
        if (i > j) std::swap(i, j);
        if (j - i <= cutoff) continue;
        if (seen.insert((uint64_t)i * n + j).second) {
            e.push_back({i, j});
            got++;
        }
    }
}

// Chung-Lu: w_i = c (i + i0)^(-1/(gamma-1)), w_0 = wmax, sum w = target; edge
// i<j with probability min(1, w_i w_j / sum w), sampled by the skipping method
// of Miller & Hagberg (2011); node labels then permuted by a seeded shuffle.
void gen_chunglu(int n, double target, double gamma, double wmax, Rng& rng, Pairs& e) {
    if (n < 2) return;
    if (wmax <= 0) wmax = std::sqrt(target);
    const double alpha = 1.0 / (gamma - 1.0);
    auto sum_for = [&](double i0) {
        double s = 0;
        for (int i = 0; i < n; i++) s += std::pow(i0 / (i + i0), alpha);
        return wmax * s;
    };
    std::vector<double> w(n);
    if (wmax * n <= target) {
        std::fill(w.begin(), w.end(), wmax);
    } else {
        double lo = 1e-9, hi = 1e12;
        for (int it = 0; it < 200; it++) {
            const double mid = std::sqrt(lo * hi);
            if (sum_for(mid) < target) lo = mid; else hi = mid;
        }
        const double i0 = std::sqrt(lo * hi);
        for (int i = 0; i < n; i++) w[i] = wmax * std::pow(i0 / (i + i0), alpha);
    }
    double S = 0;
    for (double x : w) S += x;
    for (int u = 0; u + 1 < n; u++) {
        int v = u + 1;
        double p = std::min(w[u] * w[v] / S, 1.0);
        while (v < n && p > 0) {
            if (p != 1.0) {
                const double r = 1.0 - rng.uniform();  // (0, 1]
                v += (int)std::floor(std::log(r) / std::log1p(-p));
            }
            if (v < n) {
                const double q = std::min(w[u] * w[v] / S, 1.0);
                const double r = rng.uniform();
                if (r < q / p) e.push_back({u, v});
                p = q;
                v++;
            }
        }
    }
    std::vector<int> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int i = n - 1; i > 0; i--) std::swap(perm[i], perm[(int)rng.below((uint64_t)i + 1)]);
    for (auto& x : e) {
        int a = perm[x.first], b = perm[x.second];
        if (a > b) std::swap(a, b);
        x = {a, b};
    }
}

}  // namespace

extern "C" {

spfft_status spfft_coo_generate(const char* kind, int n, double p0, double p1, double p2, double p3,
                                uint64_t seed, spfft_host_coo* out) {
    if (!kind || !out || n <= 0) return SPFFT_ERR_INVALID_ARG;
    std::memset(out, 0, sizeof(*out));
    Rng rng(seed);
    Pairs e;
    const std::string k(kind);
    if (k == "uniform") {
        gen_uniform(n, (long long)std::llround(p0), rng, e);
    } else if (k == "banded") {
        gen_banded(n, p0, p1 > 0 ? p1 : 300.0, p2 > 0 ? (int)p2 : 1200, p3, rng, e);
    } else if (k == "chunglu") {
        if (p1 <= 1.0) return SPFFT_ERR_INVALID_ARG;
        gen_chunglu(n, p0, p1, p2, rng, e);
    } else {
        return SPFFT_ERR_INVALID_ARG;
    }
    return emit(n, e, out);
}

void spfft_coo_free(spfft_host_coo* coo) {
    if (!coo) return;
    std::free(coo->row_idx);
    std::free(coo->col_idx);
    coo->row_idx = coo->col_idx = nullptr;
    coo->nnz = 0;
}

// src/mtx_reader.cu:7-77 semantics.
spfft_status spfft_read_mtx(const char* path, spfft_host_coo* out) {
    if (!path || !out) return SPFFT_ERR_INVALID_ARG;
    std::memset(out, 0, sizeof(*out));
    std::ifstream f(path);
    if (!f.is_open()) return SPFFT_ERR_INVALID_ARG;
    std::string line;
    if (!std::getline(f, line) || line.rfind("%%MatrixMarket", 0) != 0) return SPFFT_ERR_BAD_INPUT;
    const bool symmetric = line.find("symmetric") != std::string::npos ||
                           line.find("Symmetric") != std::string::npos;
    while (std::getline(f, line))
        if (!line.empty() && line[0] != '%') break;
    int rows, cols, stored;
    {
        std::istringstream ss(line);
        if (!(ss >> rows >> cols >> stored)) return SPFFT_ERR_BAD_INPUT;
    }
    if (rows <= 0 || cols <= 0 || stored < 0) return SPFFT_ERR_BAD_INPUT;
    Pairs all;
    all.reserve((size_t)stored * (symmetric ? 2 : 1));
    for (int i = 0; i < stored; i++) {
        if (!std::getline(f, line)) return SPFFT_ERR_BAD_INPUT;
        std::istringstream ss(line);
        int r, c;
        if (!(ss >> r >> c)) return SPFFT_ERR_BAD_INPUT;
        r--;
        c--;
        all.push_back({r, c});
        if (symmetric && r != c) all.push_back({c, r});
    }
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    out->rows = rows;
    out->cols = cols;
    out->nnz = (int)all.size();
    out->row_idx = (int*)std::malloc(sizeof(int) * std::max<size_t>(all.size(), 1));
    out->col_idx = (int*)std::malloc(sizeof(int) * std::max<size_t>(all.size(), 1));
    if (!out->row_idx || !out->col_idx) return SPFFT_ERR_ALLOC;
    for (size_t i = 0; i < all.size(); i++) {
        out->row_idx[i] = all[i].first;
        out->col_idx[i] = all[i].second;
    }
    return SPFFT_OK;
}

}  // extern "C"
