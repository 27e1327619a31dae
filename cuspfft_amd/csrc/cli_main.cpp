// cli_main.cpp -- benchmark driver for the two reported variants, keeping the
// reference CLI's flags and table format (reference src/main.cu:217-311 for
// the flags, :45-83 for the warm-up + median rule and the table lines,
// :486-524 and :649-659 for the variants).
//
//   spfft_cli <matrix.mtx | gen:c1 | gen:c2 | gen:c3 | gen:c5 | gen:p14> [options]
//     --csc-mixed-radix   Bluestein at next_3_smooth(2*cols-1)
//     --csc-mixed-radix-graph  the same, captured into one hipGraph and launched once
//     --csc-cufft-smooth  Bluestein at next_7_smooth(2*cols-1) (hand-written FFTs, no library)
//     --csc-stockham-stream        output streamed to pinned host memory (3-smooth length)
//     --csc-stockham-cufft-stream  the same at the 7-smooth length (reference main.cu:612-628)
//     --csc-tile N        output rows per tile (default 1024, as the reference)
//     --repeat N          1 warm-up + N timed runs; median, min, max
//     --sparse-only       accepted for compatibility (no dense baseline is built)
//     --cpu-reference     FFTW3 CPU reference; reported as unavailable when this
//                         build has no FFTW (the reference compiles it out then)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sparse_fft.h"
#include "../../include/spfft_c.h"

namespace {

COOMatrix load(const std::string& spec) {
    if (spec.rfind("gen:", 0) != 0) return read_mtx(spec);
    const std::string k = spec.substr(4);
    spfft_host_coo h;
    spfft_status st;
    if (k == "c1") st = spfft_coo_generate("uniform", 2310, 8000, 0, 0, 0, 0, &h);
    else if (k == "c2") st = spfft_coo_generate("banded", 8219, 242700, 300, 1200, 0.02, 0, &h);
    else if (k == "c3") st = spfft_coo_generate("chunglu", 16706, 121250, 2.1, 360, 0, 0, &h);
    else if (k == "c5") st = spfft_coo_generate("chunglu", 40421, 351400, 2.2, 0, 0, 0, &h);
    else if (k == "p14") st = spfft_coo_generate("uniform", 16384, 49122, 0, 0, 0, 0, &h);  // delaunay_n14's size
    else throw std::runtime_error("unknown generator " + spec);
    if (st != SPFFT_OK) throw std::runtime_error("generator failed");
    COOMatrix c;
    c.rows = h.rows;
    c.cols = h.cols;
    c.nnz = h.nnz;
    c.row_idx.assign(h.row_idx, h.row_idx + h.nnz);
    c.col_idx.assign(h.col_idx, h.col_idx + h.nnz);
    spfft_coo_free(&h);
    return c;
}

template <typename Fn>
SparseFFTResult run_repeated(int n_iters, Fn&& fn, float* mn, float* mx) {
    SparseFFTResult result{};
    std::vector<float> times;
    if (n_iters > 1) {
        result = fn();
        if (result.d_output) (void)hipFree(result.d_output);
        result.d_output = nullptr;
    }
    for (int i = 0; i < n_iters; i++) {
        if (i > 0 && result.d_output) (void)hipFree(result.d_output);
        result = fn();
        times.push_back(result.ms);
    }
    if (n_iters > 1) {
        std::sort(times.begin(), times.end());
        result.ms = times[n_iters / 2];
        *mn = times.front();
        *mx = times.back();
    }
    return result;
}

void print_bench(const char* label, float ms, size_t mem, int n, float mn, float mx) {
    if (n > 1)
        printf("%-26s  %10.3f  %12.2f  [min %.3f, max %.3f]  n=%d\n", label, ms, mem / 1e6, mn, mx, n);
    else
        printf("%-26s  %10.3f  %12.2f\n", label, ms, mem / 1e6);
}

void usage(const char* p) {
    fprintf(stderr,
            "Usage: %s <matrix.mtx | gen:c1|c2|c3|c5|p14> [--csc-mixed-radix] [--csc-mixed-radix-graph]\n"
            "          [--csc-cufft-smooth] [--csc-stockham-stream] [--csc-stockham-cufft-stream]\n"
            "          [--csc-tile N] [--repeat N] [--sparse-only] [--cpu-reference]\n",
            p);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) { usage(argv[0]); return 1; }
    if (!std::strcmp(argv[1], "--help") || !std::strcmp(argv[1], "-h")) { usage(argv[0]); return 0; }
    bool mixed = false, mixed_graph = false, smooth = false, cpu_ref = false;
    bool stream3 = false, stream7 = false;
    int tile = 1024, repeat = 1;
    try {
        for (int i = 2; i < argc; i++) {
            const std::string f(argv[i]);
            if (f == "--sparse-only") {}
            else if (f == "--csc-mixed-radix") mixed = true;
            else if (f == "--csc-mixed-radix-graph") mixed_graph = true;
            else if (f == "--csc-cufft-smooth") smooth = true;
            else if (f == "--csc-stockham-stream") stream3 = true;
            else if (f == "--csc-stockham-cufft-stream") stream7 = true;
            else if (f == "--cpu-reference") cpu_ref = true;
            else if (f == "--csc-tile") {
                if (++i >= argc) throw std::runtime_error("--csc-tile requires an integer value");
                tile = std::stoi(argv[i]);
                if (tile <= 0) throw std::runtime_error("--csc-tile must be positive");
            } else if (f == "--repeat") {
                if (++i >= argc) throw std::runtime_error("--repeat requires an integer value");
                repeat = std::stoi(argv[i]);
                if (repeat <= 0) throw std::runtime_error("--repeat must be positive");
            } else {
                throw std::runtime_error("Unknown option: " + f);
            }
        }
        printf("Matrix: %s\n", argv[1]);
        COOMatrix coo = load(argv[1]);
        printf("Size:   %d x %d\n", coo.rows, coo.cols);
        printf("nnz:    %d  (density %.4f%%)\n\n", coo.nnz, 100.0 * coo.nnz / ((double)coo.rows * coo.cols));
        printf("%-26s  %10s  %12s\n", "Method", "Time (ms)", "Mem (MB)");
        printf("%-26s  %10s  %12s\n", "--------------------------", "----------", "------------");
        if (cpu_ref)
            printf("%-26s  %10s  %12s  [%s]\n", "CPU reference (FFTW3)", "FAIL", "FAIL",
                   "FFTW3 not available in this build");
        struct V { bool on; const char* label; int kind; };
        const V vs[] = {{mixed, "Sparse CSC (mixed-radix)", 0},
                        {mixed_graph, "Sparse CSC (mr-graph)", 1},
                        {smooth, "Sparse CSC (cufft smooth)", 2},
                        {stream3, "Sparse CSC (Stockham stream)", 3},
                        {stream7, "Sparse CSC (cufft stream)", 4}};
        for (const V& v : vs) {
            if (!v.on) continue;
            try {
                float mn = 0, mx = 0;
                SparseFFTResult r = run_repeated(
                    repeat,
                    [&] {
                        if (v.kind == 4) return sparse_fft_csc_bluestein_cufft_streaming(coo, tile);
                        if (v.kind == 3) return sparse_fft_csc_stockham_streaming(coo, tile, 0);
                        if (v.kind == 2) return sparse_fft_csc_bluestein_cufft_smooth(coo, tile);
                        if (v.kind == 1) return sparse_fft_csc_bluestein_mixed_radix_graph(coo, tile);
                        return sparse_fft_csc_bluestein_mixed_radix(coo, tile);
                    },
                    &mn, &mx);
                print_bench(v.label, r.ms, r.mem_bytes, repeat, mn, mx);
                printf("  rate %-20s  %.2f FFTs/s (tile loop%s)\n", v.label, 1000.0 / r.ms,
                       v.kind >= 3 ? " with device-to-host copies" : "");
                (void)hipFree(r.d_output);
            } catch (const std::exception& e) {
                printf("%-26s  %10s  %12s  [%s]\n", v.label, "OOM", "OOM", e.what());
                (void)hipGetLastError();
            }
        }
    } catch (const std::exception& e) {
        fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
    return 0;
}
