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
//     --csc-auto          this build's native plan (direct or Rader row transform
//                         where the width has one, else Bluestein)
//     --cpu-reference     the CPU reference row: a dense FP32 2-D R2C FFT on the host,
//                         single-threaded, only the transform timed (the reference runs
//                         FFTW3 here; this build has no FFTW, so the transform is its own,
//                         cpu_fft.cpp, and the row says so); its output is the ground
//                         truth of every variant's check line, as in the reference
//     --check-ref FILE    after each variant, the reference's check line (max_abs,
//                         max_rel, rms_abs) against a reference half spectrum
//                         ("SPFFTREF", int32 rows, int32 cols, complex128 values);
//                         overrides --cpu-reference as the ground truth
//     --write-ref FILE    write the --cpu-reference spectrum in that format
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sparse_fft.h"
#include "../../include/spfft_c.h"
#include "cpu_fft.h"

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

// Timing contract of the reference driver (src/main.cu:45-70): with n > 1 one
// untimed warm-up call comes first, then n timed calls; the reported time is the
// upper median sorted[n / 2] with the extremes beside it.  Every output but the
// last is released as soon as it is replaced.
struct Timing {
    float median = 0.f, lo = 0.f, hi = 0.f;
};

template <typename Fn>
SparseFFTResult timed_calls(int n, Fn&& call, Timing* t) {
    auto release = [](SparseFFTResult& r) {
        if (r.d_output) (void)hipFree(r.d_output);
        r.d_output = nullptr;
    };
    SparseFFTResult last{};
    if (n > 1) {
        last = call();  // warm-up
        release(last);
    }
    std::vector<float> ms;
    ms.reserve(n);
    while ((int)ms.size() < n) {
        release(last);
        last = call();
        ms.push_back(last.ms);
    }
    std::sort(ms.begin(), ms.end());
    t->median = ms[ms.size() / 2];
    t->lo = ms.front();
    t->hi = ms.back();
    return last;
}

// One table row: label, time (ms), device memory (MB); with n > 1 also the range.
void print_row(const char* label, const Timing& t, size_t mem, int n) {
    printf("%-26s  %10.3f  %12.2f", label, t.median, mem / 1e6);
    if (n > 1) printf("  [min %.3f, max %.3f]  n=%d", t.lo, t.hi, n);
    printf("\n");
}

// A reference half spectrum written by the test tooling: the 8 bytes "SPFFTREF",
// int32 rows, int32 cols, then rows x (cols/2 + 1) complex128 values, row-major.
struct RefSpectrum {
    int rows = 0, cols = 0;
    std::vector<double> re_im;  // interleaved
};

RefSpectrum load_reference(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open reference spectrum " + path);
    char magic[8];
    int dims[2];
    RefSpectrum r;
    const bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "SPFFTREF", 8) == 0 &&
                    std::fread(dims, sizeof(int), 2, f) == 2 && dims[0] > 0 && dims[1] > 0;
    if (ok) {
        r.rows = dims[0];
        r.cols = dims[1];
        r.re_im.resize((size_t)r.rows * (r.cols / 2 + 1) * 2);
    }
    const bool full = ok && std::fread(r.re_im.data(), sizeof(double), r.re_im.size(), f) == r.re_im.size();
    std::fclose(f);
    if (!full) throw std::runtime_error("malformed reference spectrum " + path);
    return r;
}

void write_reference(const std::string& path, const RefSpectrum& r) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write reference spectrum " + path);
    const int dims[2] = {r.rows, r.cols};
    const bool ok = std::fwrite("SPFFTREF", 1, 8, f) == 8 && std::fwrite(dims, sizeof(int), 2, f) == 2 &&
                    std::fwrite(r.re_im.data(), sizeof(double), r.re_im.size(), f) == r.re_im.size();
    if (std::fclose(f) != 0 || !ok) throw std::runtime_error("cannot write reference spectrum " + path);
}

// The CPU reference row (reference src/cpu_reference.cpp:8-64): the binary matrix as a
// dense R x C float array, then the timed 2-D R2C transform (spfft_cpu::rfft2).
struct CpuResult {
    RefSpectrum spectrum;
    float ms = 0.f;
    size_t mem_bytes = 0;
};

CpuResult cpu_reference(const COOMatrix& coo) {
    const size_t R = coo.rows, C = coo.cols, V = C / 2 + 1;
    std::vector<float> dense(R * C, 0.f);
    for (int i = 0; i < coo.nnz; i++) {
        const int r = coo.row_idx[i], c = coo.col_idx[i];
        if (r < 0 || c < 0 || (size_t)r >= R || (size_t)c >= C)
            throw std::runtime_error("COO index out of range");
        dense[(size_t)r * C + c] = 1.f;
    }
    std::vector<spfft_cpu::cf> out(R * V);
    const auto t0 = std::chrono::steady_clock::now();
    spfft_cpu::rfft2((int)R, (int)C, dense.data(), out.data());
    const auto t1 = std::chrono::steady_clock::now();
    CpuResult res;
    res.ms = (float)std::chrono::duration<double, std::milli>(t1 - t0).count();
    res.mem_bytes = dense.size() * sizeof(float) + out.size() * sizeof(spfft_cpu::cf);
    res.spectrum.rows = (int)R;
    res.spectrum.cols = (int)C;
    res.spectrum.re_im.resize(out.size() * 2);
    for (size_t i = 0; i < out.size(); i++) {
        res.spectrum.re_im[2 * i] = out[i].real();
        res.spectrum.re_im[2 * i + 1] = out[i].imag();
    }
    return res;
}

// The reference driver's error triple (src/main.cu:168-215): over every entry of the
// half spectrum, max |x - ref|, max |x - ref| / max(1, |ref|), and the root mean square
// of |x - ref|; printed in the same line format.
void print_check(const char* label, const RefSpectrum& ref, const SparseFFTResult& r) {
    const int V = ref.cols / 2 + 1;
    std::vector<hipFloatComplex> got((size_t)ref.rows * V);
    if (r.d_output) {
        if (hipMemcpy2D(got.data(), (size_t)V * sizeof(hipFloatComplex), r.d_output,
                        (size_t)r.output_cols * sizeof(hipFloatComplex), (size_t)V * sizeof(hipFloatComplex),
                        ref.rows, hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("hipMemcpy2D of the output failed");
    } else {
        for (int u = 0; u < ref.rows; u++)
            std::copy_n(r.h_output.data() + (size_t)u * r.output_cols, V, got.data() + (size_t)u * V);
    }
    double worst_abs = 0.0, worst_rel = 0.0, sq = 0.0;
    for (size_t i = 0; i < got.size(); i++) {
        const double xr = ref.re_im[2 * i], xi = ref.re_im[2 * i + 1];
        const double e = std::hypot((double)got[i].x - xr, (double)got[i].y - xi);
        worst_abs = std::max(worst_abs, e);
        worst_rel = std::max(worst_rel, e / std::max(1.0, std::hypot(xr, xi)));
        sq += e * e;
    }
    printf("  check %-19s  max_abs %.3e  max_rel %.3e  rms_abs %.3e\n", label, worst_abs, worst_rel,
           std::sqrt(sq / (double)got.size()));
}

// The native plan of this build (row_transform AUTO: the direct or Rader transform
// where the width has one, else Bluestein), through the C ABI.
SparseFFTResult run_auto(const COOMatrix& coo, int tile) {
    spfft_opts o;
    spfft_opts_default(&o);
    o.u_tile = tile;
    spfft_result r;
    r.struct_size = sizeof(r);
    if (spfft_run_host_coo(coo.rows, coo.cols, coo.nnz, coo.row_idx.data(), coo.col_idx.data(), &o, &r) !=
        SPFFT_OK)
        throw std::runtime_error(spfft_last_error());
    SparseFFTResult out{};
    out.d_output = static_cast<hipFloatComplex*>(r.d_output);
    out.output_cols = r.output_cols;
    out.ms = r.ms;
    out.mem_bytes = r.mem_bytes;
    return out;
}

void usage(const char* p) {
    fprintf(stderr,
            "Usage: %s <matrix.mtx | gen:c1|c2|c3|c5|p14> [--csc-mixed-radix] [--csc-mixed-radix-graph]\n"
            "          [--csc-cufft-smooth] [--csc-stockham-stream] [--csc-stockham-cufft-stream]\n"
            "          [--csc-auto] [--csc-tile N] [--repeat N] [--sparse-only] [--cpu-reference]\n"
            "          [--check-ref FILE] [--write-ref FILE]\n",
            p);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) { usage(argv[0]); return 1; }
    if (!std::strcmp(argv[1], "--help") || !std::strcmp(argv[1], "-h")) { usage(argv[0]); return 0; }
    bool mixed = false, mixed_graph = false, smooth = false, cpu_ref = false;
    bool stream3 = false, stream7 = false, native = false;
    int tile = 1024, repeat = 1;
    std::string ref_path, write_path;
    try {
        for (int i = 2; i < argc; i++) {
            const std::string f(argv[i]);
            if (f == "--sparse-only") {}
            else if (f == "--csc-mixed-radix") mixed = true;
            else if (f == "--csc-mixed-radix-graph") mixed_graph = true;
            else if (f == "--csc-cufft-smooth") smooth = true;
            else if (f == "--csc-stockham-stream") stream3 = true;
            else if (f == "--csc-stockham-cufft-stream") stream7 = true;
            else if (f == "--csc-auto") native = true;
            else if (f == "--cpu-reference") cpu_ref = true;
            else if (f == "--write-ref") {
                if (++i >= argc) throw std::runtime_error("--write-ref requires a file name");
                write_path = argv[i];
            }
            else if (f == "--check-ref") {
                if (++i >= argc) throw std::runtime_error("--check-ref requires a file name");
                ref_path = argv[i];
            }
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
        if (!write_path.empty() && !cpu_ref) throw std::runtime_error("--write-ref needs --cpu-reference");
        printf("Matrix: %s\n", argv[1]);
        COOMatrix coo = load(argv[1]);
        printf("Size:   %d x %d\n", coo.rows, coo.cols);
        printf("nnz:    %d  (density %.4f%%)\n\n", coo.nnz, 100.0 * coo.nnz / ((double)coo.rows * coo.cols));
        printf("%-26s  %10s  %12s\n", "Method", "Time (ms)", "Mem (MB)");
        printf("%-26s  %10s  %12s\n", "--------------------------", "----------", "------------");
        RefSpectrum ref;
        if (cpu_ref) {
            // the reference's CPU row (src/main.cu:324-342): dense scatter untimed, only the
            // transform timed; its output becomes the ground truth of the check lines
            try {
                CpuResult cpu = cpu_reference(coo);
                printf("%-26s  %10.3f  %12.2f  [CPU, own FP32 FFT single-thread; no FFTW3 in this build]\n",
                       "CPU reference (FP32 FFT)", cpu.ms, cpu.mem_bytes / 1e6);
                ref = std::move(cpu.spectrum);
                if (!write_path.empty()) write_reference(write_path, ref);
            } catch (const std::exception& e) {
                printf("%-26s  %10s  %12s  [%s]\n", "CPU reference (FP32 FFT)", "FAIL", "FAIL", e.what());
            }
        }
        if (!ref_path.empty()) {
            ref = load_reference(ref_path);
            if (ref.rows != coo.rows || ref.cols != coo.cols)
                throw std::runtime_error("reference spectrum shape does not match the matrix");
        }
        struct V { bool on; const char* label; int kind; };
        const V vs[] = {{mixed, "Sparse CSC (mixed-radix)", 0},
                        {mixed_graph, "Sparse CSC (mr-graph)", 1},
                        {smooth, "Sparse CSC (cufft smooth)", 2},
                        {stream3, "Sparse CSC (Stockham stream)", 3},
                        {stream7, "Sparse CSC (cufft stream)", 4},
                        {native, "Sparse CSC (auto)", 5}};
        for (const V& v : vs) {
            if (!v.on) continue;
            try {
                Timing t;
                SparseFFTResult r = timed_calls(
                    repeat,
                    [&] {
                        switch (v.kind) {
                            case 1: return sparse_fft_csc_bluestein_mixed_radix_graph(coo, tile);
                            case 2: return sparse_fft_csc_bluestein_cufft_smooth(coo, tile);
                            case 3: return sparse_fft_csc_stockham_streaming(coo, tile, 0);
                            case 4: return sparse_fft_csc_bluestein_cufft_streaming(coo, tile);
                            case 5: return run_auto(coo, tile);
                            default: return sparse_fft_csc_bluestein_mixed_radix(coo, tile);
                        }
                    },
                    &t);
                print_row(v.label, t, r.mem_bytes, repeat);
                printf("  rate %-20s  %.2f FFTs/s (tile loop%s)\n", v.label, 1000.0 / t.median,
                       v.kind == 3 || v.kind == 4 ? " with device-to-host copies" : "");
                if (ref.rows) print_check(v.label, ref, r);
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
