// spfft_batch.cpp -- hipGraph capture of whole transforms and the batched form.
//
// Graph form: the CSC build and the tile loop of one plan captured once into a
// hipGraph and replayed with a single launch.  It replaces
// sparse_fft_csc_bluestein_mixed_radix_graph (reference
// src/sparse_fft_bluestein.cu:2402-2549, capture :2473-2520, single replay :2526).
//
// Batched form: many matrices of uneven shape and nnz, each with its own plan
// (its own Bluestein length, split and kernels), run as one launch sequence.
// Matrices are assigned longest-first to S worker streams forked from the
// caller's stream with events and joined back at the end; the streams of one
// slot share one row-pass workspace (the largest T and X0 among its matrices),
// so device memory is S workspaces plus the per-matrix tables, CSC arrays and
// the caller's output arena.  The whole fork / run / join sequence can be
// captured into one hipGraph (spfft_batch_capture) and replayed.
//
// All streams are created at the default priority.  The captured work has no
// memset nodes: every buffer that must start at zero is cleared by a kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <numeric>
#include <vector>

#include "../../include/spfft_c.h"
#include "spfft_plan_internal.h"

using namespace spfft_internal;

struct spfft_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~spfft_graph_s() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

struct spfft_batch_s {
    int count = 0;
    std::vector<std::unique_ptr<spfft_plan_s>> plans;
    std::vector<int64_t> out_off;  // element offset of each matrix in the output arena
    int64_t out_total = 0;
    int nslots = 1;
    std::vector<int> slot_of;      // matrix -> worker stream
    std::vector<std::vector<int>> slot_list;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> join;  // one per slot
    hipEvent_t fork = nullptr;
    std::vector<DevBuf> ws_T, ws_X0;
    ~spfft_batch_s() {
        for (hipStream_t s : streams) (void)hipStreamDestroy(s);
        for (hipEvent_t e : join) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
    }
};

namespace {

// Work estimate used for the longest-first assignment: pass 1 ~ rows * nnz,
// row pass ~ rows * M * log2(M) (halved rows in both).
double work_estimate(const spfft_plan_s& p, int nnz) {
    const double rows = p.R / 2 + 1;
    return rows * (double)nnz + 4.0 * rows * (double)p.M * std::log2((double)std::max(p.M, 2));
}

spfft_status run_batch(spfft_batch_s& b, const int* d_row, const int* d_col, const int64_t* coo_off,
                       void* d_out, hipStream_t s) {
    if (!coo_off) return fail(SPFFT_ERR_INVALID_ARG, "NULL COO offsets");
    if (!d_out) return fail(SPFFT_ERR_INVALID_ARG, "NULL output arena");
    SPFFT_HIP(hipEventRecord(b.fork, s), "hipEventRecord");
    for (int k = 0; k < b.nslots; k++) SPFFT_HIP(hipStreamWaitEvent(b.streams[k], b.fork, 0), "hipStreamWaitEvent");
    for (int k = 0; k < b.nslots; k++) {
        for (int i : b.slot_list[k]) {
            spfft_plan_s& p = *b.plans[i];
            const int64_t n = coo_off[i + 1] - coo_off[i];
            if (n < 0 || n > 0x7fffffff) return fail(SPFFT_ERR_INVALID_ARG, "bad COO offsets");
            spfft_status st = do_csc_build(p, d_row + coo_off[i], d_col + coo_off[i], (int)n, b.streams[k]);
            if (st != SPFFT_OK) return st;
            st = do_rowpass(p, static_cast<float2*>(d_out) + b.out_off[i], p.P, b.streams[k],
                            b.ws_T[k].as<float2>(), b.ws_X0[k].as<float2>());
            if (st != SPFFT_OK) return st;
        }
        SPFFT_HIP(hipEventRecord(b.join[k], b.streams[k]), "hipEventRecord");
        SPFFT_HIP(hipStreamWaitEvent(s, b.join[k], 0), "hipStreamWaitEvent");
    }
    return SPFFT_OK;
}

}  // namespace

extern "C" {

spfft_status spfft_graph_launch(spfft_graph g, void* stream) {
    if (!g) return fail(SPFFT_ERR_INVALID_ARG, "NULL graph");
    SPFFT_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream), "hipGraphLaunch");
    return SPFFT_OK;
}

spfft_status spfft_graph_destroy(spfft_graph g) {
    delete g;
    return SPFFT_OK;
}

spfft_status spfft_graph_capture(spfft_plan p, const int* d_row, const int* d_col, int nnz, void* d_out,
                                 int64_t pitch, void* stream, spfft_graph* out) {
    (void)stream;
    if (!p || !out) return fail(SPFFT_ERR_INVALID_ARG, "NULL argument");
    *out = nullptr;
    spfft_status st = ensure_csc_capacity(*p, nnz);  // no allocation may happen while capturing
    if (st != SPFFT_OK) return st;
    const bool prof = p->profiling;
    p->profiling = false;
    hipStream_t cs;
    SPFFT_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate");
    std::unique_ptr<spfft_graph_s> g(new (std::nothrow) spfft_graph_s());
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        st = do_csc_build(*p, d_row, d_col, nnz, cs);
        if (st == SPFFT_OK) st = do_rowpass(*p, d_out, pitch, cs, nullptr, nullptr);
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(cs, &graph);
        g->graph = graph;
        if (e == hipSuccess && st == SPFFT_OK) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    }
    (void)hipStreamDestroy(cs);
    p->profiling = prof;
    if (st != SPFFT_OK) return st;
    if (e != hipSuccess) return hip_fail(e, "graph capture");
    *out = g.release();
    return SPFFT_OK;
}

spfft_status spfft_batch_create(int count, const int* rows, const int* cols, const int* nnz_hint,
                                const spfft_opts* opts, int streams, spfft_batch* out) {
    if (!out || count <= 0 || !rows || !cols) return fail(SPFFT_ERR_INVALID_ARG, "bad batch arguments");
    *out = nullptr;
    std::unique_ptr<spfft_batch_s> b(new (std::nothrow) spfft_batch_s());
    if (!b) return fail(SPFFT_ERR_ALLOC, "host allocation failed");
    b->count = count;
    b->plans.resize(count);
    b->out_off.resize(count);
    std::vector<double> work(count);
    for (int i = 0; i < count; i++) {
        spfft_opts o;
        spfft_status st0 = read_opts(opts ? &opts[i] : nullptr, &o);
        if (st0 != SPFFT_OK) return fail(st0, "matrix " + std::to_string(i) + ": " + spfft_last_error());
        if (o.u_tile < 0) return fail(SPFFT_ERR_INVALID_ARG, "CSC tile size must be positive");
        spfft_plan_s* p = nullptr;
        spfft_status st = plan_create_ex(rows[i], cols[i], &o, false, &p);
        if (st != SPFFT_OK) return st;
        b->plans[i].reset(p);
        b->out_off[i] = b->out_total;
        b->out_total += (int64_t)p->R * p->P;
        work[i] = work_estimate(*p, nnz_hint ? nnz_hint[i] : 8 * cols[i]);
    }
    b->nslots = std::max(1, std::min(streams > 0 ? streams : 4, count));
    // longest processing time first onto the least loaded stream
    std::vector<int> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return work[x] > work[y]; });
    std::vector<double> load(b->nslots, 0.0);
    b->slot_of.assign(count, 0);
    b->slot_list.assign(b->nslots, {});
    for (int i : order) {
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        b->slot_of[i] = k;
        b->slot_list[k].push_back(i);
        load[k] += work[i];
    }
    b->streams.resize(b->nslots);
    b->join.resize(b->nslots);
    b->ws_T = std::vector<DevBuf>(b->nslots);
    b->ws_X0 = std::vector<DevBuf>(b->nslots);
    for (int k = 0; k < b->nslots; k++) {
        SPFFT_HIP(hipStreamCreateWithFlags(&b->streams[k], hipStreamNonBlocking), "hipStreamCreate");
        SPFFT_HIP(hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming), "hipEventCreate");
        size_t tmax = 0, xmax = 0;
        for (int i : b->slot_list[k]) {
            size_t t, x;
            workspace_bytes(*b->plans[i], &t, &x);
            tmax = std::max(tmax, t);
            xmax = std::max(xmax, x);
        }
        SPFFT_HIP(b->ws_T[k].alloc(tmax), "hipMalloc(batch tile buffer)");
        SPFFT_HIP(b->ws_X0[k].alloc(xmax), "hipMalloc(batch pass-1 buffer)");
    }
    SPFFT_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate");
    *out = b.release();
    return SPFFT_OK;
}

spfft_status spfft_batch_destroy(spfft_batch b) {
    delete b;
    return SPFFT_OK;
}

spfft_status spfft_batch_layout(spfft_batch b, int64_t* total_elems, int64_t* offsets, int* pitches,
                                spfft_plan_info* infos) {
    if (!b) return fail(SPFFT_ERR_INVALID_ARG, "NULL batch");
    // infos: count entries, each with struct_size set by the caller
    for (int i = 0; infos && i < b->count; i++) {
        spfft_status st = check_info(&infos[i]);
        if (st != SPFFT_OK) return st;
    }
    if (total_elems) *total_elems = b->out_total;
    for (int i = 0; i < b->count; i++) {
        if (offsets) offsets[i] = b->out_off[i];
        if (pitches) pitches[i] = b->plans[i]->P;
        if (infos) spfft_plan_get_info(b->plans[i].get(), &infos[i]);
    }
    return SPFFT_OK;
}

// NULL checks and the per-matrix COO ranges: 0 <= coo_off[i + 1] - coo_off[i] <= INT_MAX,
// checked in 64 bits before any use
static spfft_status check_batch_io(spfft_batch b, const int64_t* coo_off, const void* d_out) {
    if (!b) return fail(SPFFT_ERR_INVALID_ARG, "NULL batch");
    if (!coo_off) return fail(SPFFT_ERR_INVALID_ARG, "NULL coo_off");
    if (!d_out) return fail(SPFFT_ERR_INVALID_ARG, "NULL output arena");
    for (int i = 0; i < b->count; i++) {
        const int64_t n = coo_off[i + 1] - coo_off[i];
        if (coo_off[i] < 0 || n < 0 || n > (int64_t)INT_MAX)
            return fail(SPFFT_ERR_INVALID_ARG, "coo_off of matrix " + std::to_string(i) +
                                                   " is not a non-decreasing range of at most INT_MAX pairs");
    }
    return SPFFT_OK;
}

spfft_status spfft_batch_execute(spfft_batch b, const int* d_row, const int* d_col, const int64_t* coo_off,
                                 void* d_out, void* stream) {
    spfft_status st = check_batch_io(b, coo_off, d_out);
    if (st != SPFFT_OK) return st;
    for (int i = 0; i < b->count; i++) {
        st = ensure_csc_capacity(*b->plans[i], (int)(coo_off[i + 1] - coo_off[i]));
        if (st != SPFFT_OK) return st;
    }
    return run_batch(*b, d_row, d_col, coo_off, d_out, (hipStream_t)stream);
}

spfft_status spfft_batch_capture(spfft_batch b, const int* d_row, const int* d_col, const int64_t* coo_off,
                                 void* d_out, spfft_graph* out) {
    if (!out) return fail(SPFFT_ERR_INVALID_ARG, "NULL argument");
    *out = nullptr;
    spfft_status st0 = check_batch_io(b, coo_off, d_out);
    if (st0 != SPFFT_OK) return st0;
    for (int i = 0; i < b->count; i++) {
        spfft_status st = ensure_csc_capacity(*b->plans[i], (int)(coo_off[i + 1] - coo_off[i]));
        if (st != SPFFT_OK) return st;
        b->plans[i]->profiling = false;
    }
    hipStream_t cs;
    SPFFT_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate");
    std::unique_ptr<spfft_graph_s> g(new (std::nothrow) spfft_graph_s());
    spfft_status st = SPFFT_OK;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        st = run_batch(*b, d_row, d_col, coo_off, d_out, cs);
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(cs, &graph);
        g->graph = graph;
        if (e == hipSuccess && st == SPFFT_OK) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    }
    (void)hipStreamDestroy(cs);
    if (st != SPFFT_OK) return st;
    if (e != hipSuccess) return hip_fail(e, "batch graph capture");
    *out = g.release();
    return SPFFT_OK;
}

spfft_status spfft_batch_check_input(spfft_batch b, void* stream) {
    if (!b) return fail(SPFFT_ERR_INVALID_ARG, "NULL batch");
    for (int i = 0; i < b->count; i++) {
        spfft_status st = spfft_check_input(b->plans[i].get(), stream);
        if (st != SPFFT_OK) return fail(st, "matrix " + std::to_string(i) + ": " + spfft_last_error());
    }
    return SPFFT_OK;
}

// Reference-shaped graph variant: host COO in, caller-owned device output out;
// ms = GPU time of the single graph launch (the reference's timed region).
spfft_status spfft_csc_bluestein_mixed_radix_graph(int rows, int cols, int nnz, const int* h_row,
                                                   const int* h_col, int u_tile, spfft_result* res) {
    spfft_status st = check_result(res);
    if (st != SPFFT_OK) return st;
    if (u_tile <= 0) return fail(SPFFT_ERR_INVALID_ARG, "CSC tile size must be positive");
    if (rows <= 0 || cols <= 0 || nnz < 0) return fail(SPFFT_ERR_INVALID_ARG, "bad matrix shape");
    if (nnz > 0 && (!h_row || !h_col)) return fail(SPFFT_ERR_INVALID_ARG, "NULL COO pointer");
    for (int i = 0; i < nnz; i++)
        if (h_row[i] < 0 || h_row[i] >= rows || h_col[i] < 0 || h_col[i] >= cols)
            return fail(SPFFT_ERR_BAD_INPUT, "COO index outside [0, rows) x [0, cols)");
    size_t free0 = 0, total = 0;
    SPFFT_HIP(hipMemGetInfo(&free0, &total), "hipMemGetInfo");
    spfft_opts o;
    spfft_reference_opts(SPFFT_SMOOTH_3, u_tile, &o);
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
    SPFFT_HIP(dout.alloc((size_t)rows * p->P * sizeof(float2)), "hipMalloc(output)");
    spfft_graph g = nullptr;
    st = spfft_graph_capture(p, drow.as<int>(), dcol.as<int>(), nnz, dout.p, p->P, nullptr, &g);
    if (st != SPFFT_OK) return st;
    std::unique_ptr<spfft_graph_s> gg(g);
    hipStream_t s;
    SPFFT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s);
    hipError_t e = hipGraphLaunch(g->exec, s);
    (void)hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    if (e != hipSuccess) return hip_fail(e, "graph launch");
    st = spfft_check_input(p, nullptr);
    if (st != SPFFT_OK) return st;
    size_t free1 = 0;
    SPFFT_HIP(hipMemGetInfo(&free1, &total), "hipMemGetInfo");
    res->d_output = dout.p;
    dout.p = nullptr;
    res->output_cols = p->P;
    res->ms = ms;
    res->ms_e2e = ms;
    res->mem_bytes = free1 < free0 ? free0 - free1 : 0;
    res->fft_len = p->M;
    res->row_transform = SPFFT_TRANSFORM_BLUESTEIN;
    return SPFFT_OK;
}

}  // extern "C"
