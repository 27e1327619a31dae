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
#include "spfft_kernels.h"
#include "spfft_plan_internal.h"

using namespace spfft_internal;

struct spfft_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<DevBuf> keep;  // device buffers the captured work reads (descriptor copies)
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
    // descriptor-driven form (see run_batch_desc)
    struct Group {
        int rep = 0;        // member whose descriptor picks the kernel
        size_t lds = 0;     // dynamic LDS of the launch (column pass: the group's largest)
        int n = 0;          // work items
        DevBuf work;        // int4 (column pass) or int2 (row kernels) items
    };
    struct Step {  // a few members of one lane: their column passes, then their row kernels
        int lane = 0;
        std::vector<Group> colpass, rows;
    };
    struct Chunk {
        std::vector<Step> steps;  // in enqueue order; a lane's steps run in order on its stream
    };
    std::vector<hipStream_t> lane_streams;
    std::vector<hipEvent_t> lane_done;
    hipEvent_t lane_fork = nullptr;
    int mode = 0;                       // SPFFT_BATCH_*
    int req_streams = 4;                // worker streams asked for at creation
    std::vector<double> work;           // per-matrix work estimates (longest first)
    std::vector<int64_t> nnz_hint;      // per-matrix nonzero counts given at creation (else 8 cols)
    std::vector<int64_t> coo_off_h;     // the COO offsets coo_off_d holds
    bool desc_on = false;               // some members run through the descriptors
    std::vector<int> members, fallback; // descriptor members; matrices on their own launches
    std::vector<int> r_off, c_off;      // block-diagonal offsets of the members (-1: not a member)
    std::vector<int64_t> x0_off;        // member's X0 offset in the arena (elements)
    int r_total = 0, c_total = 0;
    int64_t nnz_cap = -1;
    DevBuf mat, merr, coo_off_d, mrow, mcol;
    DevBuf counts, colstart, sparse_cols, col_ptr, totals, err, scan_tmp, hist, hist_off, keys0, keys1, vals0,
        vals1, row_idx, col_rbase, sched, sched_off;
    DevBuf x0_arena, colsched_work;
    int colsched_n = 0;
    std::vector<Chunk> chunks;
    std::vector<spfft::RowPassArgs> h_desc;  // members' descriptors (output pointers patched per call)
    DevBuf d_desc;
    void* desc_out = nullptr;                // the output arena d_desc was written for
    ~spfft_batch_s() {
        for (hipStream_t s : lane_streams) (void)hipStreamDestroy(s);
        for (hipEvent_t e : lane_done) (void)hipEventDestroy(e);
        if (lane_fork) (void)hipEventDestroy(lane_fork);
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

// ---- descriptor-driven form -------------------------------------------------
//
// The members (matrices whose row pass is one on-chip kernel launch: fused, direct or
// Rader, one tile) run as one short launch sequence for the whole batch:
//   1. one CSC build of the block-diagonal matrix of all members (their COO offset by
//      the members' first rows and columns; columns padded to multiples of 64), whose
//      final pass stores each row relative to its own matrix in uint16: every member's
//      CSC is a slice of it (column starts at its first column, rows as they are);
//   2. one column-schedule launch over every member's column tiles;
//   3. per group of members with the same column-pass kernel, one launch whose
//      workgroups take (member, column tile, row block) items from a work list;
//   4. per group of members with the same row kernel shape, one launch of (member,
//      row group) items.
// Each workgroup reads its member's RowPassArgs from a device table (the descriptors).
// Members are ordered longest first, so the big matrices' work is dispatched first.
// Steps 3 and 4 run in chunks whose pass-1 rows fit kBatchX0Budget bytes, each chunk on
// concurrent lane streams forked from (and joined back to) the launch stream.  The other
// matrices (four-step row pass, several tiles, a row limit) keep their own launches on
// worker streams, concurrently.
constexpr size_t kBatchX0Budget = (size_t)24 << 30;
// Members per step of a lane.  A chunk's members are shared out longest first to the
// least loaded of the lane streams (as many as the batch's worker streams); each lane
// runs its members in steps of this many, a step being its column-pass group launches
// and then its row-kernel group launches.  The lanes run concurrently, so column
// passes (VALU-bound table sums) of one lane overlap row kernels of another.  Measured
// on the 512-matrix batch (n 500-12000): whole lanes 134.1 ms, 32 members 132.5 ms,
// 8 members 129.9 ms; one chunk (64 GiB budget) or 8 lanes change it by < 1 ms.
constexpr size_t kBatchStepMembers = 8;

size_t round_up(size_t n, size_t m) { return (n + m - 1) / m * m; }

spfft_status ensure_desc_capacity(spfft_batch_s& b, int64_t nnz) {
    if (nnz <= b.nnz_cap) return SPFFT_OK;
    if (nnz > INT_MAX) return fail(SPFFT_ERR_UNSUPPORTED, "batched CSC build: more than INT_MAX pairs in all");
    const size_t cap = (size_t)std::max<int64_t>(nnz, 1);
    const size_t hist_n = spfft::csc_sort_hist_elems((int)cap);
    const size_t scan_n = std::max(spfft::csc_scan_tmp_elems(b.c_total), spfft::csc_scan_tmp_elems((int)hist_n));
    SPFFT_HIP(b.mrow.alloc(cap * 4), "hipMalloc(batch COO)");
    SPFFT_HIP(b.mcol.alloc(cap * 4), "hipMalloc(batch COO)");
    SPFFT_HIP(b.keys0.alloc(cap * 4), "hipMalloc(batch keys)");
    SPFFT_HIP(b.keys1.alloc(cap * 4), "hipMalloc(batch keys)");
    SPFFT_HIP(b.vals0.alloc(cap * 4), "hipMalloc(batch vals)");
    SPFFT_HIP(b.vals1.alloc(cap * 4), "hipMalloc(batch vals)");
    SPFFT_HIP(b.row_idx.alloc(cap * 2), "hipMalloc(batch row_idx)");
    SPFFT_HIP(b.hist.alloc(hist_n * 4), "hipMalloc(batch hist)");
    SPFFT_HIP(b.hist_off.alloc((hist_n + 1) * 4), "hipMalloc(batch hist)");
    SPFFT_HIP(b.scan_tmp.alloc(scan_n * sizeof(int2)), "hipMalloc(batch scan)");
    b.nnz_cap = (int64_t)cap;
    return SPFFT_OK;
}

// The group launches of one step: its members' column passes grouped by kernel (whole
// W_R table, rows per lane, pairing), their row kernels grouped by shape.
spfft_status build_step(spfft_batch_s& b, const int* ks, int n, spfft_batch_s::Step& step) {
    std::vector<std::pair<std::vector<int>, std::vector<int4>>> cgroups;  // key, items
    std::vector<int> creps;
    std::vector<size_t> clds;
    std::vector<std::pair<std::vector<int>, std::vector<int2>>> rgroups;
    std::vector<int> rreps;
    for (int j = 0; j < n; j++) {
        const int k = ks[j];
        const spfft::RowPassArgs& a = b.h_desc[k];
        const std::vector<int> ckey = {a.wr_full ? 1 : 0, a.colpass_rows, a.pair_half};
        size_t g = 0;
        while (g < cgroups.size() && cgroups[g].first != ckey) g++;
        if (g == cgroups.size()) {
            cgroups.push_back({ckey, {}});
            creps.push_back(k);
            clds.push_back(0);
        }
        clds[g] = std::max(clds[g], spfft::colpass_lds_bytes(a));
        int tx = 0, ty = 0;
        spfft::colpass_grid(a, &tx, &ty);
        for (int y = 0; y < ty; y++)
            for (int x = 0; x < tx; x++) cgroups[g].second.push_back(make_int4(k, x, y, 0));
        const std::vector<int> rkey = {a.rader, a.precision, a.M1, a.fused_nb, a.fused_rpw, a.fused_nt,
                                       a.f2.nstages > 0 ? a.f2.radix[0] : 1, a.f2.nstages > 1 ? a.f2.radix[1] : 1,
                                       a.f2.nstages > 2 ? a.f2.radix[2] : 1, a.direct};
        g = 0;
        while (g < rgroups.size() && rgroups[g].first != rkey) g++;
        if (g == rgroups.size()) {
            rgroups.push_back({rkey, {}});
            rreps.push_back(k);
        }
        const int nblk = a.rader ? a.tile_rows : (a.tile_rows + a.fused_rpw - 1) / a.fused_rpw;
        for (int blk = 0; blk < nblk; blk++) rgroups[g].second.push_back(make_int2(k, blk));
    }
    for (size_t g = 0; g < cgroups.size(); g++) {
        spfft_batch_s::Group grp;
        grp.rep = creps[g];
        grp.lds = clds[g];
        grp.n = (int)cgroups[g].second.size();
        SPFFT_HIP(upload(grp.work, cgroups[g].second), "upload batch work");
        step.colpass.push_back(std::move(grp));
    }
    for (size_t g = 0; g < rgroups.size(); g++) {
        spfft_batch_s::Group grp;
        grp.rep = rreps[g];
        grp.n = (int)rgroups[g].second.size();
        SPFFT_HIP(upload(grp.work, rgroups[g].second), "upload batch work");
        step.rows.push_back(std::move(grp));
    }
    return SPFFT_OK;
}

// The members, their offsets, work lists and fixed buffers (once, at creation).
spfft_status setup_desc(spfft_batch_s& b, const std::vector<double>& work) {
    b.desc_on = false;
    b.members.clear();
    b.fallback.clear();
    bool small = true;  // uint16 rows and columns for every matrix
    for (int i = 0; i < b.count; i++) small = small && b.plans[i]->R <= 65536 && b.plans[i]->C <= 65536;
    std::vector<int> order(b.count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work[x] > work[y]; });
    for (int i : order) {
        const spfft_plan_s& p = *b.plans[i];
        const bool one_tile = p.opts.u_tile >= p.R / 2 + 1 && p.opts.row_limit == 0;
        if (b.mode == SPFFT_BATCH_AUTO && small && p.base.fused && one_tile) b.members.push_back(i);
        else b.fallback.push_back(i);
    }
    b.r_off.assign(b.count, -1);
    b.c_off.assign(b.count, -1);
    b.x0_off.assign(b.count, 0);
    int64_t rt = 0, ct = 0;
    for (int i : b.members) {
        b.r_off[i] = (int)rt;
        b.c_off[i] = (int)ct;
        rt += b.plans[i]->R;
        ct += (int64_t)round_up((size_t)b.plans[i]->C, 64);
    }
    if (b.members.empty() || ct >= (1 << 24) || rt > INT_MAX) {
        // (column keys of the batched sort must stay below 2^24: three 8-bit passes)
        for (int i : b.members) b.fallback.push_back(i);
        b.members.clear();
        return SPFFT_OK;
    }
    b.desc_on = true;
    b.r_total = (int)rt;
    b.c_total = (int)ct;
    // per-matrix {R, C, first row, first column} (not a member: R = -1)
    std::vector<int4> mat(b.count, make_int4(-1, 0, 0, 0));
    for (int i : b.members) mat[i] = make_int4(b.plans[i]->R, b.plans[i]->C, b.r_off[i], b.c_off[i]);
    SPFFT_HIP(upload(b.mat, mat), "upload batch matrices");
    SPFFT_HIP(b.merr.alloc((size_t)b.count * 4), "hipMalloc(batch err)");
    SPFFT_HIP(b.coo_off_d.alloc((size_t)(b.count + 1) * 8), "hipMalloc(batch offsets)");
    const size_t C1 = (size_t)b.c_total + 1;
    SPFFT_HIP(b.counts.alloc(C1 * 4), "hipMalloc(batch counts)");
    SPFFT_HIP(b.colstart.alloc(C1 * 4), "hipMalloc(batch colstart)");
    SPFFT_HIP(b.sparse_cols.alloc(C1 * 4), "hipMalloc(batch sparse_cols)");
    SPFFT_HIP(b.col_ptr.alloc(C1 * 4), "hipMalloc(batch col_ptr)");
    SPFFT_HIP(b.totals.alloc(8), "hipMalloc(batch totals)");
    SPFFT_HIP(b.err.alloc(4), "hipMalloc(batch err)");
    SPFFT_HIP(b.sched.alloc(C1 * 4), "hipMalloc(batch schedule)");
    SPFFT_HIP(b.sched_off.alloc((C1 / 8 + 1) * 9 * 4), "hipMalloc(batch schedule)");
    {
        std::vector<int> rb(b.c_total, 0);
        for (int i : b.members)
            std::fill(rb.begin() + b.c_off[i], rb.begin() + b.c_off[i] + (int)round_up(b.plans[i]->C, 64), b.r_off[i]);
        SPFFT_HIP(upload(b.col_rbase, rb), "upload batch row bases");
    }
    // descriptors (pointers into the shared arrays; X0 and the output set below / per call)
    const int nm = (int)b.members.size();
    b.h_desc.assign(nm, spfft::RowPassArgs{});
    for (int k = 0; k < nm; k++) {
        const int i = b.members[k];
        spfft::RowPassArgs& a = b.h_desc[k];
        batch_descriptor(*b.plans[i], b.nnz_hint[i], a);
        a.colstart = b.colstart.as<int>() + b.c_off[i];
        a.col_sched = b.sched.as<int>() + b.c_off[i];
        a.col_sched_off = b.sched_off.as<int>() + (size_t)(b.c_off[i] / 8) * 9;
        a.row_idx = b.row_idx.p;  // set again once row_idx is allocated
    }
    // chunks of members whose pass-1 rows fit the budget; the X0 arena holds the largest
    size_t arena = 0, in_chunk = 0;
    std::vector<std::vector<int>> chunk_members(1);
    for (int k = 0; k < nm; k++) {
        const spfft::RowPassArgs& a = b.h_desc[k];
        const size_t bytes = (size_t)a.tile_rows * (size_t)a.ldx * sizeof(float2);
        if (in_chunk > 0 && in_chunk + bytes > kBatchX0Budget) {
            chunk_members.emplace_back();
            in_chunk = 0;
        }
        b.x0_off[b.members[k]] = (int64_t)(in_chunk / sizeof(float2));
        in_chunk += round_up(bytes, 256);
        arena = std::max(arena, in_chunk);
        chunk_members.back().push_back(k);
    }
    SPFFT_HIP(b.x0_arena.alloc(arena), "hipMalloc(batch pass-1 rows)");
    for (int k = 0; k < nm; k++) b.h_desc[k].X0 = b.x0_arena.as<float2>() + b.x0_off[b.members[k]];
    // work lists
    {
        std::vector<int2> w;
        for (int k = 0; k < nm; k++) {
            const spfft::RowPassArgs& a = b.h_desc[k];
            const int tc = spfft::colpass_cols_per_tile(a.colpass_rows);
            for (int t = 0; t < (a.C + tc - 1) / tc; t++) w.push_back(make_int2(k, t));
        }
        b.colsched_n = (int)w.size();
        SPFFT_HIP(upload(b.colsched_work, w), "upload batch work");
    }
    // lanes: a chunk's members shared out longest first to the least loaded of the lane
    // streams, which run concurrently (different kernels then share the CUs, as the
    // worker streams of the per-matrix form do)
    const int nlanes = std::max(1, std::min(b.req_streams > 0 ? b.req_streams : 4, 8));
    if ((int)b.lane_streams.size() != nlanes) {
        for (hipStream_t st : b.lane_streams) (void)hipStreamDestroy(st);
        for (hipEvent_t e : b.lane_done) (void)hipEventDestroy(e);
        b.lane_streams.assign(nlanes, nullptr);
        b.lane_done.assign(nlanes, nullptr);
        for (int l = 0; l < nlanes; l++) {
            SPFFT_HIP(hipStreamCreateWithFlags(&b.lane_streams[l], hipStreamNonBlocking), "hipStreamCreate");
            SPFFT_HIP(hipEventCreateWithFlags(&b.lane_done[l], hipEventDisableTiming), "hipEventCreate");
        }
    }
    if (!b.lane_fork) SPFFT_HIP(hipEventCreateWithFlags(&b.lane_fork, hipEventDisableTiming), "hipEventCreate");
    b.chunks.clear();
    b.chunks.resize(chunk_members.size());
    for (size_t c = 0; c < chunk_members.size(); c++) {
        std::vector<std::vector<int>> lane_members(nlanes);
        std::vector<double> lane_load(nlanes, 0.0);
        for (int k : chunk_members[c]) {  // (already longest first)
            const int l = (int)(std::min_element(lane_load.begin(), lane_load.end()) - lane_load.begin());
            lane_members[l].push_back(k);
            lane_load[l] += work[b.members[k]];
        }
        size_t most = 0;
        for (int l = 0; l < nlanes; l++) most = std::max(most, lane_members[l].size());
        for (size_t s0 = 0; s0 < most; s0 += kBatchStepMembers)
            for (int l = 0; l < nlanes; l++) {
                if (s0 >= lane_members[l].size()) continue;
                const size_t s1 = std::min(lane_members[l].size(), s0 + kBatchStepMembers);
                spfft_batch_s::Step step;
                step.lane = l;
                const spfft_status st = build_step(b, lane_members[l].data() + s0, (int)(s1 - s0), step);
                if (st != SPFFT_OK) return st;
                b.chunks[c].steps.push_back(std::move(step));
            }
    }
    SPFFT_HIP(b.d_desc.alloc((size_t)nm * sizeof(spfft::RowPassArgs)), "hipMalloc(batch descriptors)");
    b.desc_out = nullptr;
    return SPFFT_OK;
}

// The members' descriptors for output arena d_out, into d_desc (synchronous upload).
spfft_status write_descriptors(spfft_batch_s& b, void* d_out, DevBuf& d_desc) {
    for (size_t k = 0; k < b.members.size(); k++) {
        spfft::RowPassArgs& a = b.h_desc[k];
        a.row_idx = b.row_idx.p;
        a.out = a.out_m = static_cast<float2*>(d_out) + b.out_off[b.members[k]];
    }
    if (!d_desc.p || d_desc.bytes < b.h_desc.size() * sizeof(spfft::RowPassArgs))
        SPFFT_HIP(d_desc.alloc(b.h_desc.size() * sizeof(spfft::RowPassArgs)), "hipMalloc(batch descriptors)");
    SPFFT_HIP(hipMemcpy(d_desc.p, b.h_desc.data(), b.h_desc.size() * sizeof(spfft::RowPassArgs),
                        hipMemcpyHostToDevice),
              "hipMemcpy(batch descriptors)");
    return SPFFT_OK;
}

// The matrices outside the descriptors: longest processing time first onto the least
// loaded of the worker streams, which share one row-pass workspace per stream.
spfft_status assign_slots(spfft_batch_s& b, const std::vector<double>& work) {
    for (hipStream_t st : b.streams) (void)hipStreamDestroy(st);
    for (hipEvent_t e : b.join) (void)hipEventDestroy(e);
    b.streams.clear();
    b.join.clear();
    b.nslots = std::max(1, std::min(b.req_streams > 0 ? b.req_streams : 4, std::max((int)b.fallback.size(), 1)));
    std::vector<int> order = b.fallback;
    std::sort(order.begin(), order.end(), [&](int x, int y) { return work[x] > work[y]; });
    std::vector<double> load(b.nslots, 0.0);
    b.slot_of.assign(b.count, 0);
    b.slot_list.assign(b.nslots, {});
    for (int i : order) {
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        b.slot_of[i] = k;
        b.slot_list[k].push_back(i);
        load[k] += work[i];
    }
    b.streams.resize(b.nslots);
    b.join.resize(b.nslots);
    b.ws_T = std::vector<DevBuf>(b.nslots);
    b.ws_X0 = std::vector<DevBuf>(b.nslots);
    for (int k = 0; k < b.nslots; k++) {
        SPFFT_HIP(hipStreamCreateWithFlags(&b.streams[k], hipStreamNonBlocking), "hipStreamCreate");
        SPFFT_HIP(hipEventCreateWithFlags(&b.join[k], hipEventDisableTiming), "hipEventCreate");
        size_t tmax = 0, xmax = 0;
        for (int i : b.slot_list[k]) {
            size_t t, x;
            workspace_bytes(*b.plans[i], &t, &x);
            tmax = std::max(tmax, t);
            xmax = std::max(xmax, x);
        }
        SPFFT_HIP(b.ws_T[k].alloc(tmax), "hipMalloc(batch tile buffer)");
        SPFFT_HIP(b.ws_X0[k].alloc(xmax), "hipMalloc(batch pass-1 buffer)");
    }
    return SPFFT_OK;
}

// The members' work on stream s (descriptors d_desc, COO offsets d_coo_off on the device).
spfft_status run_batch_desc(spfft_batch_s& b, const int* d_row, const int* d_col, const int64_t* d_coo_off,
                            int64_t nnz_total, const spfft::RowPassArgs* d_desc, hipStream_t s) {
    SPFFT_HIP(spfft::launch_fill_int(b.merr.as<int>(), b.count, 0, s), "batch launch");
    SPFFT_HIP(spfft::launch_batch_coo(d_row, d_col, reinterpret_cast<const long long*>(d_coo_off), b.count,
                                      b.mat.as<int4>(), b.mrow.as<int>(), b.mcol.as<int>(), b.merr.as<int>(),
                                      (long long)nnz_total, s),
              "batch launch");
    spfft::CscBuildArgs c{};
    c.R = b.r_total;
    c.C = b.c_total;
    c.nnz = (int)nnz_total;
    c.packed = false;
    c.row = b.mrow.as<int>();
    c.col = b.mcol.as<int>();
    c.counts = b.counts.as<int>();
    c.colstart = b.colstart.as<int>();
    c.sparse_cols = b.sparse_cols.p;
    c.col_ptr = b.col_ptr.as<int>();
    c.row_idx = b.row_idx.p;
    c.totals = b.totals.as<int>();
    c.err = b.err.as<int>();
    c.scan_tmp = b.scan_tmp.as<int2>();
    c.hist = b.hist.as<int>();
    c.hist_off = b.hist_off.as<int>();
    c.keys[0] = b.keys0.as<int>();
    c.keys[1] = b.keys1.as<int>();
    c.vals[0] = b.vals0.as<int>();
    c.vals[1] = b.vals1.as<int>();
    c.col_rbase = b.col_rbase.as<int>();
    SPFFT_HIP(spfft::launch_csc_build(c, s), "batch CSC build");
    SPFFT_HIP(spfft::launch_colsched_batch(d_desc, b.colsched_work.as<int2>(), b.colsched_n, s), "batch launch");
    for (const auto& ch : b.chunks) {
        SPFFT_HIP(hipEventRecord(b.lane_fork, s), "hipEventRecord");
        for (hipStream_t ls : b.lane_streams)
            SPFFT_HIP(hipStreamWaitEvent(ls, b.lane_fork, 0), "hipStreamWaitEvent");
        for (const auto& st : ch.steps) {
            hipStream_t ls = b.lane_streams[st.lane];
            for (const auto& g : st.colpass)
                SPFFT_HIP(spfft::launch_colpass_batch(b.h_desc[g.rep], true, d_desc, g.work.as<int4>(), g.n, g.lds,
                                                      ls),
                          "batch column pass");
            for (const auto& g : st.rows) {
                const spfft::RowPassArgs& rep = b.h_desc[g.rep];
                SPFFT_HIP(rep.rader ? spfft::launch_rader_batch(rep, d_desc, g.work.as<int2>(), g.n, ls)
                                    : spfft::launch_fused_batch(rep, d_desc, g.work.as<int2>(), g.n, ls),
                          "batch row pass");
            }
        }
        for (size_t l = 0; l < b.lane_streams.size(); l++) {
            SPFFT_HIP(hipEventRecord(b.lane_done[l], b.lane_streams[l]), "hipEventRecord");
            SPFFT_HIP(hipStreamWaitEvent(s, b.lane_done[l], 0), "hipStreamWaitEvent");
        }
    }
    return SPFFT_OK;
}

// The whole batch on stream s: the matrices outside the descriptors on the worker streams
// (forked from s), the members on s, then s waits for the workers.
spfft_status run_batch(spfft_batch_s& b, const int* d_row, const int* d_col, const int64_t* coo_off,
                       void* d_out, hipStream_t s, const spfft::RowPassArgs* d_desc, const int64_t* d_coo_off) {
    if (!coo_off) return fail(SPFFT_ERR_INVALID_ARG, "NULL COO offsets");
    if (!d_out) return fail(SPFFT_ERR_INVALID_ARG, "NULL output arena");
    const bool workers = !b.fallback.empty();
    if (workers) {
        SPFFT_HIP(hipEventRecord(b.fork, s), "hipEventRecord");
        for (int k = 0; k < b.nslots; k++)
            SPFFT_HIP(hipStreamWaitEvent(b.streams[k], b.fork, 0), "hipStreamWaitEvent");
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
        }
    }
    if (b.desc_on) {
        spfft_status st = run_batch_desc(b, d_row, d_col, d_coo_off, coo_off[b.count] - coo_off[0], d_desc, s);
        if (st != SPFFT_OK) return st;
    }
    if (workers)
        for (int k = 0; k < b.nslots; k++) SPFFT_HIP(hipStreamWaitEvent(s, b.join[k], 0), "hipStreamWaitEvent");
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
    b->nnz_hint.resize(count);
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
        b->nnz_hint[i] = nnz_hint ? nnz_hint[i] : 8 * (int64_t)cols[i];
    }
    b->mode = SPFFT_BATCH_AUTO;
    b->req_streams = streams;
    b->work = work;
    spfft_status st_d = setup_desc(*b, work);
    if (st_d != SPFFT_OK) return st_d;
    if ((st_d = assign_slots(*b, work)) != SPFFT_OK) return st_d;
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

// Device copies of the members' descriptors (for output arena d_out) and of the COO
// offsets, for the eager form: rewritten (after the stream has drained, since queued work
// may still read them) only when they change.
static spfft_status prepare_eager(spfft_batch b, const int64_t* coo_off, void* d_out, hipStream_t s) {
    if (!b->desc_on) return SPFFT_OK;
    const bool new_out = d_out != b->desc_out;
    const bool new_off = b->coo_off_h.size() != (size_t)b->count + 1 ||
                         !std::equal(b->coo_off_h.begin(), b->coo_off_h.end(), coo_off);
    if (!new_out && !new_off) return SPFFT_OK;
    SPFFT_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (new_out) {
        spfft_status st = write_descriptors(*b, d_out, b->d_desc);
        if (st != SPFFT_OK) return st;
        b->desc_out = d_out;
    }
    if (new_off) {
        b->coo_off_h.assign(coo_off, coo_off + b->count + 1);
        SPFFT_HIP(hipMemcpy(b->coo_off_d.p, coo_off, (size_t)(b->count + 1) * 8, hipMemcpyHostToDevice),
                  "hipMemcpy(batch offsets)");
    }
    return SPFFT_OK;
}

static spfft_status ensure_capacity_all(spfft_batch b, const int64_t* coo_off) {
    for (int i : b->fallback) {
        spfft_status st = ensure_csc_capacity(*b->plans[i], (int)(coo_off[i + 1] - coo_off[i]));
        if (st != SPFFT_OK) return st;
    }
    if (b->desc_on) {
        const bool had = b->row_idx.p != nullptr;
        spfft_status st = ensure_desc_capacity(*b, coo_off[b->count] - coo_off[0]);
        if (st != SPFFT_OK) return st;
        if (had != (b->row_idx.p != nullptr) || b->h_desc.empty() || b->h_desc[0].row_idx != b->row_idx.p)
            b->desc_out = nullptr;  // row_idx moved: the descriptors are rewritten
    }
    return SPFFT_OK;
}

spfft_status spfft_batch_execute(spfft_batch b, const int* d_row, const int* d_col, const int64_t* coo_off,
                                 void* d_out, void* stream) {
    spfft_status st = check_batch_io(b, coo_off, d_out);
    if (st != SPFFT_OK) return st;
    if ((st = ensure_capacity_all(b, coo_off)) != SPFFT_OK) return st;
    if ((st = prepare_eager(b, coo_off, d_out, (hipStream_t)stream)) != SPFFT_OK) return st;
    return run_batch(*b, d_row, d_col, coo_off, d_out, (hipStream_t)stream, b->d_desc.as<spfft::RowPassArgs>(),
                     b->coo_off_d.as<int64_t>());
}

spfft_status spfft_batch_capture(spfft_batch b, const int* d_row, const int* d_col, const int64_t* coo_off,
                                 void* d_out, spfft_graph* out) {
    if (!out) return fail(SPFFT_ERR_INVALID_ARG, "NULL argument");
    *out = nullptr;
    spfft_status st0 = check_batch_io(b, coo_off, d_out);
    if (st0 != SPFFT_OK) return st0;
    if ((st0 = ensure_capacity_all(b, coo_off)) != SPFFT_OK) return st0;
    for (int i = 0; i < b->count; i++) b->plans[i]->profiling = false;
    std::unique_ptr<spfft_graph_s> g(new (std::nothrow) spfft_graph_s());
    if (!g) return fail(SPFFT_ERR_ALLOC, "host allocation failed");
    // the graph's own copies of the descriptors and offsets (later eager calls may
    // rewrite the batch's)
    const spfft::RowPassArgs* d_desc = nullptr;
    const int64_t* d_off = nullptr;
    if (b->desc_on) {
        g->keep.resize(2);
        spfft_status st = write_descriptors(*b, d_out, g->keep[0]);
        if (st != SPFFT_OK) return st;
        b->desc_out = nullptr;  // h_desc now holds this arena's pointers; eager calls rewrite theirs
        SPFFT_HIP(g->keep[1].alloc((size_t)(b->count + 1) * 8), "hipMalloc(graph offsets)");
        SPFFT_HIP(hipMemcpy(g->keep[1].p, coo_off, (size_t)(b->count + 1) * 8, hipMemcpyHostToDevice),
                  "hipMemcpy(graph offsets)");
        d_desc = g->keep[0].as<spfft::RowPassArgs>();
        d_off = g->keep[1].as<int64_t>();
    }
    hipStream_t cs;
    SPFFT_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate");
    spfft_status st = SPFFT_OK;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        st = run_batch(*b, d_row, d_col, coo_off, d_out, cs, d_desc, d_off);
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
    if (b->desc_on) {
        SPFFT_HIP(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
        std::vector<int> err(b->count, 0);
        SPFFT_HIP(hipMemcpy(err.data(), b->merr.p, (size_t)b->count * 4, hipMemcpyDeviceToHost), "hipMemcpy(err)");
        for (int i : b->members)
            if (err[i])
                return fail(SPFFT_ERR_BAD_INPUT, "matrix " + std::to_string(i) + ": COO index outside [0, rows) x [0, cols)");
    }
    for (int i : b->fallback) {
        spfft_status st = spfft_check_input(b->plans[i].get(), stream);
        if (st != SPFFT_OK) return fail(st, "matrix " + std::to_string(i) + ": " + spfft_last_error());
    }
    return SPFFT_OK;
}

spfft_status spfft_batch_set_mode(spfft_batch b, int mode) {
    if (!b) return fail(SPFFT_ERR_INVALID_ARG, "NULL batch");
    if (mode != SPFFT_BATCH_AUTO && mode != SPFFT_BATCH_PER_MATRIX)
        return fail(SPFFT_ERR_INVALID_ARG, "batch mode must be SPFFT_BATCH_AUTO or SPFFT_BATCH_PER_MATRIX");
    if (mode == b->mode) return SPFFT_OK;
    b->mode = mode;
    std::vector<double> work(b->count);
    for (int i = 0; i < b->count; i++) work[i] = b->work[i];
    spfft_status st = setup_desc(*b, work);
    if (st != SPFFT_OK) return st;
    return assign_slots(*b, work);
}

spfft_status spfft_batch_get_mode(spfft_batch b, int* mode, int* members) {
    if (!b) return fail(SPFFT_ERR_INVALID_ARG, "NULL batch");
    if (mode) *mode = b->mode;
    if (members) *members = (int)b->members.size();
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
