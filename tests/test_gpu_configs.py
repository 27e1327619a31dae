"""Full-size GPU runs of the two largest workloads, checked against the FP64 oracle.

c5: the 40421 x 40421 Chung-Lu matrix (~351,400 nonzeros) with its default plan, the
split row transform 40421 = 83 * 487 (the width-487 Bluestein row kernel at M = 1024 over
the 83 runs of each row, then k_splitcols<83>), and with the four-step Bluestein row pass
at M = 82944 (k_fft1<64> + k_conv_fast<36,36> + k_ifft1<64>); spot rows against the FP64
oracle plus the size-independent invariants.

c4: the 512-matrix batch (n log-uniform in [500, 12000]) with the shipped selector's
plans, eager against hipGraph replays bit for bit, and sampled rows of every 16th
matrix against the FP64 oracle.
"""
import numpy as np
import pytest

from test_gpu_path import REL_L2_PLAIN, _dev, _invariants


@pytest.mark.gpu
@pytest.mark.parametrize("transform", ["default", "bluestein"])
def test_c5_full_size(gpu, oracle_mod, transform):
    import torch
    sp = gpu
    m = sp.config_matrix("c5")
    kw = {} if transform == "default" else dict(row_transform=sp.TRANSFORM_BLUESTEIN)
    plan = sp.Plan(m.rows, m.cols, **kw)
    info = plan.info
    print(f"[c5 {transform}] {m.rows}x{m.cols} nnz={m.nnz} M={info['fft_len']} m1={info['m1']} "
          f"m2={info['m2']} t_rows={info['t_rows']} kernel={info['row_kernel']}", flush=True)
    if transform == "default":
        assert info["row_transform"] == sp.TRANSFORM_SPLIT and (info["m1"], info["m2"]) == (83, 487)
    else:
        assert info["fft_len"] == 82944 and info["row_kernel"] == sp.ROW_FOUR_STEP
        assert info["t_rows"] == info["u_tile"]
    assert info["u_tile"] == m.rows // 2 + 1
    out = plan.alloc_output()
    plan.execute(_dev(m.row_idx), _dev(m.col_idx), out)
    torch.cuda.synchronize()
    plan.check_input()
    R = m.rows
    rng = np.random.default_rng(5)
    rows = np.unique(np.concatenate([[0, 1, R // 2, R // 2 + 1, R - 1],
                                     rng.choice(R, 28, replace=False)])).astype(np.int32)
    assert len(rows) >= 32
    ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, u_list=rows, nthreads=8)
    got = out[torch.as_tensor(rows).cuda(), : m.cols // 2 + 1].cpu().numpy()
    err = oracle_mod.rel_l2(got, ref)
    print(f"[c5] spot rel_l2 = {err:.3e} over {len(rows)} rows", flush=True)
    assert err <= REL_L2_PLAIN
    _invariants(out, m)


@pytest.mark.gpu
def test_t_rows_subtiles_bit_identical(gpu):
    """The four-step kernels over sub-tiles of T compute each row exactly as over one
    tile: the results are bit-identical (c5's width, 600 rows)."""
    import torch
    sp = gpu
    m = sp.generate("chunglu", 40421, 9000, 2.2, 0, seed=3)
    m = sp.COOMatrix(600, m.cols, m.row_idx[m.row_idx < 600].copy(), m.col_idx[m.row_idx < 600].copy())
    outs = []
    for t_rows in (0, 301, 7):
        plan = sp.Plan(m.rows, m.cols, t_rows=t_rows, row_transform=sp.TRANSFORM_BLUESTEIN)
        print(f"[t_rows] request {t_rows} -> {plan.info['t_rows']}", flush=True)
        out = plan.alloc_output()
        plan.execute(_dev(m.row_idx), _dev(m.col_idx), out)
        torch.cuda.synchronize()
        outs.append(out[:, : m.cols // 2 + 1].clone())
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])


@pytest.mark.gpu
def test_c4_batch_512_graph_vs_eager(gpu, oracle_mod):
    import torch
    sp = gpu
    from cuspfft_amd import autotune, selector
    mats = autotune.batch_matrices(512)
    assert len(mats) == 512 and all(500 <= m.rows <= 12000 for m in mats)
    opts = [dict(selector.choose_plan(m, fp32_only=True), u_tile=m.rows) for m in mats]
    print(f"[c4] 512 matrices, {sum(m.nnz for m in mats)} nnz, plans chosen", flush=True)
    d_row, d_col, off = sp.Batch.pack(mats)
    b = sp.Batch([(m.rows, m.cols) for m in mats], opts=opts, nnz_hint=[m.nnz for m in mats])
    eager = b.alloc_output()
    b.execute(d_row, d_col, off, eager)
    torch.cuda.synchronize()
    b.check_input()
    g = b.capture(d_row, d_col, off, b.alloc_output())
    replay = g._keep[3]
    for k in range(2):
        replay.zero_()
        g.launch()
        torch.cuda.synchronize()
        # the padding columns of each row are never written: compare the written entries
        for i in range(512):
            V = mats[i].cols // 2 + 1
            assert torch.equal(b.output_of(replay, i)[:, :V], b.output_of(eager, i)[:, :V]), (k, i)
        print(f"[c4] replay {k} matches eager", flush=True)
    worst = 0.0
    for i in range(0, 512, 16):
        m = mats[i]
        rows = np.array(sorted({0, 1, m.rows // 2, m.rows - 1}), np.int32)
        ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, u_list=rows, nthreads=8)
        for arena in (eager, replay):
            X = b.output_of(arena, i)[torch.as_tensor(rows).cuda(), : m.cols // 2 + 1].cpu().numpy()
            worst = max(worst, oracle_mod.rel_l2(X, ref))
    print(f"[c4] worst sampled rel_l2 = {worst:.3e}", flush=True)
    assert worst <= REL_L2_PLAIN
    # the shared descriptor-driven launches against every matrix on its own launches
    mode, members = b.mode()
    print(f"[c4] {members} of 512 matrices in the shared launches", flush=True)
    assert mode == sp.BATCH_AUTO and members > 256
    del g, replay
    b.set_mode(sp.BATCH_PER_MATRIX)
    per = b.alloc_output()
    b.execute(d_row, d_col, off, per)
    torch.cuda.synchronize()
    b.check_input()
    for i in range(512):
        V = mats[i].cols // 2 + 1
        assert torch.equal(b.output_of(per, i)[:, :V], b.output_of(eager, i)[:, :V]), i
