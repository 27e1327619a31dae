"""GPU parity tests of the split row transform (cols = n1 * n2: an on-chip row pass of
width n2 over the n1 runs of each row -- the width-n2 fused plan, or k_raderruns for an
odd prime n2 -- then k_splitcols' length-n1 DFTs across them), through the C ABI,
against the FP64 oracle.

Bar: relative L2 error <= 9e-6 (the plain form's), every output row compared for the
small cases; the streamed output and graph replays bit-identical to the eager run.
"""
import numpy as np
import pytest

from test_gpu_path import REL_L2_PLAIN, _dev, _random_coo, _run

pytestmark = pytest.mark.gpu

# rows, cols, nnz, u_tile, forced (row_transform SPLIT) or AUTO, expected (n1, n2)
SPLIT_CASES = [
    (40, 43243, 4000, 0, False, (83, 521)),    # AUTO: no on-chip Bluestein length; odd prime n1 > 31
    (41, 7217, 3000, 0, True, (7, 1031)),      # register codelet n1, Bluestein sub-pass
    (100, 8198, 5000, 32, True, (2, 4099)),    # even cols, four tiles (the last one partial)
    (64, 3072, 2500, 0, True, (2, 1536)),      # the sub-pass at a direct (smooth) length
    (17, 38147, 2000, 0, False, (37, 1031)),   # odd prime n1 = 37
    (8, 66167, 900, 0, False, (127, 521)),     # n1 = 127, cols > 65536: int32 CSC indices
    (2, 43243, 50, 0, False, (83, 521)),       # rows 0 and 1 only: no mirror rows but row 1
    (1, 43243, 20, 0, False, (83, 521)),       # one row
    # run length 487 (odd prime, 486 = 2 * 243): Rader's convolution over the runs
    (40, 40421, 4000, 0, False, (83, 487)),    # the 40421 x 40421 workload's width
    (100, 40421, 5000, 16, False, (83, 487)),  # tiles of 16 rows, the last one of 3
    (41, 3409, 3000, 0, True, (7, 487)),       # 147 runs: the last workgroup's 3 of 8
    (35, 974, 1500, 0, True, (2, 487)),        # 36 runs
    (1, 974, 30, 0, True, (2, 487)),           # one row, two runs
]


@pytest.mark.parametrize("rows,cols,nnz,tile,forced,n12", SPLIT_CASES)
def test_split_transform(gpu, oracle_mod, rows, cols, nnz, tile, forced, n12):
    import torch
    sp = gpu
    rng = np.random.default_rng(rows * 7 + cols)
    r, c = _random_coo(rng, rows, cols, nnz)
    kw = dict(u_tile=tile)
    if forced:
        kw["row_transform"] = sp.TRANSFORM_SPLIT
    print(f"[split] {rows}x{cols} nnz={len(r)} tile={tile} forced={forced}", flush=True)
    plan, X = _run(sp, rows, cols, r, c, **kw)
    info = plan.info
    assert info["row_transform"] == sp.TRANSFORM_SPLIT
    assert (info["m1"], info["m2"]) == n12
    ref = oracle_mod.sparse_rfft2_f64(rows, cols, r, c, nthreads=8)
    err = oracle_mod.rel_l2(X, ref)
    print(f"  n1={info['m1']} n2={info['m2']} len={info['fft_len']} rel_l2={err:.3e}")
    assert err <= REL_L2_PLAIN
    # streamed output: the same bits
    host = plan.alloc_host_output()
    host.fill_(complex(float("nan"), float("nan")))
    plan.execute_to_host(_dev(r), _dev(c), host)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(host[:, : cols // 2 + 1].numpy(), X)


@pytest.mark.parametrize("c0", [0, 1, 82, 83, 521, 43242])
def test_split_single_nonzero_closed_form(gpu, c0):
    """One nonzero at (r0, c0): X[u][v] = exp(-2 pi i (u r0 / R + v c0 / C)) -- covers the
    first and last entries of the runs (n1 = 0, n2 = 0) and both ends of the row."""
    sp = gpu
    rows, cols, r0 = 12, 43243, 5
    plan, X = _run(sp, rows, cols, np.array([r0], np.int32), np.array([c0], np.int32))
    assert plan.info["row_transform"] == sp.TRANSFORM_SPLIT
    u = np.arange(rows)[:, None]
    v = np.arange(cols // 2 + 1)[None, :]
    ph = ((u * r0) % rows) / rows + ((v * c0) % cols) / cols
    assert np.abs(X - np.exp(-2j * np.pi * ph)).max() < 2e-5


@pytest.mark.parametrize("c0", [0, 1, 82, 83, 83 * 485 + 7, 40420])
def test_split_rader_runs_single_nonzero(gpu, c0):
    """Width 83 * 487 (Rader over the runs): one nonzero at (r0, c0) against the closed
    form; c0 = n1 + 83 n2 covers the runs' x[0] (n2 = 0), which enters each output
    through the inverse codelet, and their last entries."""
    sp = gpu
    rows, cols, r0 = 10, 40421, 3
    plan, X = _run(sp, rows, cols, np.array([r0], np.int32), np.array([c0], np.int32))
    assert plan.info["row_transform"] == sp.TRANSFORM_SPLIT and plan.info["fft_len"] == 486
    u = np.arange(rows)[:, None]
    v = np.arange(cols // 2 + 1)[None, :]
    ph = ((u * r0) % rows) / rows + ((v * c0) % cols) / cols
    assert np.abs(X - np.exp(-2j * np.pi * ph)).max() < 2e-5


def test_split_matches_bluestein_plan(gpu, oracle_mod):
    """A width where both exist: the split plan and the on-chip Bluestein plan agree with
    the oracle, and run to run each gives the same bits."""
    sp = gpu
    rows, cols = 301, 7217
    rng = np.random.default_rng(11)
    r, c = _random_coo(rng, rows, cols, 20000)
    _, Xs = _run(sp, rows, cols, r, c, row_transform=sp.TRANSFORM_SPLIT)
    _, Xs2 = _run(sp, rows, cols, r, c, row_transform=sp.TRANSFORM_SPLIT)
    _, Xb = _run(sp, rows, cols, r, c, row_transform=sp.TRANSFORM_BLUESTEIN)
    np.testing.assert_array_equal(Xs, Xs2)
    ref = oracle_mod.sparse_rfft2_f64(rows, cols, r, c, nthreads=8)
    es, eb = oracle_mod.rel_l2(Xs, ref), oracle_mod.rel_l2(Xb, ref)
    print(f"[split vs bluestein] rel_l2 split={es:.3e} bluestein={eb:.3e}", flush=True)
    assert es <= REL_L2_PLAIN and eb <= REL_L2_PLAIN


def test_batch_graph_with_split_plans(gpu, oracle_mod):
    """A batch with split plans (their own launches on the worker streams, the tile buffer
    from the per-stream workspace) next to on-chip ones: eager against the FP64 oracle,
    two graph replays bit-identical to eager."""
    import torch
    sp = gpu
    sizes = (7217, 1500, 8198, 2310)
    mats = [sp.generate("uniform", n, n * 4, seed=700 + i) for i, n in enumerate(sizes)]
    opts = [dict(row_transform=sp.TRANSFORM_SPLIT if n in (7217, 8198) else sp.TRANSFORM_AUTO)
            for n in sizes]
    b = sp.Batch([(m.rows, m.cols) for m in mats], opts=opts, nnz_hint=[m.nnz for m in mats], streams=2)
    assert [inf["row_transform"] for inf in b.infos][0] == sp.TRANSFORM_SPLIT
    print("[batch] split plans", flush=True)
    d_r, d_c, off = sp.Batch.pack(mats)
    arena = b.alloc_output()
    b.execute(d_r, d_c, off, arena)
    torch.cuda.synchronize()
    b.check_input()
    eager = arena.clone()
    for i, m in enumerate(mats):
        u = np.array([0, 1, 2, m.rows // 3, m.rows // 2, m.rows // 2 + 1, m.rows - 1])
        X = b.output_of(arena, i)[:, : m.cols // 2 + 1].cpu().numpy()[u]
        ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, u_list=u, nthreads=8)
        assert oracle_mod.rel_l2(X, ref) <= REL_L2_PLAIN, i
    g = b.capture(d_r, d_c, off, arena)
    for rep in range(2):
        arena.zero_()
        g.launch()
        torch.cuda.synchronize()
        for i, m in enumerate(mats):
            V = m.cols // 2 + 1
            np.testing.assert_array_equal(b.output_of(arena, i)[:, :V].cpu().numpy(),
                                          b.output_of(eager, i)[:, :V].cpu().numpy())
