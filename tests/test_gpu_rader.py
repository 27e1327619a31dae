"""GPU parity tests of the generated Rader row-kernel shapes (rader_shapes.h): widths C = p
(an odd prime, one convolution of length p - 1 per row) and C = 2p (two), p - 1 7-smooth,
through the C ABI, against the FP64 oracle.

Bar: relative L2 error <= 9e-6 (the plain form's) over every output row; the closed form of
a single nonzero to 2e-5 per entry; graph replays of a batch bit-identical to its eager run.
"""
import importlib.util
import os

import numpy as np
import pytest

from test_gpu_path import REL_L2_PLAIN, _dev, _random_coo, _run

pytestmark = pytest.mark.gpu

_CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cuspfft_amd", "csrc")


def _generated():
    spec = importlib.util.spec_from_file_location("gen_rader_shapes", os.path.join(_CSRC, "gen_rader_shapes.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.shapes()


# (width, (M1, NT, R1, R2, R3, NS)) of every generated shape
GEN = _generated()


@pytest.mark.parametrize("cols,shape", GEN, ids=[f"C{w}" for w, _ in GEN])
def test_rader_generated_shape(gpu, oracle_mod, cols, shape):
    """Every generated shape on a small matrix (rows odd or even by width, one tile and
    the paired column pass where rows are even): all output rows against the oracle."""
    sp = gpu
    m1, nt, r1, r2, r3, ns = shape
    rows = 10 if cols % 4 else 9
    rng = np.random.default_rng(cols)
    r, c = _random_coo(rng, rows, cols, 80)
    print(f"[rader-gen] {rows}x{cols} shape {shape}", flush=True)
    plan, X = _run(sp, rows, cols, r, c, u_tile=rows)
    info = plan.info
    p = cols // ns
    assert info["rader"] == 1 and info["fft_len"] == p - 1 and info["m1"] == m1 and info["m2"] == r1 * r2 * r3
    assert info["row_threads"] == nt
    ref = oracle_mod.sparse_rfft2_f64(rows, cols, r, c, nthreads=8)
    err = oracle_mod.rel_l2(X, ref)
    print(f"  rel_l2={err:.3e}")
    assert err <= REL_L2_PLAIN


# larger cases: square matrices, several tiles, rows = 1 and 2
GEN_BIG = [(1153, 1153, 12000, 1153), (1153, 400, 6000, 64), (12097, 300, 9000, 300), (12097, 1, 40, 1),
           (12097, 2, 60, 2), (8402, 2000, 20000, 512), (2402, 2402, 20000, 2402), (15361, 64, 4000, 16),
           (16202, 128, 5000, 128)]


@pytest.mark.parametrize("cols,rows,nnz,tile", GEN_BIG)
def test_rader_generated_large(gpu, oracle_mod, cols, rows, nnz, tile):
    """Odd prime widths (1153, 12097, 15361) and widths 2p (2402, 8402, 16202) at larger
    sizes: the FP64 oracle on every row (sampled rows above 1000 rows), Rader against the
    Bluestein plan of the same matrix; the streamed output (rows to pinned host memory)
    bit-identical to the device output up to 2402 rows."""
    import torch
    sp = gpu
    assert cols in sp.rader_widths()
    rng = np.random.default_rng(cols + 3 * rows)
    r, c = _random_coo(rng, rows, cols, nnz)
    print(f"[rader-gen-large] {rows}x{cols} nnz={len(r)} tile={tile}", flush=True)
    plan, X = _run(sp, rows, cols, r, c, u_tile=tile)
    ns = 1 if cols % 2 else 2
    assert plan.info["rader"] == 1 and plan.info["fft_len"] == cols // ns - 1
    _, Xb = _run(sp, rows, cols, r, c, u_tile=tile, row_transform=sp.TRANSFORM_BLUESTEIN)
    if rows <= 1000:
        u = np.arange(rows)
    else:
        u = np.unique(np.array([0, 1, 2, 3, rows // 4, rows // 2 - 1, rows // 2, rows // 2 + 1, rows - 2, rows - 1]))
    ref = oracle_mod.sparse_rfft2_f64(rows, cols, r, c, u_list=u.astype(np.int32), nthreads=8)
    err, errb = oracle_mod.rel_l2(X[u], ref), oracle_mod.rel_l2(Xb[u], ref)
    print(f"  rel_l2 rader={err:.3e} bluestein={errb:.3e}")
    assert err <= REL_L2_PLAIN and errb <= REL_L2_PLAIN
    if rows <= 2402:
        host = plan.alloc_host_output()
        host.fill_(complex(float("nan"), float("nan")))
        plan.execute_to_host(_dev(r), _dev(c), host)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(host[:, : cols // 2 + 1].numpy(), X)


@pytest.mark.parametrize("r0,c0", [(0, 0), (1, 1), (5, 1152), (37, 576), (100, 577)])
def test_rader_odd_width_single_nonzero(gpu, r0, c0):
    """One nonzero at (r0, c0) of a 1153 x 1153 matrix: X[u][v] = W_R^{u r0} W_C^{v c0}
    (the positions cover column 0, which the kernel adds to every output, and the
    columns of the convolution's first and last entries)."""
    sp = gpu
    rows = cols = 1153
    plan, X = _run(sp, rows, cols, np.array([r0], np.int32), np.array([c0], np.int32), u_tile=rows)
    assert plan.info["rader"] == 1
    u = np.arange(rows)[:, None]
    v = np.arange(cols // 2 + 1)[None, :]
    ref = np.exp(-2j * np.pi * (((u * r0) % rows) / rows + ((v * c0) % cols) / cols))
    assert np.abs(X - ref).max() < 2e-5


def test_batch_graph_with_generated_rader_widths(gpu, oracle_mod):
    """A batch mixing odd prime widths, generated widths 2p, a direct width and a Bluestein
    one (shared launches: one row-kernel launch per shape): eager against the
    FP64 oracle, two graph replays bit-identical to the eager run, and the per-matrix
    launch mode bit-identical to the shared one."""
    import torch
    sp = gpu
    sizes = (1153, 2402, 1500, 12097, 1024, 1153, 8402)
    mats = [sp.generate("uniform", n, n * (5 if n < 10000 else 1), seed=300 + i) for i, n in enumerate(sizes)]
    b = sp.Batch([(m.rows, m.cols) for m in mats], opts=[dict(u_tile=m.rows) for m in mats],
                 nnz_hint=[m.nnz for m in mats], streams=2)
    assert [inf["rader"] for inf in b.infos] == [1, 1, 0, 1, 0, 1, 1]
    print("[batch] generated rader widths", flush=True)
    d_r, d_c, off = sp.Batch.pack(mats)
    arena = b.alloc_output()
    b.execute(d_r, d_c, off, arena)
    torch.cuda.synchronize()
    b.check_input()
    eager = arena.clone()
    for i, m in enumerate(mats):
        X = b.output_of(arena, i)[:, : m.cols // 2 + 1].cpu().numpy()
        u = np.unique(np.array([0, 1, 2, m.rows // 3, m.rows // 2, m.rows // 2 + 1, m.rows - 1]))
        ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, u_list=u.astype(np.int32),
                                          nthreads=8)
        assert oracle_mod.rel_l2(X[u], ref) <= REL_L2_PLAIN, i
    g = b.capture(d_r, d_c, off, arena)
    for rep in range(2):
        arena.zero_()
        g.launch()
        torch.cuda.synchronize()
        for i, m in enumerate(mats):
            V = m.cols // 2 + 1
            np.testing.assert_array_equal(b.output_of(arena, i)[:, :V].cpu().numpy(),
                                          b.output_of(eager, i)[:, :V].cpu().numpy())
    b.set_mode(sp.BATCH_PER_MATRIX)
    arena.zero_()
    b.execute(d_r, d_c, off, arena)
    torch.cuda.synchronize()
    for i, m in enumerate(mats):
        V = m.cols // 2 + 1
        np.testing.assert_array_equal(b.output_of(arena, i)[:, :V].cpu().numpy(),
                                      b.output_of(eager, i)[:, :V].cpu().numpy())
