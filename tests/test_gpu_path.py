"""GPU parity tests of the hot path, through the C ABI (cuspfft_amd -> libspfft_hip.so).

Bars:
  * CSC build: bit-exact against the oracle's restatement of make_compact_csc
    (reference include/sparse_fft_internal.cuh:200-222) and its uint16 packing.
  * Plain FP32 transform: relative L2 error <= 9e-6 against the FP64 oracle.
    FFTW single precision is ~1e-7 from exact, so by the triangle inequality
    this keeps the plain form within 1e-5 of FFTW's output.
  * Full-size configurations: spot rows against the FP64 oracle plus
    size-independent invariants (X[0][0] = nnz, Parseval, X[u][v] = X[v][u]
    for symmetric inputs).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL_L2_PLAIN = 9e-6


def _dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, np.int32)).cuda()


def _run(sp, rows, cols, r, c, **kw):
    import torch
    plan = sp.Plan(rows, cols, **kw)
    out = plan.alloc_output()
    plan.execute(_dev(r), _dev(c), out)
    torch.cuda.synchronize()
    plan.check_input()
    return plan, out[:, : cols // 2 + 1].cpu().numpy()


def _random_coo(rng, rows, cols, nnz, shuffle=True):
    nnz = min(nnz, rows * cols)
    idx = rng.choice(rows * cols, nnz, replace=False)
    if not shuffle:
        idx = np.sort(idx)
    return (idx // cols).astype(np.int32), (idx % cols).astype(np.int32)


def _csc_case(sp, oracle_mod, rows, cols, r, c):
    import torch
    plan = sp.Plan(rows, cols)
    plan.csc_build(_dev(r), _dev(c))
    got = plan.csc_get()
    sc, cp, ri = oracle_mod.compact_csc(rows, cols, r, c)
    assert got["packed"] == oracle_mod.can_pack_u16(rows, cols)
    assert got["n_sparse_cols"] == len(sc)
    np.testing.assert_array_equal(got["sparse_cols"].astype(np.int64), sc)
    np.testing.assert_array_equal(got["col_ptr"], cp)
    np.testing.assert_array_equal(got["row_idx"].astype(np.int64), ri)
    counts = np.bincount(c, minlength=cols)
    np.testing.assert_array_equal(got["col_start"], np.concatenate([[0], np.cumsum(counts)]))
    del plan
    torch.cuda.synchronize()


@pytest.mark.parametrize("case", ["c1", "c1_shuffled", "c3", "small_cols", "empty_cols",
                                  "full_column", "single", "edge_65536", "int32_path"])
def test_csc_bit_exact(gpu, oracle_mod, case):
    sp = gpu
    rng = np.random.default_rng(7)
    print(f"[csc] case={case}", flush=True)
    if case in ("c1", "c1_shuffled", "c3"):
        m = sp.config_matrix("c3" if case == "c3" else "c1")
        r, c = m.row_idx, m.col_idx
        if case == "c1_shuffled":
            p = rng.permutation(len(r))
            r, c = r[p], c[p]
        _csc_case(sp, oracle_mod, m.rows, m.cols, r, c)
    elif case == "small_cols":
        r, c = _random_coo(rng, 3000, 200, 20000)
        _csc_case(sp, oracle_mod, 3000, 200, r, c)
    elif case == "empty_cols":
        r, c = _random_coo(rng, 100, 1000, 300)
        c = (c // 7) * 7  # only every 7th column can be active
        key = np.unique(r.astype(np.int64) * 1000 + c)
        r, c = (key // 1000).astype(np.int32), (key % 1000).astype(np.int32)
        p = rng.permutation(len(r))
        _csc_case(sp, oracle_mod, 100, 1000, r[p], c[p])
    elif case == "full_column":
        rows, cols = 5000, 40
        r = np.arange(rows, dtype=np.int32)[::-1].copy()
        c = np.full(rows, 17, np.int32)
        r2, c2 = _random_coo(rng, rows, cols, 3000)
        keep = c2 != 17
        r = np.concatenate([r, r2[keep]])
        c = np.concatenate([c, c2[keep]])
        p = rng.permutation(len(r))
        _csc_case(sp, oracle_mod, rows, cols, r[p], c[p])
    elif case == "single":
        _csc_case(sp, oracle_mod, 9, 11, np.array([4], np.int32), np.array([10], np.int32))
    elif case == "edge_65536":
        rows, cols = 8, 65536
        r, c = _random_coo(rng, rows, cols, 5000)
        c[:3] = [0, 65535, 32768]
        key = np.unique(r.astype(np.int64) * cols + c)
        r, c = (key // cols).astype(np.int32), (key % cols).astype(np.int32)
        _csc_case(sp, oracle_mod, rows, cols, r, c)
    elif case == "int32_path":
        rows, cols = 6, 70000
        r, c = _random_coo(rng, rows, cols, 4000)
        _csc_case(sp, oracle_mod, rows, cols, r, c)


ROW_KERNELS = pytest.mark.parametrize("row_kernel", [0, 1], ids=["auto", "four_step"])


@ROW_KERNELS
@pytest.mark.parametrize("rule", [3, 7])
def test_c1_full_vs_fp64(gpu, oracle_mod, rule, row_kernel):
    sp = gpu
    m = sp.config_matrix("c1")
    print(f"[fft] c1 rule={rule} row_kernel={row_kernel}", flush=True)
    plan, X = _run(sp, m.rows, m.cols, m.row_idx, m.col_idx, length_rule=rule, row_kernel=row_kernel)
    assert plan.info["row_kernel"] == (row_kernel or 2)
    ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, nthreads=8)
    err = oracle_mod.rel_l2(X, ref)
    print(f"c1 rule={rule} M={plan.info['fft_len']} rel_l2={err:.3e}")
    assert err <= REL_L2_PLAIN
    assert abs(X[0, 0] - m.nnz) <= 1e-6 * m.nnz * 10


SHAPES = [(1, 1, 1), (1, 7, 3), (7, 1, 4), (3, 5, 6), (17, 31, 100), (64, 64, 700),
          (127, 128, 2000), (130, 257, 3000), (200, 3, 250), (33, 1000, 5000)]


@pytest.mark.parametrize("rows,cols,nnz", SHAPES)
@pytest.mark.parametrize("tile", [1, 7, 128])
def test_small_shapes(gpu, oracle_mod, rows, cols, nnz, tile):
    sp = gpu
    rng = np.random.default_rng(rows * 1000 + cols)
    r, c = _random_coo(rng, rows, cols, nnz)
    print(f"[fft] shape {rows}x{cols} nnz={len(r)} tile={tile}", flush=True)
    for rule in (3, 7):
        _, X = _run(sp, rows, cols, r, c, length_rule=rule, u_tile=tile)
        ref = oracle_mod.sparse_rfft2_f64(rows, cols, r, c)
        assert oracle_mod.rel_l2(X, ref) <= REL_L2_PLAIN, (rows, cols, rule)


def test_empty_matrix(gpu):
    sp = gpu
    z = np.zeros(0, np.int32)
    _, X = _run(sp, 50, 60, z, z)
    assert np.all(X == 0)


@pytest.mark.parametrize("rows,cols,r0,c0", [(2310, 2310, 17, 1999), (31, 8219, 30, 8218),
                                             (1000, 999, 0, 0)])
def test_single_nonzero_closed_form(gpu, rows, cols, r0, c0):
    sp = gpu
    _, X = _run(sp, rows, cols, np.array([r0], np.int32), np.array([c0], np.int32))
    u = np.arange(rows)[:, None]
    v = np.arange(cols // 2 + 1)[None, :]
    ph = ((u * r0) % rows) / rows + ((v * c0) % cols) / cols
    ref = np.exp(-2j * np.pi * ph)
    assert np.abs(X - ref).max() < 2e-5


def _invariants(X_dev, m):
    """X[0][0] = nnz, half-spectrum Parseval, and symmetry X[u][v] = X[v][u]."""
    import torch
    V = m.cols // 2 + 1
    Xv = X_dev[:, :V]
    assert abs(complex(Xv[0, 0].item()) - m.nnz) <= 1e-5 * m.nnz
    w = torch.full((V,), 2.0, dtype=torch.float64, device=Xv.device)
    w[0] = 1.0
    if m.cols % 2 == 0:
        w[-1] = 1.0
    pars = (Xv.abs().double() ** 2 * w[None, :]).sum().item()
    expect = float(m.rows) * m.cols * m.nnz
    assert abs(pars / expect - 1) < 1e-5, pars / expect
    n = min(V, m.rows)
    S = Xv[:n, :n]
    sym = ((S - S.transpose(0, 1)).abs().double().norm() / S.abs().double().norm()).item()
    assert sym < 1e-5, sym


@ROW_KERNELS
@pytest.mark.parametrize("cfg,rule", [("c2", 3), ("c2", 7), ("c3", 3), ("c3", 7)])
def test_full_config_spot_rows_and_invariants(gpu, oracle_mod, cfg, rule, row_kernel):
    import torch
    sp = gpu
    m = sp.config_matrix(cfg)
    print(f"[fft] {cfg} rule={rule} row_kernel={row_kernel}", flush=True)
    plan = sp.Plan(m.rows, m.cols, length_rule=rule, u_tile=256, row_kernel=row_kernel)
    assert plan.info["row_kernel"] == (row_kernel or 2)
    out = plan.alloc_output()
    plan.execute(_dev(m.row_idx), _dev(m.col_idx), out)
    torch.cuda.synchronize()
    plan.check_input()
    rng = np.random.default_rng(3)
    rows = np.unique(np.concatenate([[0, 1, m.rows - 1], rng.choice(m.rows, 29, replace=False)]))
    ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, u_list=rows, nthreads=8)
    got = out[torch.as_tensor(rows).cuda(), : m.cols // 2 + 1].cpu().numpy()
    err = oracle_mod.rel_l2(got, ref)
    print(f"{cfg} rule={rule} M={plan.info['fft_len']} spot rel_l2={err:.3e}")
    assert err <= REL_L2_PLAIN
    _invariants(out, m)


def test_tile_invariance(gpu, oracle_mod):
    """Any tiling computes the same transform: identical for repeated runs of one
    plan, and within FP32 rounding across tile sizes (the column pass derives the
    twiddle of a lane's second row by one complex product from its first, and the
    rows that are first or second in a tile depend on the tiling)."""
    sp = gpu
    m = sp.config_matrix("c1")
    outs = []
    for tile in (1, 100, 2310, 5000):
        plan, X = _run(sp, m.rows, m.cols, m.row_idx, m.col_idx, u_tile=tile)
        outs.append(X)
        _, X2 = _run(sp, m.rows, m.cols, m.row_idx, m.col_idx, u_tile=tile)
        np.testing.assert_array_equal(X, X2)
    for X in outs[1:]:
        assert oracle_mod.rel_l2(X, outs[0]) <= 1e-6


def test_errors(gpu):
    sp = gpu
    m = sp.config_matrix("c1")
    with pytest.raises(RuntimeError, match="CSC tile size must be positive"):
        sp.sparse_fft_csc_bluestein_mixed_radix(m, 0)
    with pytest.raises(sp.SpfftError) as ei:
        sp.Plan(10, 10, u_tile=-3)
    assert ei.value.status == 1
    # out-of-range rows and columns (too large and negative): the kernels skip
    # them in the column counts and store invalid rows as 0, so every access
    # stays in bounds; check_input reports the bad input afterwards
    for bad_r, bad_c in (([0, 5, 10], [0, 1, 2]), ([-7, 3, 2, 1], [4, 1, 70000, -1])):
        plan = sp.Plan(10, 10)
        out = plan.alloc_output()
        plan.execute(_dev(np.array(bad_r, np.int32)), _dev(np.array(bad_c, np.int32)), out)
        with pytest.raises(sp.SpfftError) as ei:
            plan.check_input()
        assert ei.value.status == 5
    # a valid transform on the same plan afterwards is unaffected
    plan.execute(_dev(np.array([1, 2], np.int32)), _dev(np.array([3, 4], np.int32)), out)
    plan.check_input()
    got = out[:, :6].cpu().numpy()
    u, v = np.arange(10)[:, None], np.arange(6)[None, :]
    want = np.exp(-2j * np.pi * (u * 1 + v * 3) / 10) + np.exp(-2j * np.pi * (u * 2 + v * 4) / 10)
    np.testing.assert_allclose(got, want, atol=2e-6)


def test_reference_shaped_entry_points(gpu, oracle_mod):
    """The one-shot C entry points: host COO in, caller-owned device output out."""
    sp = gpu
    L = sp.load()
    m = sp.config_matrix("c1")
    ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, nthreads=8)
    for fn in (L.spfft_csc_bluestein_mixed_radix, L.spfft_csc_bluestein_cufft_smooth):
        res = sp.Result()
        st = fn(m.rows, m.cols, m.nnz, m.row_idx.ctypes.data, m.col_idx.ctypes.data, 128,
                ctypes.byref(res))
        assert st == 0, L.spfft_last_error()
        assert res.output_cols == (m.cols // 2 + 1 + 3) // 4 * 4
        assert res.ms > 0 and res.ms_e2e >= res.ms
        X = sp.device_ptr_to_host(res.d_output, m.rows, res.output_cols)[:, : m.cols // 2 + 1]
        assert L.spfft_free_output(res.d_output) == 0
        assert oracle_mod.rel_l2(X, ref) <= REL_L2_PLAIN
    res = sp.Result()
    st = L.spfft_csc_bluestein_mixed_radix(m.rows, m.cols, m.nnz, m.row_idx.ctypes.data,
                                           m.col_idx.ctypes.data, 0, ctypes.byref(res))
    assert st == 1 and b"CSC tile size must be positive" in L.spfft_last_error()


@pytest.mark.parametrize("cols", [16706, 16384])
def test_reference_entry_points_run_bluestein(gpu, oracle_mod, cols):
    """On widths where the native plan takes the Rader (c3's 16706) or direct (16384)
    row transform, the reference-shaped entry points still run the reference's
    algorithm, Bluestein at next_3_smooth / next_7_smooth(2 cols - 1), as reported in
    spfft_result, and match the FP64 oracle."""
    sp = gpu
    L = sp.load()
    m = sp.generate("uniform", cols, 2500, seed=11)
    rows = np.array([0, 1, 2, cols // 2, cols - 1], np.int32)
    ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, u_list=rows, nthreads=8)
    for fn, M in ((L.spfft_csc_bluestein_mixed_radix, sp.next_3_smooth(2 * cols - 1)),
                  (L.spfft_csc_bluestein_cufft_smooth, sp.next_7_smooth(2 * cols - 1)),
                  (L.spfft_csc_bluestein_mixed_radix_graph, sp.next_3_smooth(2 * cols - 1))):
        print(f"[ref-entry] cols={cols} {fn.__name__}", flush=True)
        res = sp.Result()
        st = fn(m.rows, m.cols, m.nnz, m.row_idx.ctypes.data, m.col_idx.ctypes.data, 128, ctypes.byref(res))
        assert st == 0, L.spfft_last_error()
        assert res.fft_len == M and res.row_transform == sp.TRANSFORM_BLUESTEIN
        X = sp.device_ptr_to_host(res.d_output, m.rows, res.output_cols)[rows, : m.cols // 2 + 1]
        assert L.spfft_free_output(res.d_output) == 0
        assert oracle_mod.rel_l2(X, ref) <= REL_L2_PLAIN
    # the native API on the same width: the Rader / direct transform
    res = sp.Result()
    o = sp._opts()
    st = L.spfft_run_host_coo(m.rows, m.cols, m.nnz, m.row_idx.ctypes.data, m.col_idx.ctypes.data,
                              ctypes.byref(o), ctypes.byref(res))
    assert st == 0, L.spfft_last_error()
    want = sp.TRANSFORM_RADER if cols == 16706 else sp.TRANSFORM_DIRECT
    assert res.row_transform == want
    X = sp.device_ptr_to_host(res.d_output, m.rows, res.output_cols)[rows, : m.cols // 2 + 1]
    assert L.spfft_free_output(res.d_output) == 0
    assert oracle_mod.rel_l2(X, ref) <= REL_L2_PLAIN


def test_python_mirror(gpu, oracle_mod):
    sp = gpu
    m = sp.config_matrix("c1")
    r = sp.sparse_fft_csc_bluestein_cufft_smooth(m, 64)
    X = r.d_output[:, : m.cols // 2 + 1].cpu().numpy()
    ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, nthreads=8)
    assert r.output_cols == (m.cols // 2 + 1 + 3) // 4 * 4
    assert oracle_mod.rel_l2(X, ref) <= REL_L2_PLAIN


def _split_cases():
    """(cols, fft_len, m1) combinations covering many register codelets (k_fft1/k_ifft1),
    fast convolution pairs and the generic convolution fallback."""
    cases = []
    for m1 in (2, 3, 4, 6, 8, 9, 12, 16, 18, 24, 27, 32, 36, 48, 54, 64):
        cases.append((2310, 5184, m1))
    cases += [(20, 64, 1), (3, 5, 1)]  # M1 = 1: the whole row in the convolution kernel
    for m1 in (5, 7, 10, 14, 15, 20, 21, 25, 28, 30, 35, 40, 42, 45, 49, 56, 60, 63):
        cases.append((2310, 0, m1))  # 7-smooth M chosen below per m1
    cases += [(300, 686, 0), (300, 625, 0), (1000, 2401, 0), (100, 3125, 0), (5000, 10125, 0)]
    return cases


@pytest.mark.parametrize("cols,fft_len,m1", _split_cases())
def test_m1_splits(gpu, oracle_mod, cols, fft_len, m1):
    sp = gpu
    if fft_len == 0:
        # smallest 7-smooth M >= 2*cols-1 divisible by m1
        M = 2 * cols - 1
        while M % m1 or sp.next_7_smooth(M) != M:
            M += 1
        fft_len = M
    rows = 97
    rng = np.random.default_rng(cols + fft_len + m1)
    r, c = _random_coo(rng, rows, cols, 3000)
    print(f"[split] cols={cols} M={fft_len} m1={m1}", flush=True)
    plan, X = _run(sp, rows, cols, r, c, fft_len=fft_len, m1=m1, u_tile=50, row_kernel=1)
    assert plan.info["fft_len"] == fft_len
    if m1:
        assert plan.info["m1"] == m1
    ref = oracle_mod.sparse_rfft2_f64(rows, cols, r, c)
    err = oracle_mod.rel_l2(X, ref)
    print(f"  plan {plan.info} rel_l2={err:.3e}")
    assert err <= REL_L2_PLAIN


# Bluestein lengths with a fused row-kernel shape (one per instantiated shape)
FUSED_LENGTHS = [34992, 33600, 36864, 35840, 17496, 16464, 5184, 4704,
                 1024, 1152, 1296, 1458, 1536, 1728, 1944, 2048, 2304, 2592, 2916, 3072, 3456,
                 3888, 4096, 4374, 4608, 5832, 6144, 6912, 7776, 8192, 8748, 9216, 10368, 11664,
                 12288, 13122, 13824, 15552, 16384, 18432, 20736, 23328, 24576, 26244, 27648,
                 31104, 32768]


@pytest.mark.parametrize("M", FUSED_LENGTHS)
@pytest.mark.parametrize("rows,tile", [(13, 128), (2, 128), (40, 5)])
def test_fused_row_kernel(gpu, oracle_mod, M, rows, tile):
    """k_rowfused for every instantiated shape: against the FP64 oracle and against
    the four-step kernels at the same M, for the longest row the length allows
    (cols = (M + 1) / 2) and a shorter one, odd and even row counts, row counts
    that are not a multiple of the rows per workgroup, and several tiles."""
    sp = gpu
    for cols in ((M + 1) // 2, M // 3 + 5):
        rng = np.random.default_rng(M + cols + rows)
        r, c = _random_coo(rng, rows, cols, 4 * cols)
        print(f"[fused] M={M} rows={rows} cols={cols} tile={tile}", flush=True)
        plan, X = _run(sp, rows, cols, r, c, fft_len=M, u_tile=tile, row_kernel=2)
        assert plan.info["row_kernel"] == 2 and plan.info["fft_len"] == M
        _, X4 = _run(sp, rows, cols, r, c, fft_len=M, u_tile=tile, row_kernel=1)
        ref = oracle_mod.sparse_rfft2_f64(rows, cols, r, c)
        err, err4 = oracle_mod.rel_l2(X, ref), oracle_mod.rel_l2(X4, ref)
        print(f"  plan m1={plan.info['m1']} radix={plan.info['conv_radix']} rel_l2 fused={err:.3e} "
              f"four-step={err4:.3e}")
        assert err <= REL_L2_PLAIN and err4 <= REL_L2_PLAIN
        assert oracle_mod.rel_l2(X, X4) <= 2 * REL_L2_PLAIN


# rows_per_lane 32: the matrix-core column pass (k_colpass_mx, 1024-row blocks)
@pytest.mark.parametrize("rows_per_lane", [1, 2, 4, 8, 16, 32])
@pytest.mark.parametrize("table", [1, 2])
@pytest.mark.parametrize("shape", [(2310, 2310, 16000, 128), (300, 4000, 9000, 100), (4095, 33, 2000, 4095)])
def test_colpass_variants(gpu, oracle_mod, rows_per_lane, table, shape):
    sp = gpu
    rows, cols, nnz, tile = shape
    rng = np.random.default_rng(rows + cols)
    r, c = _random_coo(rng, rows, cols, nnz)
    print(f"[colpass] {shape} rows/lane={rows_per_lane} table={table}", flush=True)
    plan, X = _run(sp, rows, cols, r, c, u_tile=tile, colpass_rows=rows_per_lane,
                   colpass_table=table)
    assert plan.info["colpass_rows"] == rows_per_lane
    assert plan.info["colpass_full_table"] == (1 if table == 2 else 0)
    ref = oracle_mod.sparse_rfft2_f64(rows, cols, r, c)
    assert oracle_mod.rel_l2(X, ref) <= REL_L2_PLAIN


# Half-period row pairing (even R, one tile of all rows): rows u <= R/4 and R/2 - u
# from one phase lookup.  R/2 odd and even (the self-paired row u = R/4), the
# smallest R, and int32 indices (R > 65536).
@pytest.mark.parametrize("rows_per_lane", [2, 4, 8, 16, 32])
@pytest.mark.parametrize("table", [1, 2])
@pytest.mark.parametrize("shape", [(2310, 2310, 16000), (300, 4000, 9000), (2, 7, 9), (4, 3, 7),
                                   (6, 40, 100), (1026, 513, 5000), (70000, 50, 3000)])
def test_colpass_pairing(gpu, oracle_mod, rows_per_lane, table, shape):
    sp = gpu
    rows, cols, nnz = shape
    if table == 2 and rows * 8 > 64 * 1024:
        pytest.skip("whole W_R table only up to 8192 rows")
    rng = np.random.default_rng(rows * 7 + cols)
    r, c = _random_coo(rng, rows, cols, nnz)
    print(f"[pair] {shape} rows/lane={rows_per_lane} table={table}", flush=True)
    plan, X = _run(sp, rows, cols, r, c, u_tile=rows, colpass_rows=rows_per_lane,
                   colpass_table=table)
    assert plan.info["colpass_paired"] == 1
    _, X0 = _run(sp, rows, cols, r, c, u_tile=rows, colpass_rows=rows_per_lane,
                 colpass_table=table, colpass_pair=sp.PAIR_OFF)
    ref = oracle_mod.sparse_rfft2_f64(rows, cols, r, c)
    err, err0 = oracle_mod.rel_l2(X, ref), oracle_mod.rel_l2(X0, ref)
    print(f"  rel_l2 paired={err:.3e} unpaired={err0:.3e}")
    assert err <= REL_L2_PLAIN and err0 <= REL_L2_PLAIN
    assert oracle_mod.rel_l2(X, X0) <= 1e-6  # two summation orders of the same terms
    assert abs(X[0, 0] - nnz) <= 1e-5 * nnz


def test_full_config_paired(gpu, oracle_mod):
    """c3 (16706 rows: R/2 = 8353 odd) with the default one-tile plan, which pairs
    rows u and R/2 - u in the column pass: spot rows on both sides of the pairing
    boundary R/4 against the FP64 oracle, plus the invariants."""
    import torch
    sp = gpu
    m = sp.config_matrix("c3")
    print("[fft] c3 paired", flush=True)
    plan = sp.Plan(m.rows, m.cols, u_tile=m.rows)
    assert plan.info["colpass_paired"] == 1
    out = plan.alloc_output()
    plan.execute(_dev(m.row_idx), _dev(m.col_idx), out)
    torch.cuda.synchronize()
    plan.check_input()
    q, h = m.rows // 4, m.rows // 2
    rows = np.array([0, 1, q - 1, q, q + 1, h - 1, h, h + 1, m.rows - q, m.rows - 1], np.int32)
    ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, u_list=rows, nthreads=8)
    got = out[torch.as_tensor(rows).cuda(), : m.cols // 2 + 1].cpu().numpy()
    err = oracle_mod.rel_l2(got, ref)
    print(f"c3 paired spot rel_l2={err:.3e}")
    assert err <= REL_L2_PLAIN
    _invariants(out, m)


# ---- precision form: BF16x3 MFMA length-16/32 sub-FFTs -------------------------
REL_L2_BF16X3 = 1e-5


@pytest.mark.parametrize("shape", [(2310, 2310, 16000), (97, 300, 3000), (50, 37, 400), (1, 5, 3),
                                   (333, 1000, 9000)])
@pytest.mark.parametrize("rule", [3, 7])
def test_bf16x3_small(gpu, oracle_mod, shape, rule):
    sp = gpu
    rows, cols, nnz = shape
    rng = np.random.default_rng(rows * 7 + cols)
    r, c = _random_coo(rng, rows, cols, nnz)
    print(f"[bf16x3] {shape} rule={rule}", flush=True)
    plan, X = _run(sp, rows, cols, r, c, length_rule=rule, precision=sp.PREC_BF16X3, u_tile=64)
    assert plan.info["precision"] == 1 and plan.info["m1"] in (16, 32)
    ref = oracle_mod.sparse_rfft2_f64(rows, cols, r, c)
    err = oracle_mod.rel_l2(X, ref)
    print(f"  M={plan.info['fft_len']} m1={plan.info['m1']} rel_l2={err:.3e}")
    assert err <= REL_L2_BF16X3


@pytest.mark.parametrize("m1", [16, 32])
def test_bf16x3_forced_m1(gpu, oracle_mod, m1):
    sp = gpu
    m = sp.config_matrix("c1")
    plan, X = _run(sp, m.rows, m.cols, m.row_idx, m.col_idx, precision=sp.PREC_BF16X3, m1=m1)
    assert plan.info["m1"] == m1
    ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, nthreads=8)
    assert oracle_mod.rel_l2(X, ref) <= REL_L2_BF16X3


@pytest.mark.parametrize("cfg", ["c2", "c3"])
def test_bf16x3_full_config(gpu, oracle_mod, cfg):
    import torch
    sp = gpu
    m = sp.config_matrix(cfg)
    plan = sp.Plan(m.rows, m.cols, precision=sp.PREC_BF16X3, u_tile=m.rows)
    out = plan.alloc_output()
    plan.execute(_dev(m.row_idx), _dev(m.col_idx), out)
    torch.cuda.synchronize()
    plan.check_input()
    rng = np.random.default_rng(5)
    rows = np.unique(np.concatenate([[0, 1, m.rows // 2, m.rows - 1], rng.choice(m.rows, 28, replace=False)]))
    ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, u_list=rows, nthreads=8)
    got = out[torch.as_tensor(rows).cuda(), : m.cols // 2 + 1].cpu().numpy()
    err = oracle_mod.rel_l2(got, ref)
    print(f"{cfg} bf16x3 M={plan.info['fft_len']} spot rel_l2={err:.3e}")
    assert err <= REL_L2_BF16X3
    _invariants(out, m)


# ---- graph and batched forms ---------------------------------------------------
def test_graph_replays_match_eager(gpu, oracle_mod):
    """The captured graph is replayed several times with eager work in between;
    every replay must equal the eager result bit for bit."""
    import torch
    sp = gpu
    m = sp.config_matrix("c1")
    plan = sp.Plan(m.rows, m.cols, u_tile=700)
    d_r, d_c = _dev(m.row_idx), _dev(m.col_idx)
    eager = plan.alloc_output()
    plan.execute(d_r, d_c, eager)
    torch.cuda.synchronize()
    ref = eager[:, : m.cols // 2 + 1].cpu().numpy()
    out = plan.alloc_output()
    g = sp.capture_plan(plan, d_r, d_c, out)
    for rep in range(4):
        out.zero_()
        g.launch()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out[:, : m.cols // 2 + 1].cpu().numpy(), ref)
        plan.execute(d_r, d_c, eager)  # eager work between replays
        torch.cuda.synchronize()
    exact = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, nthreads=8)
    assert oracle_mod.rel_l2(ref, exact) <= REL_L2_PLAIN


def test_reference_graph_entry_point(gpu, oracle_mod):
    sp = gpu
    L = sp.load()
    m = sp.config_matrix("c1")
    res = sp.Result()
    st = L.spfft_csc_bluestein_mixed_radix_graph(m.rows, m.cols, m.nnz, m.row_idx.ctypes.data,
                                                 m.col_idx.ctypes.data, 128, ctypes.byref(res))
    assert st == 0, L.spfft_last_error()
    X = sp.device_ptr_to_host(res.d_output, m.rows, res.output_cols)[:, : m.cols // 2 + 1]
    assert L.spfft_free_output(res.d_output) == 0
    ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, nthreads=8)
    assert oracle_mod.rel_l2(X, ref) <= REL_L2_PLAIN and res.ms > 0


def test_python_mirror_graph(gpu, oracle_mod):
    sp = gpu
    m = sp.config_matrix("c1")
    r = sp.sparse_fft_csc_bluestein_mixed_radix_graph(m, 128)
    X = r.d_output[:, : m.cols // 2 + 1].cpu().numpy()
    ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, nthreads=8)
    assert r.ms > 0 and oracle_mod.rel_l2(X, ref) <= REL_L2_PLAIN
    with pytest.raises(RuntimeError, match="CSC tile size must be positive"):
        sp.sparse_fft_csc_bluestein_mixed_radix_graph(m, 0)


def test_cli_reference_flags(gpu, oracle_mod, tmp_path):
    """The benchmark driver with the reference's flags runs all three variants
    through the C++ drop-in entry points and prints the reference's table lines."""
    import os
    import subprocess
    sp = gpu
    cli = os.path.join(os.path.dirname(sp.LIB_PATH), "spfft_cli")
    res = subprocess.run([cli, "gen:c1", "--csc-mixed-radix", "--csc-mixed-radix-graph",
                          "--csc-cufft-smooth", "--csc-stockham-stream", "--csc-stockham-cufft-stream",
                          "--csc-tile", "256", "--repeat", "3"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    for label in ("Sparse CSC (mixed-radix)", "Sparse CSC (mr-graph)", "Sparse CSC (cufft smooth)",
                  "Sparse CSC (Stockham stream)", "Sparse CSC (cufft stream)"):
        assert label in res.stdout
    assert "OOM" not in res.stdout
    # the reference's check lines against an FP64 reference spectrum of the same matrix,
    # and this build's native plan (--csc-auto)
    m = sp.config_matrix("c1")
    ref_path = os.path.join(str(tmp_path), "c1.ref")
    oracle_mod.write_reference(ref_path, m.rows, m.cols, m.row_idx, m.col_idx, nthreads=8)
    res = subprocess.run([cli, "gen:c1", "--csc-mixed-radix", "--csc-cufft-smooth", "--csc-auto",
                          "--csc-stockham-stream", "--repeat", "3", "--check-ref", ref_path],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    checks = [ln for ln in res.stdout.splitlines() if ln.startswith("  check ")]
    assert len(checks) == 4, res.stdout
    for ln in checks:
        f = ln.split()
        max_abs, max_rel, rms = float(f[f.index("max_abs") + 1]), float(f[f.index("max_rel") + 1]), \
            float(f[f.index("rms_abs") + 1])
        # |X| reaches nnz = 16000 at X[0][0] and is ~sqrt(nnz) = 126 on average: FP32
        # rounding leaves absolute errors of ~1e-4 (rel-L2 ~2e-7), larger at the largest
        # entries; max_rel divides by max(1, |ref|), so small entries count absolutely
        assert max_abs < 2e-2 and max_rel < 1e-3 and rms < 1e-3, ln
    assert "Sparse CSC (auto)" in res.stdout
    # --cpu-reference: the host FP32 transform's row, and its output as the ground truth of
    # every variant's check line (the reference's behaviour, src/main.cu:324-342)
    res = subprocess.run([cli, "gen:c1", "--cpu-reference", "--csc-mixed-radix", "--csc-cufft-smooth",
                          "--csc-auto"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "CPU reference (FP32 FFT)" in res.stdout and "FAIL" not in res.stdout
    checks = [ln for ln in res.stdout.splitlines() if ln.startswith("  check ")]
    assert len(checks) == 3, res.stdout
    for ln in checks:
        f = ln.split()
        max_abs, max_rel = float(f[f.index("max_abs") + 1]), float(f[f.index("max_rel") + 1])
        assert max_abs < 2e-2 and max_rel < 1e-3, ln
    bad = subprocess.run([cli, "gen:c1", "--csc-mixed-radix", "--csc-tile", "0"],
                         capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "--csc-tile must be positive" in bad.stderr


# ---- direct length-C transform (smooth widths, no Bluestein) ------------------
DIRECT_CASES = [(1024, 300, 5000, 128), (1024, 7, 40, 3), (1152, 1001, 9000, 1001), (2048, 64, 900, 64),
                (4704, 90, 3000, 17), (5184, 2310, 16000, 2310), (16384, 33, 4000, 7),
                (16384, 16384, 98244, 16384), (17496, 20, 600, 20), (18432, 5, 100, 1)]


@pytest.mark.parametrize("cols,rows,nnz,tile", DIRECT_CASES)
def test_direct_transform(gpu, oracle_mod, cols, rows, nnz, tile):
    """The direct length-cols FFT against the FP64 oracle and the Bluestein plan of the
    same matrix; the streamed output of the direct plan is bit-identical to it."""
    import torch
    sp = gpu
    rng = np.random.default_rng(cols + rows)
    r, c = _random_coo(rng, rows, cols, nnz)
    print(f"[direct] {rows}x{cols} nnz={len(r)} tile={tile}", flush=True)
    plan, X = _run(sp, rows, cols, r, c, u_tile=tile)
    assert plan.info["direct"] == 1 and plan.info["fft_len"] == cols
    _, Xb = _run(sp, rows, cols, r, c, u_tile=tile, row_transform=sp.TRANSFORM_BLUESTEIN)
    if rows * cols <= 40_000_000:
        ref = oracle_mod.sparse_rfft2_f64(rows, cols, r, c, nthreads=8)
    else:
        u = np.unique(np.concatenate([[0, 1, rows // 2, rows - 1], rng.choice(rows, 12, replace=False)]))
        ref = oracle_mod.sparse_rfft2_f64(rows, cols, r, c, u_list=u, nthreads=8)
        X, Xb = X[u], Xb[u]
    err, errb = oracle_mod.rel_l2(X, ref), oracle_mod.rel_l2(Xb, ref)
    print(f"  rel_l2 direct={err:.3e} bluestein={errb:.3e}")
    assert err <= REL_L2_PLAIN and errb <= REL_L2_PLAIN
    if rows <= 2310:
        host = plan.alloc_host_output()
        host.fill_(complex(float("nan"), float("nan")))
        plan.execute_to_host(_dev(r), _dev(c), host)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(host[:, : cols // 2 + 1].numpy(), X)


def test_direct_single_nonzero_closed_form(gpu):
    sp = gpu
    rows, cols, r0, c0 = 1000, 16384, 999, 12345
    plan, X = _run(sp, rows, cols, np.array([r0], np.int32), np.array([c0], np.int32), u_tile=rows)
    assert plan.info["direct"] == 1
    u = np.arange(rows)[:, None]
    v = np.arange(cols // 2 + 1)[None, :]
    ref = np.exp(-2j * np.pi * (((u * r0) % rows) / rows + ((v * c0) % cols) / cols))
    assert np.abs(X - ref).max() < 2e-5


# ---- Rader row transform: widths 2p, p an odd prime with a shape for p - 1 ------
RADER_CASES = [(1858, 50, 2000, 50), (1858, 7, 40, 3), (1858, 1858, 12000, 1858), (4178, 100, 5000, 33),
               (4178, 4177, 20000, 4177), (16706, 64, 3000, 64), (16706, 5, 300, 2), (1858, 40, 0, 40),
               (1858, 1, 5, 1), (1858, 2, 9, 2)]


@pytest.mark.parametrize("cols,rows,nnz,tile", RADER_CASES)
def test_rader_transform(gpu, oracle_mod, cols, rows, nnz, tile):
    """Rader's convolutions against the FP64 oracle and the Bluestein plan of the same
    matrix (odd and even rows, several tiles, the paired column pass); the streamed
    output of the Rader plan is bit-identical to it."""
    import torch
    sp = gpu
    assert cols in sp.rader_widths()
    rng = np.random.default_rng(cols + rows)
    r, c = _random_coo(rng, rows, cols, nnz)
    print(f"[rader] {rows}x{cols} nnz={len(r)} tile={tile}", flush=True)
    plan, X = _run(sp, rows, cols, r, c, u_tile=tile)
    assert plan.info["rader"] == 1 and plan.info["fft_len"] == cols // 2 - 1
    _, Xb = _run(sp, rows, cols, r, c, u_tile=tile, row_transform=sp.TRANSFORM_BLUESTEIN)
    ref = oracle_mod.sparse_rfft2_f64(rows, cols, r, c, nthreads=8)
    err, errb = oracle_mod.rel_l2(X, ref), oracle_mod.rel_l2(Xb, ref)
    print(f"  rel_l2 rader={err:.3e} bluestein={errb:.3e}")
    assert err <= REL_L2_PLAIN and errb <= REL_L2_PLAIN
    if rows <= 2310:
        host = plan.alloc_host_output()
        host.fill_(complex(float("nan"), float("nan")))
        plan.execute_to_host(_dev(r), _dev(c), host)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(host[:, : cols // 2 + 1].numpy(), X)


@pytest.mark.parametrize("cols,rows,nnz,tile", RADER_CASES)
def test_rader_mfma_transform(gpu, oracle_mod, cols, rows, nnz, tile):
    """The precision form of the Rader plan (k_rowrader_mfma: the length-29 DFTs as
    BF16x3 DFT-matrix products on the matrix cores) against the FP64 oracle."""
    sp = gpu
    rng = np.random.default_rng(cols + rows)
    r, c = _random_coo(rng, rows, cols, nnz)
    print(f"[rader-mfma] {rows}x{cols} nnz={len(r)} tile={tile}", flush=True)
    plan, X = _run(sp, rows, cols, r, c, u_tile=tile, precision=sp.PREC_BF16X3)
    assert plan.info["rader"] == 1 and plan.info["precision"] == 1 and plan.info["m1"] == 29
    ref = oracle_mod.sparse_rfft2_f64(rows, cols, r, c, nthreads=8)
    err = oracle_mod.rel_l2(X, ref)
    print(f"  rel_l2 rader-mfma={err:.3e}")
    assert err <= REL_L2_BF16X3
    # run to run the same bits
    _, X2 = _run(sp, rows, cols, r, c, u_tile=tile, precision=sp.PREC_BF16X3)
    np.testing.assert_array_equal(X, X2)


@pytest.mark.parametrize("c0", [0, 1, 2, 3, 8352, 16705])
def test_rader_mfma_single_nonzero_closed_form(gpu, c0):
    """A single nonzero in each column class through the precision form: a pure tone."""
    sp = gpu
    rows, cols, r0 = 40, 16706, 17
    plan, X = _run(sp, rows, cols, np.array([r0], np.int32), np.array([c0], np.int32), u_tile=rows,
                   precision=sp.PREC_BF16X3)
    assert plan.info["rader"] == 1 and plan.info["precision"] == 1
    u = np.arange(rows)[:, None]
    v = np.arange(cols // 2 + 1)[None, :]
    ref = np.exp(-2j * np.pi * (((u * r0) % rows) / rows + ((v * c0) % cols) / cols))
    assert np.abs(X - ref).max() < 1e-4


def test_rader_row_limit_and_tiles(gpu):
    """A Rader plan computing only rows u < 37 (and their mirror rows) in tiles of 16 rows
    matches the full one-tile run bit for bit on those rows (both with the unpaired column
    pass, whose rows u < 64 each take their own phase lookup)."""
    import torch
    sp = gpu
    rows, cols = 300, 4178
    r, c = _random_coo(np.random.default_rng(11), rows, cols, 6000)
    full, X = _run(sp, rows, cols, r, c, u_tile=rows, colpass_pair=sp.PAIR_OFF)
    plan = sp.Plan(rows, cols, u_tile=16, row_limit=37)
    assert plan.info["rader"] == 1 and full.info["rader"] == 1
    out = plan.alloc_output()
    out.fill_(complex(float("nan"), float("nan")))
    plan.execute(_dev(r), _dev(c), out)
    torch.cuda.synchronize()
    V = cols // 2 + 1
    got = out[:, :V].cpu().numpy()
    u = np.arange(37)
    np.testing.assert_array_equal(got[u], X[u])
    np.testing.assert_array_equal(got[rows - u[1:]], X[rows - u[1:]])


@pytest.mark.parametrize("c0", [0, 1, 2, 3, 8352, 16705])
def test_rader_single_nonzero_closed_form(gpu, c0):
    """One nonzero per column class: columns 0 and 1 (x_s[0] of the two halves), the
    first columns of each Rader run and the last column."""
    sp = gpu
    rows, cols, r0 = 1000, 16706, 999
    plan, X = _run(sp, rows, cols, np.array([r0], np.int32), np.array([c0], np.int32), u_tile=rows)
    assert plan.info["rader"] == 1
    u = np.arange(rows)[:, None]
    v = np.arange(cols // 2 + 1)[None, :]
    ref = np.exp(-2j * np.pi * (((u * r0) % rows) / rows + ((v * c0) % cols) / cols))
    assert np.abs(X - ref).max() < 2e-5


# ---- streamed output: device chunks copied to pinned host memory --------------
STREAM_CASES = [((2310, 2310, 16000), 128, {}), ((2310, 2310, 16000), 2310, {}),
                ((2310, 2310, 16000), 500, dict(row_kernel=1)),
                ((2310, 2310, 16000), 300, dict(precision=1)),
                ((7, 9, 30), 1, {}), ((8, 9, 30), 3, {}), ((2, 5, 6), 1, {}), ((1, 5, 3), 4, {}),
                ((301, 1000, 5000), 50, {}), ((300, 4000, 9000), 151, {}), ((300, 4000, 9000), 77, {})]


@pytest.mark.parametrize("shape,tile,kw", STREAM_CASES)
def test_streamed_output(gpu, oracle_mod, shape, tile, kw):
    """execute_to_host gives exactly the device-output transform: every row (tile
    rows and mirror rows, host buffer pre-filled with NaN) on two calls of one plan."""
    import torch
    sp = gpu
    rows, cols, nnz = shape
    rng = np.random.default_rng(rows + 3 * cols + tile)
    r, c = _random_coo(rng, rows, cols, nnz)
    print(f"[stream] {shape} tile={tile} {kw}", flush=True)
    plan, X = _run(sp, rows, cols, r, c, u_tile=tile, **kw)
    V = cols // 2 + 1
    dr, dc = _dev(r), _dev(c)
    for _ in range(2):
        host = plan.alloc_host_output()
        host.fill_(complex(float("nan"), float("nan")))
        plan.execute_to_host(dr, dc, host)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(host[:, :V].numpy(), X)
    ref = oracle_mod.sparse_rfft2_f64(rows, cols, r, c)
    assert oracle_mod.rel_l2(X, ref) <= (REL_L2_BF16X3 if kw.get("precision") else REL_L2_PLAIN)


def test_streaming_entry_points(gpu, oracle_mod):
    """The reference-shaped streamed variants: h_output set, d_output None, same result."""
    sp = gpu
    m = sp.config_matrix("c1")
    V = m.cols // 2 + 1
    ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, nthreads=8)
    for fn in (sp.sparse_fft_csc_stockham_streaming, sp.sparse_fft_csc_bluestein_cufft_streaming):
        res = fn(m, 200)
        assert res.d_output is None and res.output_cols == (V + 3) // 4 * 4 and res.ms > 0
        assert oracle_mod.rel_l2(res.h_output[:, :V].numpy(), ref) <= REL_L2_PLAIN
    with pytest.raises(RuntimeError):
        sp.sparse_fft_csc_stockham_streaming(m, 0)


def test_batch_graph_with_rader_widths(gpu, oracle_mod):
    """A batch mixing Rader widths (1858, 4178, 16706), a direct width and a Bluestein one,
    one tile per matrix: eager against the FP64 oracle (sampled rows for the 16706 one), and
    two graph replays bit-identical to the eager run."""
    import torch
    sp = gpu
    sizes = (1858, 1500, 4178, 1024, 16706)
    mats = [sp.generate("uniform", n, n * (6 if n < 10000 else 2), seed=90 + i) for i, n in enumerate(sizes)]
    b = sp.Batch([(m.rows, m.cols) for m in mats], opts=[dict(u_tile=m.rows) for m in mats],
                 nnz_hint=[m.nnz for m in mats], streams=2)
    assert [inf["rader"] for inf in b.infos] == [1, 0, 1, 0, 1]
    assert [inf["direct"] for inf in b.infos] == [0, 0, 0, 1, 0]
    print("[batch] rader widths", flush=True)
    d_r, d_c, off = sp.Batch.pack(mats)
    arena = b.alloc_output()
    b.execute(d_r, d_c, off, arena)
    torch.cuda.synchronize()
    b.check_input()
    eager = arena.clone()
    for i, m in enumerate(mats):
        X = b.output_of(arena, i)[:, : m.cols // 2 + 1].cpu().numpy()
        if m.rows <= 5000:
            ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, nthreads=8)
        else:
            u = np.array([0, 1, 2, m.rows // 4, m.rows // 2, m.rows // 2 + 1, m.rows - 2, m.rows - 1])
            ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, u_list=u, nthreads=8)
            X = X[u]
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


def _batch_mats(n, seed):
    import cuspfft_amd as sp
    rng = np.random.default_rng(seed)
    mats = []
    for i in range(n):
        size = int(np.exp(rng.uniform(np.log(60), np.log(1500))))
        deg = rng.uniform(4, 30)
        if i % 2 == 0:
            mats.append(sp.generate("uniform", size, size * deg / 2, seed=seed * 1000 + i))
        else:
            mats.append(sp.generate("chunglu", size, size * deg, rng.uniform(2.0, 2.5), 0,
                                    seed=seed * 1000 + i))
    return mats


@pytest.mark.parametrize("streams", [1, 3])
def test_batch_execute_and_graph(gpu, oracle_mod, streams):
    import torch
    sp = gpu
    mats = _batch_mats(12, 3)
    opts = [dict(length_rule=3 if i % 3 else 7, precision=1 if i % 4 == 1 else 0,
                 u_tile=max(m.rows // (1 + i % 3), 1)) for i, m in enumerate(mats)]
    b = sp.Batch([(m.rows, m.cols) for m in mats], opts=opts, nnz_hint=[m.nnz for m in mats],
                 streams=streams)
    d_r, d_c, off = sp.Batch.pack(mats)
    arena = b.alloc_output()
    b.execute(d_r, d_c, off, arena)
    torch.cuda.synchronize()
    b.check_input()
    eager = arena.clone()
    for i, m in enumerate(mats):
        X = b.output_of(arena, i)[:, : m.cols // 2 + 1].cpu().numpy()
        ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, nthreads=8)
        tol = REL_L2_BF16X3 if b.infos[i]["precision"] == 1 else REL_L2_PLAIN
        assert oracle_mod.rel_l2(X, ref) <= tol, i
    g = b.capture(d_r, d_c, off, arena)
    for rep in range(3):
        arena.zero_()
        g.launch()
        torch.cuda.synchronize()
        for i, m in enumerate(mats):
            V = m.cols // 2 + 1
            np.testing.assert_array_equal(b.output_of(arena, i)[:, :V].cpu().numpy(),
                                          b.output_of(eager, i)[:, :V].cpu().numpy())
        b.execute(d_r, d_c, off, eager)
        torch.cuda.synchronize()


def test_batch_shared_launches_match_per_matrix(gpu, oracle_mod):
    """The descriptor-driven batch (one CSC build of the block-diagonal matrix, one launch
    per kernel shape over (matrix, block) work items) against every matrix on its own
    launches: bit-identical outputs, also through graph replays; matrices that cannot be
    members (four-step plans, several tiles) run beside them; a bad index is reported for
    its own matrix."""
    import torch
    sp = gpu
    mats = _batch_mats(24, 7)
    for cols, rows in ((1858, 300), (4178, 97), (1024, 500), (2048, 64), (16706, 40)):
        mats.append(_rect(sp, rows, cols, rows * 12, cols))
    opts = [dict(u_tile=m.rows) for m in mats]
    opts[3] = dict(u_tile=max(mats[3].rows // 3, 1))                 # several tiles
    opts[5] = dict(u_tile=mats[5].rows, row_kernel=sp.ROW_FOUR_STEP)  # four-step kernels
    opts[-1] = dict(u_tile=mats[-1].rows, precision=sp.PREC_BF16X3)  # the MFMA Rader kernel
    shapes = [(m.rows, m.cols) for m in mats]
    d_r, d_c, off = sp.Batch.pack(mats)
    auto = sp.Batch(shapes, opts=opts, nnz_hint=[m.nnz for m in mats], streams=2)
    mode, members = auto.mode()
    print(f"[batch] {len(mats)} matrices, {members} members", flush=True)
    assert mode == sp.BATCH_AUTO and 0 < members <= len(mats) - 2
    per = sp.Batch(shapes, opts=opts, nnz_hint=[m.nnz for m in mats], streams=2)
    per.set_mode(sp.BATCH_PER_MATRIX)
    assert per.mode() == (sp.BATCH_PER_MATRIX, 0)
    a_auto, a_per = auto.alloc_output(), per.alloc_output()
    auto.execute(d_r, d_c, off, a_auto)
    per.execute(d_r, d_c, off, a_per)
    torch.cuda.synchronize()
    auto.check_input()
    per.check_input()
    for i, m in enumerate(mats):
        V = m.cols // 2 + 1
        assert torch.equal(auto.output_of(a_auto, i)[:, :V], per.output_of(a_per, i)[:, :V]), i
        if i % 4 == 0 or i >= 24:
            X = auto.output_of(a_auto, i)[:, :V].cpu().numpy()
            ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, nthreads=8)
            tol = REL_L2_BF16X3 if auto.infos[i]["precision"] == 1 else REL_L2_PLAIN
            assert oracle_mod.rel_l2(X, ref) <= tol, i
    g = auto.capture(d_r, d_c, off, auto.alloc_output())
    replay = g._keep[3]
    for rep in range(2):
        replay.zero_()
        g.launch()
        torch.cuda.synchronize()
        for i, m in enumerate(mats):
            V = m.cols // 2 + 1
            assert torch.equal(auto.output_of(replay, i)[:, :V], auto.output_of(a_auto, i)[:, :V]), (rep, i)
        auto.execute(d_r, d_c, off, a_auto)  # eager work between replays
        torch.cuda.synchronize()
    # an index outside a member matrix: reported for that matrix
    k = next(i for i, inf in enumerate(auto.infos)
             if inf["row_kernel"] == sp.ROW_FUSED and inf["u_tile"] >= mats[i].rows // 2 + 1 and mats[i].nnz)
    col = d_c.clone()
    col[int(off[k])] = mats[k].cols + 5
    auto.execute(d_r, col, off, a_auto)
    with pytest.raises(sp.SpfftError, match=f"matrix {k}:"):
        auto.check_input()


def _rect(sp, rows, cols, nnz, seed):
    rng = np.random.default_rng(seed)
    r, c = _random_coo(rng, rows, cols, nnz)
    return sp.COOMatrix(rows, cols, r, c)


def test_batch_graph_with_direct_widths(gpu, oracle_mod):
    """A batch mixing direct-transform widths (1024, 1152, 2048) with Bluestein ones,
    one tile per matrix (paired column passes for even rows): eager against the FP64
    oracle, and three graph replays bit-identical to eager runs."""
    import torch
    sp = gpu
    mats = [sp.generate("uniform", n, n * 6, seed=70 + i) for i, n in enumerate((1024, 1500, 1152, 2048, 999))]
    b = sp.Batch([(m.rows, m.cols) for m in mats], opts=[dict(u_tile=m.rows) for m in mats],
                 nnz_hint=[m.nnz for m in mats], streams=2)
    assert [inf["direct"] for inf in b.infos] == [1, 0, 1, 1, 0]
    d_r, d_c, off = sp.Batch.pack(mats)
    arena = b.alloc_output()
    b.execute(d_r, d_c, off, arena)
    torch.cuda.synchronize()
    b.check_input()
    eager = arena.clone()
    for i, m in enumerate(mats):
        X = b.output_of(arena, i)[:, : m.cols // 2 + 1].cpu().numpy()
        ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, nthreads=8)
        assert oracle_mod.rel_l2(X, ref) <= REL_L2_PLAIN, i
    g = b.capture(d_r, d_c, off, arena)
    for rep in range(3):
        arena.zero_()
        g.launch()
        torch.cuda.synchronize()
        for i, m in enumerate(mats):
            V = m.cols // 2 + 1
            np.testing.assert_array_equal(b.output_of(arena, i)[:, :V].cpu().numpy(),
                                          b.output_of(eager, i)[:, :V].cpu().numpy())

