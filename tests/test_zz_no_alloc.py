"""spfft_csc_build_noalloc, the allocation-free CSC build, against spfft_csc_build
(+ spfft_csc_get) through the C ABI: bit-exact on two calls in a row, the second with
a smaller input, so the shared scratch array's leftovers from the first call must not
leak into the second."""
import ctypes

import numpy as np
import pytest


def _existing(sp, rows, cols, r, c):
    import torch
    plan = sp.Plan(rows, cols)
    plan.csc_build(torch.as_tensor(r).cuda(), torch.as_tensor(c).cuda())
    return plan.csc_get()


def _noalloc(sp, rows, cols, r, c):
    import torch
    L = sp.load()
    nnz = len(r)
    packed = rows <= 65536 and cols <= 65536
    it = torch.int16 if packed else torch.int32
    d_r = torch.as_tensor(r).cuda()
    d_c = torch.as_tensor(c).cuda()
    sc = torch.full((cols,), -1, dtype=it, device="cuda")
    cp = torch.full((cols + 1,), -1, dtype=torch.int32, device="cuda")
    ri = torch.full((max(nnz, 1),), -1, dtype=it, device="cuda")
    cs = torch.full((cols + 1,), -1, dtype=torch.int32, device="cuda")
    info = torch.full((3,), -1, dtype=torch.int32, device="cuda")
    vp = ctypes.c_void_p
    st = L.spfft_csc_build_noalloc(rows, cols, vp(d_r.data_ptr()), vp(d_c.data_ptr()), nnz, vp(sc.data_ptr()),
                                   vp(cp.data_ptr()), vp(ri.data_ptr()), vp(cs.data_ptr()), vp(info.data_ptr()),
                                   vp(torch.cuda.current_stream().cuda_stream))
    assert st == 0, L.spfft_last_error()
    torch.cuda.synchronize()
    nk, nsp, err = info.cpu().tolist()
    npt = np.uint16 if packed else np.int32
    return {"n_sparse_cols": nsp, "nnz": nk, "err": err,
            "sparse_cols": sc[:nsp].cpu().numpy().view(npt), "col_ptr": cp[: nsp + 1].cpu().numpy(),
            "row_idx": ri[:nnz].cpu().numpy().view(npt), "col_start": cs.cpu().numpy()}


def _same(a, b):
    assert a["n_sparse_cols"] == b["n_sparse_cols"]
    for k in ("sparse_cols", "col_ptr", "row_idx", "col_start"):
        assert a[k].dtype == b[k].dtype and np.array_equal(a[k], b[k]), k


@pytest.mark.gpu
def test_noalloc_csc_matches_existing_entry_point(gpu, oracle_mod):
    sp = gpu
    L = sp.load()
    L.spfft_csc_build_noalloc.restype = ctypes.c_int
    cases = [sp.config_matrix("c3"),                    # 16706^2, 121,250 nnz (two sort passes)
             sp.config_matrix("c1"),                    # smaller: 2310^2, 16,000 nnz
             sp.generate("uniform", 200, 300, seed=4)]  # smaller again: one sort pass
    for m in cases:
        print(f"[noalloc] {m.rows}x{m.cols} nnz={m.nnz}", flush=True)
        got = _noalloc(sp, m.rows, m.cols, m.row_idx, m.col_idx)
        assert got["err"] == 0 and got["nnz"] == m.nnz
        _same(got, _existing(sp, m.rows, m.cols, m.row_idx, m.col_idx))
        sc, cp, ri = oracle_mod.compact_csc(m.rows, m.cols, m.row_idx, m.col_idx)
        assert np.array_equal(got["sparse_cols"].astype(np.int32), sc)
        assert np.array_equal(got["col_ptr"], cp)
        assert np.array_equal(got["row_idx"].astype(np.int32), ri)
    # an out-of-range pair is counted in d_info[2] and dropped, as spfft_csc_build does
    r = np.array([0, 5, 2, 1], np.int32)
    c = np.array([1, 2, 9, 3], np.int32)
    got = _noalloc(sp, 5, 6, r, c)
    assert got["err"] != 0 and got["nnz"] == 2


def test_noalloc_rejects_oversize_without_a_device():
    import cuspfft_amd as sp
    L = sp.load()
    vp = ctypes.c_void_p
    one = vp(8)
    for rows, cols, nnz in ((10, 65537, 5), (10, 100, (1 << 20) + 1), (0, 10, 1)):
        st = L.spfft_csc_build_noalloc(rows, cols, one, one, nnz, one, one, one, one, one, None)
        assert st == 1, (rows, cols, nnz)
    assert b"nnz <=" in L.spfft_last_error() or b"shape" in L.spfft_last_error()
