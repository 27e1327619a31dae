"""Pins the CPU oracle (oracle/spfft_oracle.cpp) before it is trusted as the checker.

Each oracle function is checked against something other than its own formula:
hand-built fixtures, a brute-force stable sort, committed golden vectors made by
numpy.fft.rfft2 on the dense matrix, direct O(n^2) DFTs, closed forms, and
invariants that a dropped term, a wrong sign or a transposed operand breaks.
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "small_cases.npz")


def _is_smooth(m, primes):
    for p in primes:
        while m % p == 0:
            m //= p
    return m == 1


# ---- smooth lengths and radix order (sparse_fft_bluestein.cu:1975-2011) ----
def test_smooth_length_constants(oracle_mod):
    # M3 / M7 for the benchmark sizes (closed form: smallest 3- / 7-smooth >= 2C-1)
    table = {2310: (5184, 4704), 8219: (17496, 16464), 16706: (34992, 33600),
             40421: (82944, 81000)}
    for C, (m3, m7) in table.items():
        assert oracle_mod.next_3_smooth(2 * C - 1) == m3
        assert oracle_mod.next_7_smooth(2 * C - 1) == m7


def test_smooth_length_bruteforce(oracle_mod):
    for n in range(1, 3000):
        m3 = oracle_mod.next_3_smooth(n)
        assert m3 >= n and _is_smooth(m3, (2, 3))
        assert not any(_is_smooth(k, (2, 3)) for k in range(n, m3))
        m7 = oracle_mod.next_7_smooth(n)
        assert m7 >= n and _is_smooth(m7, (2, 3, 5, 7))
        assert not any(_is_smooth(k, (2, 3, 5, 7)) for k in range(n, m7))


def test_factor_order(oracle_mod):
    # benzene's M = 17,496 -> [9, 9, 9, 3, 2, 2, 2] (reference README.md:433-437)
    assert oracle_mod.factor_3_smooth(17496) == [9, 9, 9, 3, 2, 2, 2]
    assert oracle_mod.factor_3_smooth(5184) == [9, 9, 2, 2, 2, 2, 2, 2]
    assert oracle_mod.factor_3_smooth(1) == []
    with pytest.raises(RuntimeError):
        oracle_mod.factor_3_smooth(10)
    for n in (2, 3, 6, 12, 81, 4096, 34992):
        f = oracle_mod.factor_3_smooth(n)
        assert int(np.prod(f)) == n
        assert f == sorted(f, reverse=True)


# ---- CSC (sparse_fft_internal.cuh:102-124, 200-222) ----
def test_csc_handbuilt(oracle_mod):
    # 4x6 matrix, input deliberately not in column order; column 2 and 4 empty,
    # column 5 holds rows 3, 0, 2 in input order (stable, not sorted)
    r = np.array([3, 1, 0, 0, 2, 1, 2, 3], np.int32)
    c = np.array([5, 0, 5, 1, 5, 3, 0, 3], np.int32)
    sc, cp, ri = oracle_mod.compact_csc(4, 6, r, c)
    np.testing.assert_array_equal(sc, [0, 1, 3, 5])
    np.testing.assert_array_equal(cp, [0, 2, 3, 5, 8])
    np.testing.assert_array_equal(ri, [1, 2, 0, 1, 3, 3, 0, 2])
    # empty matrix
    sc, cp, ri = oracle_mod.compact_csc(3, 3, np.zeros(0, np.int32), np.zeros(0, np.int32))
    assert len(sc) == 0 and list(cp) == [0] and len(ri) == 0
    # a single full column
    r = np.arange(9, dtype=np.int32)[::-1].copy()
    sc, cp, ri = oracle_mod.compact_csc(9, 4, r, np.full(9, 2, np.int32))
    np.testing.assert_array_equal(sc, [2])
    np.testing.assert_array_equal(cp, [0, 9])
    np.testing.assert_array_equal(ri, r)


def test_csc_bruteforce(oracle_mod):
    rng = np.random.default_rng(11)
    for trial in range(30):
        R, C = rng.integers(1, 300, size=2)
        nnz = int(rng.integers(0, min(R * C, 4000) + 1))
        idx = rng.choice(R * C, nnz, replace=False)
        r = (idx // C).astype(np.int32)
        c = (idx % C).astype(np.int32)
        sc, cp, ri = oracle_mod.compact_csc(R, C, r, c)
        order = sorted(range(nnz), key=lambda i: int(c[i]))  # stable
        np.testing.assert_array_equal(ri, r[order])
        np.testing.assert_array_equal(sc, np.unique(c))
        counts = np.array([np.sum(c == x) for x in np.unique(c)], dtype=np.int64)
        np.testing.assert_array_equal(cp, np.concatenate([[0], np.cumsum(counts)]))


def test_pack_u16_rule(oracle_mod):
    assert oracle_mod.can_pack_u16(65536, 65536)
    assert not oracle_mod.can_pack_u16(65537, 10)
    assert not oracle_mod.can_pack_u16(10, 65537)


# ---- chirp and Bluestein kernel (sparse_fft_internal.cuh:168-191) ----
@pytest.mark.parametrize("C", [1, 2, 7, 2310, 8219, 16706, 40421])
def test_chirp_closed_form(oracle_mod, C):
    h = oracle_mod.bluestein_chirp(C)
    k = np.arange(C, dtype=np.int64)
    exact = np.exp(-1j * np.pi * ((k * k) % (2 * C)) / C)  # exact integer reduction
    ulp = np.finfo(np.float32).eps
    assert np.abs(h.real - exact.real).max() <= 2 * ulp
    assert np.abs(h.imag - exact.imag).max() <= 2 * ulp


def test_bluestein_b_structure(oracle_mod):
    C, M = 2310, 5184
    b = oracle_mod.bluestein_b(C, M)
    h = oracle_mod.bluestein_chirp(C)
    assert b[0] == 1
    np.testing.assert_array_equal(b[1:C], np.conj(h[1:C]))
    np.testing.assert_array_equal(b[M - C + 1:][::-1], b[1:C])
    assert np.all(b[C:M - C + 1] == 0)


# ---- FP64 FFT ----
@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 12, 97, 1000, 1024, 2310, 8219])
def test_fft64_vs_numpy(oracle_mod, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    got = oracle_mod.fft64(x, -1)
    assert np.linalg.norm(got - ref) <= 1e-12 * max(np.linalg.norm(ref), 1)
    inv = oracle_mod.fft64(x, +1)
    assert np.linalg.norm(inv - np.fft.ifft(x) * n) <= 1e-12 * max(np.linalg.norm(ref), 1)


@pytest.mark.parametrize("n", [3, 6, 11, 16, 25])
def test_fft64_vs_direct_dft(oracle_mod, n):
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    F = np.exp(-2j * np.pi * np.outer(k, k) / n)
    assert np.abs(oracle_mod.fft64(x, -1) - F @ x).max() < 1e-12


# ---- whole 2-D transform ----
def test_golden_fixtures(oracle_mod):
    g = np.load(GOLDEN)
    names = sorted({k.split("__")[0] for k in g.files})
    assert len(names) >= 10
    for name in names:
        R, C = (int(v) for v in g[f"{name}__shape"])
        r, c = g[f"{name}__row"], g[f"{name}__col"]
        sc, cp, ri = oracle_mod.compact_csc(R, C, r, c)
        np.testing.assert_array_equal(sc, g[f"{name}__sparse_cols"])
        np.testing.assert_array_equal(cp, g[f"{name}__col_ptr"])
        np.testing.assert_array_equal(ri, g[f"{name}__row_idx"])
        X = oracle_mod.sparse_rfft2_f64(R, C, r, c)
        ref = g[f"{name}__X"]
        assert np.abs(X - ref).max() <= 1e-11 * max(1.0, np.abs(ref).max()), name
        for rule in (3, 7):
            Y = oracle_mod.sparse_rfft2_f32(R, C, r, c, smooth=rule)
            assert oracle_mod.rel_l2(Y, ref) < 2e-6, (name, rule)


def test_dense_numpy_medium(oracle_mod):
    rng = np.random.default_rng(5)
    for R, C, nnz in [(300, 301, 4000), (257, 512, 3000), (1000, 999, 20000)]:
        idx = rng.choice(R * C, nnz, replace=False)
        r, c = (idx // C).astype(np.int32), (idx % C).astype(np.int32)
        D = np.zeros((R, C))
        D[r, c] = 1
        ref = np.fft.rfft2(D)
        X = oracle_mod.sparse_rfft2_f64(R, C, r, c, nthreads=8)
        assert oracle_mod.rel_l2(X, ref) < 1e-13


def _c1():
    import cuspfft_amd as sp
    return sp.config_matrix("c1")


def test_invariants_c1(oracle_mod):
    m = _c1()
    X = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, nthreads=8)
    V = m.cols // 2 + 1
    assert abs(X[0, 0] - m.nnz) < 1e-9
    # Parseval over the half spectrum
    w = np.full(V, 2.0)
    w[0] = 1.0
    if m.cols % 2 == 0:
        w[-1] = 1.0
    assert abs((np.abs(X) ** 2 * w).sum() / (m.rows * m.cols * m.nnz) - 1) < 1e-12
    # symmetric input -> X[u][v] = X[v][u]
    assert np.abs(X[:V, :V] - X[:V, :V].T).max() < 1e-9
    # column 0 is the DFT of the row degrees, row 0 the DFT of the column degrees
    rdeg = np.bincount(m.row_idx, minlength=m.rows).astype(float)
    cdeg = np.bincount(m.col_idx, minlength=m.cols).astype(float)
    assert np.abs(X[:, 0] - np.fft.fft(rdeg)).max() < 1e-9
    assert np.abs(X[0, :] - np.fft.rfft(cdeg)).max() < 1e-9


def test_invariants_catch_planted_mistakes(oracle_mod):
    """The invariants above are not vacuous: a sign error (conjugate spectrum) or a
    transposed operand (rows and columns swapped) of a non-symmetric input fails them."""
    rng = np.random.default_rng(9)
    R, C = 40, 60
    idx = rng.choice(R * C, 300, replace=False)
    r, c = (idx // C).astype(np.int32), (idx % C).astype(np.int32)
    X = oracle_mod.sparse_rfft2_f64(R, C, r, c)
    rdeg = np.bincount(r, minlength=R).astype(float)
    assert np.abs(X[:, 0] - np.fft.fft(rdeg)).max() < 1e-9
    assert np.abs(np.conj(X)[:, 0] - np.fft.fft(rdeg)).max() > 1e-3
    Xt = oracle_mod.sparse_rfft2_f64(C, R, c, r)  # transposed operand
    cdeg = np.bincount(c, minlength=C).astype(float)
    assert np.abs(Xt[0, :] - np.fft.rfft(cdeg)[: R // 2 + 1]).max() > 1e-3


def test_f32_port_accuracy(oracle_mod):
    m = _c1()
    rows = np.arange(0, m.rows, 37, dtype=np.int32)
    X64 = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, u_list=rows)
    for rule in (3, 7):
        X32 = oracle_mod.sparse_rfft2_f32(m.rows, m.cols, m.row_idx, m.col_idx, smooth=rule,
                                          u_list=rows)
        assert oracle_mod.rel_l2(X32, X64) < 1e-6
