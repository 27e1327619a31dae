"""CPU tests of the C ABI library: it loads without a GPU, exports every symbol the
headers declare, and its host-only functions (planner, smooth lengths, generators,
Matrix Market reader) behave as documented.  No device code runs here."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_c_functions():
    txt = open(os.path.join(ROOT, "include", "spfft_c.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(spfft_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import cuspfft_amd as sp
    sp.load()
    out = subprocess.run(["nm", "-D", "--defined-only", sp.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (\S+)", out))
    missing = [f for f in _declared_c_functions() if f not in exported]
    assert not missing, missing
    # the C++ drop-in entry points of sparse_fft.h / mtx_reader.h
    dem = subprocess.run(["nm", "-DC", "--defined-only", sp.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for sym in ("sparse_fft_csc_bluestein_mixed_radix(COOMatrix const&, int)",
                "sparse_fft_csc_bluestein_mixed_radix_graph(COOMatrix const&, int)",
                "sparse_fft_csc_bluestein_cufft_smooth(COOMatrix const&, int)",
                "read_mtx("):
        assert sym in dem, sym


def test_smooth_lengths_match_oracle(oracle_mod):
    import cuspfft_amd as sp
    for n in list(range(1, 400)) + [4619, 16411, 33411, 80841]:
        assert sp.next_3_smooth(n) == oracle_mod.next_3_smooth(n)
        assert sp.next_7_smooth(n) == oracle_mod.next_7_smooth(n)
    for m in (1, 6, 5184, 17496, 34992, 82944):
        assert sp.factor_3_smooth(m) == oracle_mod.factor_3_smooth(m)


def test_planner_covers_many_shapes():
    import cuspfft_amd as sp
    # (p^k + 1) / 2 makes 2*cols - 1 a pure prime power (e.g. 5^7 = 78125), whose
    # smooth length may have no supported split: the planner moves to the next one
    powers = [(p ** k + 1) // 2 for p, k in ((3, 9), (3, 10), (5, 5), (5, 6), (5, 7), (7, 5), (7, 6))]
    for cols in list(range(1, 200)) + list(range(200, 70000, 997)) + powers:
        for rule in (3, 7):
            info = sp.describe_plan(17, cols, length_rule=rule)
            M = info["fft_len"]
            if info["row_transform"] == sp.TRANSFORM_SPLIT:
                # cols = m1 * m2, the width-m2 row pass at its direct or Bluestein length
                assert info["m1"] * info["m2"] == cols and 2 <= info["m1"] <= 127
                assert M == info["m2"] or M >= 2 * info["m2"] - 1
                continue
            assert M >= 2 * cols - 1
            assert info["m1"] * info["m2"] == M
            assert 1 <= info["m1"] <= 64
            assert int(np.prod(info["conv_radix"])) == info["m2"] or info["m2"] == 1
            assert info["pitch"] % 4 == 0 and info["pitch"] >= cols // 2 + 1


def test_planner_config_choices():
    import cuspfft_amd as sp
    for cols, m3, m7 in [(2310, 5184, 4704), (8219, 17496, 16464), (16706, 34992, 33600)]:
        a = sp.describe_plan(cols, cols, length_rule=3, row_transform=sp.TRANSFORM_BLUESTEIN)
        b = sp.describe_plan(cols, cols, length_rule=7, row_transform=sp.TRANSFORM_BLUESTEIN)
        assert a["fft_len"] == m3 and b["fft_len"] == m7
        assert a["conv_fast"] == 1 and b["conv_fast"] == 1


def test_planner_colpass_pairing():
    """Rows u and R/2 - u share the column pass only for even R with one tile of
    every computed row and no sampled row limit."""
    import cuspfft_amd as sp
    assert sp.describe_plan(16706, 16706, u_tile=16706)["colpass_paired"] == 1
    assert sp.describe_plan(16706, 16706, u_tile=8354)["colpass_paired"] == 1  # = R/2 + 1
    assert sp.describe_plan(16706, 16706, u_tile=8353)["colpass_paired"] == 0
    assert sp.describe_plan(8219, 8219, u_tile=8219)["colpass_paired"] == 0   # odd R
    assert sp.describe_plan(2, 5, u_tile=2)["colpass_paired"] == 1
    assert sp.describe_plan(2310, 2310, u_tile=2310, row_limit=64)["colpass_paired"] == 0
    assert sp.describe_plan(2310, 2310, u_tile=2310, colpass_pair=sp.PAIR_OFF)["colpass_paired"] == 0
    assert sp.describe_plan(2310, 2310, u_tile=2310, colpass_rows=1)["colpass_paired"] == 0
    assert sp.describe_plan(2310, 2310, u_tile=2310, colpass_rows=2)["colpass_paired"] == 1
    with pytest.raises(sp.SpfftError):
        sp.describe_plan(2310, 2310, colpass_pair=2)


def test_planner_direct_transform():
    """Widths with a single-batch fused shape take the direct length-cols FFT by default;
    Bluestein-only options and other widths keep the Bluestein row pass."""
    import cuspfft_amd as sp
    d = sp.direct_lengths()

    def smooth7(n):
        for q in (2, 3, 5, 7):
            while n % q == 0:
                n //= q
        return n == 1
    assert {m for m in d if smooth7(m)} <= set(sp.fused_lengths())
    assert {1024, 2048, 4096, 8192, 16384} <= set(d)
    for cols in (1024, 4704, 16384, 18432):
        info = sp.describe_plan(300, cols)
        assert info["direct"] == 1 and info["fft_len"] == cols and info["row_kernel"] == sp.ROW_FUSED
        b = sp.describe_plan(300, cols, row_transform=sp.TRANSFORM_BLUESTEIN)
        assert b["direct"] == 0 and b["fft_len"] >= 2 * cols - 1
    for kw in (dict(precision=sp.PREC_BF16X3), dict(row_kernel=sp.ROW_FOUR_STEP), dict(fft_len=4096)):
        assert sp.describe_plan(300, 2048, **kw)["direct"] == 0
    for cols in (2310, 16706, 20736, 1000):  # no single-batch shape of that length
        assert sp.describe_plan(300, cols)["direct"] == 0
        with pytest.raises(sp.SpfftError):
            sp.describe_plan(300, cols, row_transform=sp.TRANSFORM_DIRECT)
    with pytest.raises(sp.SpfftError):
        sp.describe_plan(300, 2048, row_transform=5)


def test_planner_split_transform():
    """Widths without an on-chip row pass (no direct, Rader or fused Bluestein shape) split
    as cols = n1 * n2 with the smallest n1 that has a column kernel and an on-chip width-n2
    row pass; TRANSFORM_SPLIT forces it where it exists; the four-step options, the
    precision form or an explicit length keep the four-step kernels."""
    import cuspfft_amd as sp
    c5 = sp.describe_plan(40421, 40421)  # 83 * 487: Bluestein M 82944 has no on-chip shape
    assert c5["row_transform"] == sp.TRANSFORM_SPLIT
    # 487 is an odd prime with a run-kernel shape for 486: Rader's length 486
    assert (c5["m1"], c5["m2"], c5["fft_len"], c5["row_threads"]) == (83, 487, 486, 256)
    r = sp.describe_plan(41, 3409, row_transform=sp.TRANSFORM_SPLIT)
    assert (r["m1"], r["m2"], r["fft_len"]) == (7, 487, 486)
    assert c5["row_kernel"] == sp.ROW_FUSED and c5["t_rows"] == 0
    f = sp.describe_plan(41, 7217, row_transform=sp.TRANSFORM_SPLIT)
    assert (f["m1"], f["m2"], f["fft_len"]) == (7, 1031, 2304)  # 1030 = 2 * 5 * 103: Bluestein
    d = sp.describe_plan(64, 3072, row_transform=sp.TRANSFORM_SPLIT)
    assert (d["m1"], d["m2"], d["fft_len"]) == (2, 1536, 1536)  # direct width-1536 pass
    assert sp.describe_plan(41, 7217)["row_transform"] == sp.TRANSFORM_BLUESTEIN  # fused on chip
    for kw in (dict(row_kernel=sp.ROW_FOUR_STEP), dict(precision=sp.PREC_BF16X3), dict(fft_len=82944),
               dict(t_rows=1024)):
        info = sp.describe_plan(40421, 40421, **kw)
        assert info["row_transform"] == sp.TRANSFORM_BLUESTEIN and info["row_kernel"] == sp.ROW_FOUR_STEP
    for cols in (40427, 7919):  # primes: no split
        with pytest.raises(sp.SpfftError):
            sp.describe_plan(30, cols, row_transform=sp.TRANSFORM_SPLIT)
    assert sp.describe_plan(30, 40427)["row_transform"] == sp.TRANSFORM_BLUESTEIN


def test_planner_split_n1_option():
    """With TRANSFORM_SPLIT, m1 names the split's n1: it must divide cols, have a column
    kernel and leave an on-chip width-n2 pass; m1 = 0 keeps the smallest such n1."""
    import cuspfft_amd as sp
    S = sp.TRANSFORM_SPLIT
    i = sp.describe_plan(300, 9490, row_transform=S, m1=10)
    assert (i["m1"], i["m2"]) == (10, 949)
    assert sp.describe_plan(300, 9490, row_transform=S)["m1"] == 2  # the smallest n1
    i = sp.describe_plan(300, 9490, row_transform=S, m1=5)
    assert (i["m1"], i["m2"]) == (5, 1898)
    for bad in (3, 7, 949):  # 3, 7 do not divide 9490; 949 has no column kernel
        with pytest.raises(sp.SpfftError):
            sp.describe_plan(300, 9490, row_transform=S, m1=bad)
    i = sp.describe_plan(100, 40421, row_transform=S, m1=83)
    assert (i["m1"], i["m2"], i["fft_len"]) == (83, 487, 486)


CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cuspfft_amd", "csrc")


def _load_generator():
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen_rader_shapes", os.path.join(CSRC, "gen_rader_shapes.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _smooth7(n):
    for d in (2, 3, 5, 7):
        while n % d == 0:
            n //= d
    return n == 1


def _prime(n):
    return n > 2 and all(n % d for d in range(2, int(n ** 0.5) + 1))


def test_planner_rader_transform():
    """Widths 2p and odd prime widths p (p with a shape for p - 1) take Rader's
    convolutions of length p - 1 by default; Bluestein when asked, or when an option
    rules the on-chip kernels out; other widths reject TRANSFORM_RADER.  The widths are
    the hand-written shapes' 2p with p - 1 = 29 m (1858, 4178, 16706) and every prime p
    with 7-smooth p - 1 in the generated range (widths p and 2p from 1024 to 16384)."""
    import cuspfft_amd as sp
    widths = sp.rader_widths()
    assert widths == sorted(widths) and 16706 in widths
    assert {1858, 4178, 16706} <= set(widths)
    # the generated shapes: rader_shapes.h is what gen_rader_shapes.py writes, and the
    # library has exactly its widths besides the hand-written ones, all of them primes p
    # or 2p with 7-smooth p - 1 in the generated range (the others of the range have no
    # M1 <= 32 made of whole prime-power blocks of p - 1 within the thread and register
    # limits, e.g. 3137: 3136 = 2^6 7^2)
    gen = _load_generator()
    with open(os.path.join(CSRC, "rader_shapes.h")) as f:
        assert f.read() == gen.render()
    gw = {w for w, _ in gen.shapes()}
    assert gw == set(widths) - {1858, 4178, 16706} and len(gw) >= 100
    want = {w for w in range(1024, 16385)
            if (_prime(w) and _smooth7(w - 1)) or (w % 2 == 0 and _prime(w // 2) and _smooth7(w // 2 - 1))}
    assert gw <= want and 3137 in want - gw
    odd = [w for w in widths if w % 2]
    assert len(odd) >= 50 and 12097 in odd
    for cols in widths:
        ns = 1 if cols % 2 else 2
        p = cols // ns
        assert _prime(p)
        for rows in (7, 64, cols):
            info = sp.describe_plan(rows, cols)
            assert info["rader"] == 1 and info["direct"] == 0 and info["fft_len"] == p - 1
            assert info["m1"] * info["m2"] == p - 1 and info["row_kernel"] == sp.ROW_FUSED
            assert sp.describe_plan(rows, cols, row_transform=sp.TRANSFORM_RADER)["rader"] == 1
            b = sp.describe_plan(rows, cols, row_transform=sp.TRANSFORM_BLUESTEIN)
            assert b["rader"] == 0 and b["fft_len"] >= 2 * cols - 1
        # the precision form has Rader shapes of its own (k_rowrader_mfma, M1 = 29) at the
        # hand-written widths only
        if cols in (1858, 4178, 16706):
            mf = sp.describe_plan(64, cols, precision=sp.PREC_BF16X3)
            assert mf["rader"] == 1 and mf["precision"] == 1 and mf["m1"] == 29 and mf["fft_len"] == p - 1
        else:
            assert sp.describe_plan(64, cols, precision=sp.PREC_BF16X3)["rader"] == 0
        for kw in (dict(row_kernel=sp.ROW_FOUR_STEP),
                   dict(fft_len=sp.next_3_smooth(2 * cols - 1)), dict(conv_rows=4)):
            assert sp.describe_plan(64, cols, **kw)["rader"] == 0
            with pytest.raises(sp.SpfftError):
                sp.describe_plan(64, cols, row_transform=sp.TRANSFORM_RADER, **kw)
    # prime halves or primes without a shape (8362 = 2 * 37 * 113; c2's 8219: 8218 =
    # 2 * 7 * 587), odd composite widths, smooth widths
    for cols in (2 * 8363, 8219, 16707, 2310, 2048, 6):
        assert sp.describe_plan(64, cols)["rader"] == 0
        with pytest.raises(sp.SpfftError):
            sp.describe_plan(64, cols, row_transform=sp.TRANSFORM_RADER)


def test_planner_fused_row_kernel():
    """The planner picks the fused row kernel for the lengths it has a shape for
    (the configurations' 3- and 7-smooth lengths), the four-step kernels
    otherwise or when asked, and rejects impossible requests."""
    import cuspfft_amd as sp
    for cols, M in [(2310, 5184), (2310, 4704), (8219, 17496), (8219, 16464), (16706, 34992),
                    (16706, 33600)]:
        rule = 3 if sp.next_3_smooth(2 * cols - 1) == M else 7
        info = sp.describe_plan(cols, cols, length_rule=rule, row_transform=sp.TRANSFORM_BLUESTEIN)
        assert info["fft_len"] == M and info["row_kernel"] == sp.ROW_FUSED
        assert info["m1"] * info["m2"] == M and int(np.prod(info["conv_radix"])) == info["m2"]
        four = sp.describe_plan(cols, cols, length_rule=rule, row_kernel=sp.ROW_FOUR_STEP)
        assert four["row_kernel"] == sp.ROW_FOUR_STEP and four["fft_len"] == M
    # no shape for this length, the BF16x3 form, explicit convolution rows -> four-step
    assert sp.describe_plan(100, 1000, fft_len=2000)["row_kernel"] == sp.ROW_FOUR_STEP
    assert sp.describe_plan(2310, 2310, precision=sp.PREC_BF16X3)["row_kernel"] == sp.ROW_FOUR_STEP
    assert sp.describe_plan(2310, 2310, conv_rows=2)["row_kernel"] == sp.ROW_FOUR_STEP
    with pytest.raises(sp.SpfftError) as ei:
        sp.describe_plan(100, 1000, fft_len=2000, row_kernel=sp.ROW_FUSED)
    assert ei.value.status == 4
    with pytest.raises(sp.SpfftError) as ei:
        sp.describe_plan(100, 2310, row_kernel=sp.ROW_FUSED, precision=sp.PREC_BF16X3)
    assert ei.value.status == 4
    with pytest.raises(sp.SpfftError) as ei:
        sp.describe_plan(100, 100, row_kernel=5)
    assert ei.value.status == 1


def test_fused_length_ladder():
    """Every 3-smooth length from 1024 to 36864 except the powers of 3 has a fused
    row-kernel shape, so the default (3-smooth) plan of any matrix with
    512 < cols <= 18432 runs the fused kernel; the GPU tests cover exactly the
    library's list."""
    import cuspfft_amd as sp
    from test_gpu_path import FUSED_LENGTHS
    have = sp.fused_lengths()
    assert sorted(have) == sorted(FUSED_LENGTHS)
    smooth3 = sorted({2 ** a * 3 ** b for a in range(16) for b in range(10)
                      if 1024 <= 2 ** a * 3 ** b <= 36864})
    missing = [m for m in smooth3 if m not in have]
    assert missing == [2187, 6561, 19683]
    for cols in (513, 700, 1500, 4000, 9000, 12000, 16000, 18432):
        info = sp.describe_plan(cols, cols)
        assert info["fft_len"] in have and info["row_kernel"] == sp.ROW_FUSED


def test_planner_errors():
    import cuspfft_amd as sp
    with pytest.raises(sp.SpfftError) as e:
        sp.describe_plan(10, 10, u_tile=-1)
    assert "CSC tile size must be positive" in str(e.value)
    with pytest.raises(sp.SpfftError):
        sp.describe_plan(10, 10, fft_len=11)  # not 7-smooth
    with pytest.raises(sp.SpfftError):
        sp.describe_plan(10, 10, fft_len=18)  # < 2*cols - 1
    with pytest.raises(sp.SpfftError):
        sp.describe_plan(10, 10, fft_len=20, m1=3)  # m1 does not divide M
    with pytest.raises(sp.SpfftError):
        sp.describe_plan(0, 10)


@pytest.mark.parametrize("name", ["c1", "c2", "c3"])
def test_generators_deterministic_symmetric(name):
    import cuspfft_amd as sp
    a = sp.config_matrix(name)
    b = sp.config_matrix(name)
    np.testing.assert_array_equal(a.row_idx, b.row_idx)
    np.testing.assert_array_equal(a.col_idx, b.col_idx)
    assert not np.any(a.row_idx == a.col_idx)  # zero diagonal
    key = a.row_idx.astype(np.int64) * a.cols + a.col_idx
    assert np.all(np.diff(key) > 0)  # canonical: sorted by (row, col), no duplicates
    tkey = np.sort(a.col_idx.astype(np.int64) * a.cols + a.row_idx)
    np.testing.assert_array_equal(key, tkey)  # symmetric
    c = sp.config_matrix(name, seed=1)
    assert len(c.row_idx) != len(a.row_idx) or np.any(c.row_idx != a.row_idx)


def test_generator_sizes():
    import cuspfft_amd as sp
    assert sp.config_matrix("c1").nnz == 16000
    assert abs(sp.config_matrix("c2").nnz / 242700 - 1) < 0.01
    m3 = sp.config_matrix("c3")
    assert abs(m3.nnz / 121250 - 1) < 0.02
    assert np.bincount(m3.col_idx).max() <= 400


def test_read_mtx(tmp_path):
    import cuspfft_amd as sp
    p = tmp_path / "m.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n% comment\n4 4 4\n"
                 "2 1\n3 3\n4 2\n2 1\n")
    m = sp.read_mtx(str(p))
    assert (m.rows, m.cols) == (4, 4)
    got = list(zip(m.row_idx.tolist(), m.col_idx.tolist()))
    assert got == [(0, 1), (1, 0), (1, 3), (2, 2), (3, 1)]
    q = tmp_path / "g.mtx"
    q.write_text("%%MatrixMarket matrix coordinate real general\n3 5 2\n3 5 7.5\n1 1 -2\n")
    m = sp.read_mtx(str(q))
    assert list(zip(m.row_idx.tolist(), m.col_idx.tolist())) == [(0, 0), (2, 4)]
    with pytest.raises(RuntimeError):
        sp.read_mtx(str(tmp_path / "missing.mtx"))


def test_codelets_on_host(tmp_path):
    """Every register codelet RegFft<L> (L 7-smooth <= 64) against a direct FP64 DFT,
    compiled with hipcc and run on the host."""
    exe = tmp_path / "codelet_check"
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-I",
                    os.path.join(ROOT, "cuspfft_amd", "csrc"), "--offload-arch=gfx950",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    os.path.join(ROOT, "tests", "cpp", "codelet_check.hip"), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert out.stdout.strip().endswith("OK")


def _describe_raw(rows, cols, opts, info):
    import ctypes
    import cuspfft_amd as sp
    return sp.load().spfft_plan_describe(rows, cols, ctypes.byref(opts), ctypes.byref(info))


def test_struct_size_mismatch_is_rejected():
    """A caller built against another layout of spfft_opts / spfft_plan_info / spfft_result
    (struct_size differs from the library's sizeof) gets SPFFT_ERR_INVALID_ARG, never a
    read or write past the end of its struct."""
    import ctypes
    import cuspfft_amd as sp
    L = sp.load()
    o = sp.Opts()
    L.spfft_opts_default(ctypes.byref(o))
    assert o.struct_size == ctypes.sizeof(sp.Opts)
    info = sp.PlanInfo()
    assert _describe_raw(100, 100, o, info) == 0
    assert info.struct_size == ctypes.sizeof(sp.PlanInfo) and info.fft_len >= 199

    # an older, shorter spfft_opts: the fields up to row_transform, without row_persist
    class ShortOpts(ctypes.Structure):
        _fields_ = sp.Opts._fields_[:-1]
    so = ShortOpts()
    so.struct_size = ctypes.sizeof(ShortOpts)
    so.length_rule, so.u_tile = 3, 0
    rc = L.spfft_plan_describe(100, 100, ctypes.cast(ctypes.byref(so), ctypes.POINTER(sp.Opts)),
                               ctypes.byref(info))
    assert rc == 1 and "struct_size" in L.spfft_last_error().decode()
    o.struct_size = 0
    assert _describe_raw(100, 100, o, info) == 1
    L.spfft_opts_default(ctypes.byref(o))
    bad = sp.PlanInfo()
    bad.struct_size = ctypes.sizeof(sp.PlanInfo) - 4
    assert _describe_raw(100, 100, o, bad) == 1
    assert "spfft_plan_info" in L.spfft_last_error().decode()
    # spfft_result: checked before any device work (no GPU needed for the rejection)
    res = sp.Result()
    res.struct_size = 8
    r = np.zeros(1, np.int32)
    rc = L.spfft_csc_bluestein_mixed_radix(4, 4, 1, r.ctypes.data, r.ctypes.data, 128, ctypes.byref(res))
    assert rc == 1 and "spfft_result" in L.spfft_last_error().decode()


def test_reference_opts_run_bluestein_at_the_reference_length():
    """The reference-shaped entry points plan the reference's algorithm: Bluestein at
    next_3_smooth(2 cols - 1) (mixed radix) or next_7_smooth (cuFFT smooth), also for
    widths where the native plan takes the Rader (c3) or direct (p14) transform."""
    import ctypes
    import cuspfft_amd as sp
    L = sp.load()
    for cols, native in ((16706, "rader"), (16384, "direct"), (2310, None), (8219, None)):
        for rule, nxt in ((3, sp.next_3_smooth), (7, sp.next_7_smooth)):
            o = sp.Opts()
            L.spfft_reference_opts(rule, 128, ctypes.byref(o))
            assert o.struct_size == ctypes.sizeof(sp.Opts) and o.u_tile == 128
            info = sp.PlanInfo()
            assert _describe_raw(cols, cols, o, info) == 0
            assert info.fft_len == nxt(2 * cols - 1), (cols, rule, info.fft_len)
            assert info.row_transform == sp.TRANSFORM_BLUESTEIN and not info.rader and not info.direct
            d = sp.describe_plan(cols, cols, **sp.reference_opts(rule, 128))
            assert d["fft_len"] == nxt(2 * cols - 1)
        auto = sp.describe_plan(cols, cols)
        if native:
            assert auto[native] == 1


def test_automatic_tile():
    """u_tile 0: one tile of all R/2 + 1 computed rows (the column pass then pairs rows
    for even R).  t_rows: rows of the four-step tile buffer T per sub-tile, 0 = the tile."""
    import cuspfft_amd as sp
    for n in (2310, 8219, 16706, 16384, 40421):
        info = sp.describe_plan(n, n)
        assert info["u_tile"] == n // 2 + 1
    assert sp.describe_plan(16706, 16706)["colpass_paired"] == 1
    assert sp.describe_plan(16706, 16706)["t_rows"] == 0  # on-chip row kernel: no T
    c5 = sp.describe_plan(40421, 40421, row_transform=sp.TRANSFORM_BLUESTEIN)
    assert c5["row_kernel"] == sp.ROW_FOUR_STEP and c5["fft_len"] == 82944
    assert c5["t_rows"] == c5["u_tile"] == 20211
    assert sp.describe_plan(40421, 40421, t_rows=7)["t_rows"] == 7
    assert sp.describe_plan(40421, 40421, u_tile=5, t_rows=7)["t_rows"] == 5
    with pytest.raises(sp.SpfftError):
        sp.describe_plan(40421, 40421, t_rows=-1)


def _read_spfftref(path):
    raw = open(path, "rb").read()
    assert raw[:8] == b"SPFFTREF"
    R, C = (int(x) for x in np.frombuffer(raw[8:16], np.int32))
    X = np.frombuffer(raw[16:], np.float64).reshape(R, C // 2 + 1, 2)
    return X[..., 0] + 1j * X[..., 1]


@pytest.mark.parametrize("rows,cols,nnz", [(2310, 2310, 0), (97, 131, 900), (64, 60, 500), (1, 5, 3),
                                           (7, 1, 4), (33, 2 * 8353, 400), (128, 1024, 2000)])
def test_cli_cpu_reference(tmp_path, rows, cols, nnz):
    """The CLI's --cpu-reference transform (host FP32 FFT, no GPU): mixed radix (c1's
    2310 = 2 3 5 7 11, 64 x 60, powers of two), Bluestein for prime factors above 31
    (97 x 131, c3's width 2 x 8353), 1-row and 1-column matrices; its spectrum, written
    with --write-ref, against numpy.fft.rfft2 of the dense matrix in FP64."""
    import cuspfft_amd as sp
    cli = os.path.join(os.path.dirname(sp.LIB_PATH), "spfft_cli")
    if nnz == 0:
        m = sp.config_matrix("c1")
        spec, r, c = "gen:c1", m.row_idx, m.col_idx
    else:
        rng = np.random.default_rng(rows * 1000 + cols)
        idx = rng.choice(rows * cols, min(nnz, rows * cols), replace=False)
        r, c = idx // cols, idx % cols
        spec = os.path.join(str(tmp_path), "m.mtx")
        with open(spec, "w") as f:
            f.write("%%MatrixMarket matrix coordinate pattern general\n")
            f.write(f"{rows} {cols} {len(idx)}\n")
            for a, b in zip(r, c):
                f.write(f"{a + 1} {b + 1}\n")
    out = os.path.join(str(tmp_path), "ref.bin")
    res = subprocess.run([cli, spec, "--cpu-reference", "--write-ref", out], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "CPU reference (FP32 FFT)" in res.stdout and "FAIL" not in res.stdout
    X = _read_spfftref(out)
    d = np.zeros((rows, cols))
    d[r, c] = 1.0
    ref = np.fft.rfft2(d)
    assert X.shape == ref.shape
    err = np.linalg.norm(X - ref) / np.linalg.norm(ref)
    assert err < 1e-6, err
    if nnz == 0:
        bad = subprocess.run([cli, spec, "--write-ref", out], capture_output=True, text=True, timeout=60)
        assert bad.returncode == 1 and "--write-ref needs --cpu-reference" in bad.stderr
