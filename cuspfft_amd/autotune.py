"""Autotuner for the per-matrix plan (Bluestein length M, its split M = M1 * M2,
row-pass kernel (four-step or fused), arithmetic form, column-pass rows per lane).

For each matrix it times every candidate plan on the GPU with the library's own
HIP-event stage timing (csc_build, colpass, fft1, conv, ifft1_finalize), keeps
the fastest, and records all timings in a JSON-lines timing DB.  The DB is the
training set of the learned plan selector (cuspfft_amd.selector); exhaustive
tuning stays its fallback and the judge of its regret.  Every candidate meets
the accuracy bar (FP32 codelets ~2e-7, BF16x3 MFMA ~4e-6 relative L2 vs FP64),
so the choice never changes what is computed beyond that bound.
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

from . import (Plan, SpfftError, describe_plan, fused_lengths, rader_widths, PREC_BF16X3, ROW_FOUR_STEP,
               ROW_FUSED, TRANSFORM_BLUESTEIN, TRANSFORM_DIRECT, TRANSFORM_RADER, TRANSFORM_SPLIT)

# M1 lengths tried as alternatives to the planner's choice (register codelets
# with a fast convolution kernel for M / M1).
ALT_M1 = (16, 18, 24, 27, 32, 36, 40, 48, 54, 64, 20, 25, 28, 42, 45, 49, 56, 60, 63, 12, 8, 9)
# column-pass rows per lane; 32: the matrix-core kernel (1024-row blocks)
COLPASS_ROWS = (1, 2, 4, 8, 32)
# threads per workgroup tried for the on-chip row kernels that have more than one shape
ROW_THREADS = (64, 192, 576, 704)
# tile candidates of full-size tuning: rows of T per sub-tile (four-step kernels) and
# rows per u-tile (on-chip row kernels; several tiles give up the paired column pass)
T_ROWS = (1024, 4096)
U_TILES = (1024, 4096)


def plan_key(opts: dict) -> str:
    key = "r{length_rule}_M{fft_len}_p{precision}_m{m1}_k{row_kernel}_c{colpass_rows}".format(
        length_rule=opts.get("length_rule", 3), fft_len=opts.get("fft_len", 0),
        precision=opts.get("precision", 0), m1=opts.get("m1", 0),
        row_kernel=opts.get("row_kernel", 0), colpass_rows=opts.get("colpass_rows", 0))
    for k, tag in (("row_transform", "x"), ("row_threads", "n"), ("t_rows", "t"), ("u_tile", "u")):
        if opts.get(k):
            key += f"_{tag}{opts[k]}"
    return key


def candidates(rows: int, cols: int, max_m1: int = 3) -> list[dict]:
    """Row-pass structure candidates: for both length rules, the four-step
    kernels with the planner's M1 and up to max_m1-1 alternatives with a fast
    convolution kernel, the BF16x3 MFMA form, and the fused row kernel at the
    rule's M when it has a shape; plus the fused row kernel at the fused lengths
    M in [2*cols-1, 1.12 (2*cols-1)] (at most four; else the smallest above)."""
    out = []
    # Bluestein candidates say so where the width has another transform (AUTO would take
    # it: the direct or Rader transform, or for a length rule without an on-chip shape the
    # split transform)
    autos = [describe_plan(rows, cols, length_rule=rule) for rule in (3, 7)]
    other = any(a["direct"] or a["rader"] or a["row_transform"] == TRANSFORM_SPLIT for a in autos)
    blue = dict(row_transform=TRANSFORM_BLUESTEIN) if other else {}
    for rule in (3, 7):
        base = describe_plan(rows, cols, length_rule=rule, row_kernel=ROW_FOUR_STEP, **blue)
        M = base["fft_len"]
        m1s = [base["m1"]]
        for d in ALT_M1:
            if len(m1s) >= max_m1:
                break
            if d in m1s or M % d:
                continue
            try:
                info = describe_plan(rows, cols, fft_len=M, m1=d, row_kernel=ROW_FOUR_STEP, **blue)
            except SpfftError:
                continue
            if info["conv_fast"]:
                m1s.append(d)
        for m1 in m1s:
            # M pinned: the rule's length may lack this split (the planner then moves on)
            out.append(dict(length_rule=rule, fft_len=M, m1=m1, precision=0, row_kernel=ROW_FOUR_STEP, **blue))
        out.append(dict(length_rule=rule, m1=0, precision=PREC_BF16X3, row_kernel=ROW_FOUR_STEP, **blue))
        if describe_plan(rows, cols, length_rule=rule, **blue)["row_kernel"] == ROW_FUSED:
            out.append(dict(length_rule=rule, m1=0, precision=0, row_kernel=ROW_FUSED, **blue))
    # the fused row kernel at the fused lengths up to 12% above 2*cols-1: a longer length
    # with cheaper (power-of-two) stage radices can be faster (c3: 36864 beats 34992)
    fl = [M for M in fused_lengths() if 2 * cols - 1 <= M <= 1.12 * (2 * cols - 1)]
    for M in (fl or [M for M in fused_lengths() if M >= 2 * cols - 1])[:4]:
        out.append(dict(length_rule=3, fft_len=M, m1=0, precision=0, row_kernel=ROW_FUSED, **blue))
    # the transforms other than Bluestein where the width has one: the direct length-cols
    # FFT, and Rader's two length-(p - 1) convolutions (cols = 2p) in both arithmetic forms
    # and every workgroup size with a shape
    for xf in (TRANSFORM_DIRECT, TRANSFORM_RADER):
        for prec in (0, PREC_BF16X3):
            for nt in (0,) + ROW_THREADS:
                try:
                    describe_plan(rows, cols, row_transform=xf, precision=prec, row_threads=nt)
                except SpfftError:
                    continue
                out.append(dict(length_rule=3, m1=0, precision=prec, row_transform=xf, row_threads=nt))
    # identical plans (e.g. both rules give the same M) are tried once
    seen, uniq = set(), []
    for o in out:
        info = describe_plan(rows, cols, **o)
        sig = (info["fft_len"], info["m1"], info["precision"], info["row_kernel"], info["direct"],
               info["rader"], info["row_threads"])
        if sig not in seen:
            seen.add(sig)
            if not (info["direct"] or info["rader"]):
                o = dict(o, m1=info["m1"])
            if not o.get("row_threads"):
                o.pop("row_threads", None)
            uniq.append(o)
    return uniq


def tile_candidates(rows: int, opts: dict, info: dict) -> list[dict]:
    """Tile variants of a plan for full-size tuning: sub-tiles of the tile buffer T
    (four-step kernels) or several u-tiles (on-chip row kernels), each below the
    R/2 + 1 computed rows."""
    full = rows // 2 + 1
    if info["row_kernel"] == ROW_FOUR_STEP:
        return [dict(opts, t_rows=t) for t in T_ROWS if t < full]
    return [dict(opts, u_tile=t) for t in U_TILES if t < full]


def time_plan(rows: int, cols: int, d_row, d_col, opts: dict, reps: int = 3, row_limit: int = 0):
    """GPU time per transform (ms) of each stage, from HIP events on the launch
    stream; one untimed warm-up first.  With row_limit > 0 only the first
    row_limit output rows are run (tile-sampled timing) and the row-pass stages
    are scaled to the full R/2 + 1 computed rows."""
    import torch
    full = rows // 2 + 1
    lim = row_limit if 0 < row_limit < full else 0
    o = dict(opts)
    u_tile = o.pop("u_tile", 0)
    plan = Plan(rows, cols, u_tile=lim or u_tile or rows, row_limit=lim, **o)
    out = plan.alloc_output()
    # untimed runs until about 20 ms of GPU work (at least one, at most 50): after an
    # idle spell the clocks take that long to come back up
    t0 = time.perf_counter()
    for _ in range(50):
        plan.execute(d_row, d_col, out)
        torch.cuda.synchronize()
        if time.perf_counter() - t0 > 0.02:
            break
    plan.profile(True)
    for _ in range(reps):
        plan.execute(d_row, d_col, out)
    prof = plan.profile_read()
    plan.profile(False)
    plan.check_input()
    scale = full / lim if lim else 1.0
    stages = {k: v[0] / reps * (1.0 if k == "csc_build" else scale) for k, v in prof.items()}
    return sum(stages.values()), stages, plan.info


def tune_matrix(m, reps: int = 3, max_m1: int = 3, record: dict | None = None, row_limit: int = 0,
                tiles: bool | None = None):
    """Exhaustive tuning of one matrix.  The planner's default plan is timed
    first and flagged.  With tiles (default: when every row is timed, row_limit
    absent or above R/2 + 1) the best plan's tile variants are timed last.
    Returns (best_opts, best_ms, trials)."""
    import torch
    d_row = torch.as_tensor(np.ascontiguousarray(m.row_idx, np.int32)).cuda()
    d_col = torch.as_tensor(np.ascontiguousarray(m.col_idx, np.int32)).cuda()
    trials = []

    def trial(o, default=False):
        ms, st, info = time_plan(m.rows, m.cols, d_row, d_col, o, reps, row_limit)
        o = dict(o, row_kernel=info["row_kernel"])  # the kernel the plan resolved to
        t = dict(opts=o, ms=ms, stages=st, fft_len=info["fft_len"], m1=info["m1"], m2=info["m2"],
                 direct=info["direct"], rader=info["rader"], row_threads=info["row_threads"],
                 split=int(info["row_transform"] == TRANSFORM_SPLIT))
        if default:
            t["default"] = True
        trials.append(t)
        return t

    trial(dict(length_rule=3), default=True)
    for o in candidates(m.rows, m.cols, max_m1):
        trial(o)
    best = min(trials, key=lambda t: t["ms"])
    # column-pass rows per lane only changes the column-pass kernel
    base_opts = {k: v for k, v in best["opts"].items() if k != "colpass_rows"}
    for r in COLPASS_ROWS:
        trial(dict(base_opts, colpass_rows=r))
    best = min(trials, key=lambda t: t["ms"])
    if tiles is None:
        tiles = not (0 < row_limit < m.rows // 2 + 1)
    if tiles:
        info = describe_plan(m.rows, m.cols, **best["opts"])
        for o in tile_candidates(m.rows, best["opts"], info):
            trial(o)
        best = min(trials, key=lambda t: t["ms"])
    if record is not None:
        record.update(features_of(m), trials=trials, best=best["opts"], best_ms=best["ms"])
    return best["opts"], best["ms"], trials


def features_of(m) -> dict:
    """Matrix features used by the learned selector: shape, nnz, and the
    column-degree histogram in 8 log2 bins (degree 1, 2-3, 4-7, ..., >= 128)."""
    deg = np.bincount(np.asarray(m.col_idx), minlength=m.cols)
    deg = deg[deg > 0]
    hist = np.zeros(8)
    if len(deg):
        b = np.minimum(np.floor(np.log2(deg)).astype(int), 7)
        hist = np.bincount(b, minlength=8)[:8] / len(deg)
    return dict(rows=int(m.rows), cols=int(m.cols), nnz=int(m.nnz),
                deg_hist=[float(x) for x in hist], active_cols=int(len(deg)))


def append_db(path: str, record: dict):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "a") as f:
        f.write(json.dumps(record) + "\n")


def load_db(path: str) -> list[dict]:
    if not os.path.exists(path):
        return []
    with open(path) as f:
        return [json.loads(line) for line in f if line.strip()]


def full_run_estimate_ms(rows: int, cols: int, nnz: int) -> float:
    """Rough GPU time of one transform: column pass ~ (R/2 + 1) nnz table sums at
    2.9e12/s, row pass ~ (R/2 + 1) M log2 M at 1.5e12/s with M = 2 cols."""
    r = rows // 2 + 1
    M = 2 * cols
    return 1e3 * (r * nnz / 2.9e12 + r * M * math.log2(M) / 1.5e12)


def tune_many(mats, db_path: str | None = None, reps: int = 2, max_m1: int = 2, log=None,
              row_limit: int = 0, tags=None, sample_above_ms: float = 0.0):
    """Tunes a list of matrices; returns the list of best opts (and appends to the DB).
    With sample_above_ms > 0, row_limit (tile-sampled timing) applies only to matrices
    whose full transform is estimated above that many ms; the others are timed in
    full (one tile with the paired column pass, as they run)."""
    best = []
    t0 = time.time()
    for i, m in enumerate(mats):
        rec = {} if tags is None else dict(tags[i])
        lim = row_limit
        if sample_above_ms > 0 and full_run_estimate_ms(m.rows, m.cols, m.nnz) <= sample_above_ms:
            lim = 0
        if lim:
            rec["row_limit"] = lim
        o, ms, _ = tune_matrix(m, reps=reps, max_m1=max_m1, record=rec, row_limit=lim)
        best.append(o)
        if db_path:
            append_db(db_path, rec)
        if log and (i % 32 == 0 or i == len(mats) - 1):
            log(f"[tune] {i + 1}/{len(mats)} n={m.rows} nnz={m.nnz} best={plan_key(o)} "
                f"{ms:.3f} ms ({time.time() - t0:.0f} s)")
    return best


def batch_matrices(count: int = 512, seed0: int = 0, n_lo: int = 500, n_hi: int = 12000):
    """The seeded batch workload: n log-uniform in [n_lo, n_hi], average degree
    uniform in [4, 30]; the first half uniform random pairs, the second half
    Chung-Lu with degree exponent uniform in [2.0, 2.5]; matrix i uses seed seed0 + i."""
    from . import generate
    mats = []
    half = count // 2
    for i in range(count):
        rng = np.random.default_rng(seed0 + i)
        n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        deg = float(rng.uniform(4.0, 30.0))
        if i < half:
            mats.append(generate("uniform", n, n * deg / 2.0, seed=seed0 + i))
        else:
            mats.append(generate("chunglu", n, n * deg, float(rng.uniform(2.0, 2.5)), 0.0,
                                 seed=seed0 + i))
    return mats


def training_specs(count: int = 2000, seed0: int = 100000, n_lo: int = 256, n_hi: int = 40000,
                   d_lo: float = 1e-5, d_hi: float = 1e-2, nnz_cap: int = 16_000_000,
                   per_rader_width: int = 0) -> list[dict]:
    """The selector's seeded training set, as specs for make_matrix: n log-uniform in
    [n_lo, n_hi], density log-uniform in [d_lo, d_hi] (stored nonzeros = density * n^2,
    at least 16, at most n (n - 1) / 2 and nnz_cap; the default cap is above the
    set's largest product, 1e-2 * 40000^2), even seeds uniform random pairs, odd seeds
    Chung-Lu with degree exponent uniform in [2.0, 2.5]; then per_rader_width matrices
    at each width with a Rader transform (n = that width, density drawn the same way),
    so that the network sees Rader's plans."""
    specs = []

    def spec(seed, i, n=None):
        rng = np.random.default_rng(seed)
        draw_n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        n = draw_n if n is None else n
        dens = math.exp(rng.uniform(math.log(d_lo), math.log(d_hi)))
        nnz = int(min(max(dens * n * n, 16), nnz_cap, n * (n - 1) * 0.5))
        if i % 2 == 0:
            return dict(seed=seed, kind="uniform", n=n, nnz=nnz, exponent=0.0)
        return dict(seed=seed, kind="chunglu", n=n, nnz=nnz, exponent=float(rng.uniform(2.0, 2.5)))

    for i in range(count):
        specs.append(spec(seed0 + i, i))
    k = count
    for w in sorted(rader_widths()):
        for j in range(per_rader_width):
            specs.append(spec(seed0 + k, j, n=w))
            k += 1
    return specs


def make_matrix(spec: dict):
    """The COOMatrix of a training spec (training_specs)."""
    from . import generate
    n, nnz, seed = spec["n"], spec["nnz"], spec["seed"]
    if spec["kind"] == "uniform":
        return generate("uniform", n, max(nnz // 2, 1), seed=seed)
    return generate("chunglu", n, max(nnz, 2), spec["exponent"], 0.0, seed=seed)


def training_matrices(count: int = 2000, seed0: int = 100000, n_lo: int = 256, n_hi: int = 40000,
                      d_lo: float = 1e-5, d_hi: float = 1e-2, nnz_cap: int = 16_000_000,
                      per_rader_width: int = 0):
    """The matrices of training_specs(...).  Returns (matrices, tags) with
    tags[i] = {"seed", "kind"}."""
    specs = training_specs(count, seed0, n_lo, n_hi, d_lo, d_hi, nnz_cap, per_rader_width)
    return [make_matrix(s) for s in specs], [{"seed": s["seed"], "kind": s["kind"]} for s in specs]
