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

from . import Plan, SpfftError, describe_plan, fused_lengths, PREC_BF16X3, ROW_FOUR_STEP, ROW_FUSED

# M1 lengths tried as alternatives to the planner's choice (register codelets
# with a fast convolution kernel for M / M1).
ALT_M1 = (16, 18, 24, 27, 32, 36, 40, 48, 54, 64, 20, 25, 28, 42, 45, 49, 56, 60, 63, 12, 8, 9)
COLPASS_ROWS = (1, 2, 4, 8)


def plan_key(opts: dict) -> str:
    return "r{length_rule}_M{fft_len}_p{precision}_m{m1}_k{row_kernel}_c{colpass_rows}".format(
        length_rule=opts.get("length_rule", 3), fft_len=opts.get("fft_len", 0),
        precision=opts.get("precision", 0), m1=opts.get("m1", 0),
        row_kernel=opts.get("row_kernel", 0), colpass_rows=opts.get("colpass_rows", 0))


def candidates(rows: int, cols: int, max_m1: int = 3) -> list[dict]:
    """Row-pass structure candidates: for both length rules, the four-step
    kernels with the planner's M1 and up to max_m1-1 alternatives with a fast
    convolution kernel, the BF16x3 MFMA form, and the fused row kernel at the
    rule's M when it has a shape; plus the fused row kernel at the fused lengths
    M in [2*cols-1, 1.12 (2*cols-1)] (at most four; else the smallest above)."""
    out = []
    for rule in (3, 7):
        base = describe_plan(rows, cols, length_rule=rule, row_kernel=ROW_FOUR_STEP)
        M = base["fft_len"]
        m1s = [base["m1"]]
        for d in ALT_M1:
            if len(m1s) >= max_m1:
                break
            if d in m1s or M % d:
                continue
            try:
                info = describe_plan(rows, cols, fft_len=M, m1=d, row_kernel=ROW_FOUR_STEP)
            except SpfftError:
                continue
            if info["conv_fast"]:
                m1s.append(d)
        for m1 in m1s:
            # M pinned: the rule's length may lack this split (the planner then moves on)
            out.append(dict(length_rule=rule, fft_len=M, m1=m1, precision=0, row_kernel=ROW_FOUR_STEP))
        out.append(dict(length_rule=rule, m1=0, precision=PREC_BF16X3, row_kernel=ROW_FOUR_STEP))
        if describe_plan(rows, cols, length_rule=rule)["row_kernel"] == ROW_FUSED:
            out.append(dict(length_rule=rule, m1=0, precision=0, row_kernel=ROW_FUSED))
    # the fused row kernel at the fused lengths up to 12% above 2*cols-1: a longer length
    # with cheaper (power-of-two) stage radices can be faster (c3: 36864 beats 34992)
    fl = [M for M in fused_lengths() if 2 * cols - 1 <= M <= 1.12 * (2 * cols - 1)]
    for M in (fl or [M for M in fused_lengths() if M >= 2 * cols - 1])[:4]:
        out.append(dict(length_rule=3, fft_len=M, m1=0, precision=0, row_kernel=ROW_FUSED))
    # identical plans (e.g. both rules give the same M) are tried once
    seen, uniq = set(), []
    for o in out:
        info = describe_plan(rows, cols, **o)
        sig = (info["fft_len"], info["m1"], info["precision"], info["row_kernel"])
        if sig not in seen:
            seen.add(sig)
            o = dict(o, m1=info["m1"])
            uniq.append(o)
    return uniq


def time_plan(rows: int, cols: int, d_row, d_col, opts: dict, reps: int = 3, row_limit: int = 0):
    """GPU time per transform (ms) of each stage, from HIP events on the launch
    stream; one untimed warm-up first.  With row_limit > 0 only the first
    row_limit output rows are run (tile-sampled timing) and the row-pass stages
    are scaled to the full R/2 + 1 computed rows."""
    import torch
    full = rows // 2 + 1
    lim = row_limit if 0 < row_limit < full else 0
    plan = Plan(rows, cols, u_tile=lim or rows, row_limit=lim, **opts)
    out = plan.alloc_output()
    plan.execute(d_row, d_col, out)
    torch.cuda.synchronize()
    plan.profile(True)
    for _ in range(reps):
        plan.execute(d_row, d_col, out)
    prof = plan.profile_read()
    plan.profile(False)
    plan.check_input()
    scale = full / lim if lim else 1.0
    stages = {k: v[0] / reps * (1.0 if k == "csc_build" else scale) for k, v in prof.items()}
    return sum(stages.values()), stages, plan.info


def tune_matrix(m, reps: int = 3, max_m1: int = 3, record: dict | None = None, row_limit: int = 0):
    """Exhaustive tuning of one matrix.  The planner's default plan is timed
    first and flagged.  Returns (best_opts, best_ms, trials)."""
    import torch
    d_row = torch.as_tensor(np.ascontiguousarray(m.row_idx, np.int32)).cuda()
    d_col = torch.as_tensor(np.ascontiguousarray(m.col_idx, np.int32)).cuda()
    trials = []

    def trial(o, default=False):
        ms, st, info = time_plan(m.rows, m.cols, d_row, d_col, o, reps, row_limit)
        o = dict(o, row_kernel=info["row_kernel"])  # the kernel the plan resolved to
        t = dict(opts=o, ms=ms, stages=st, fft_len=info["fft_len"], m1=info["m1"], m2=info["m2"],
                 direct=info["direct"], rader=info["rader"])
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


def tune_many(mats, db_path: str | None = None, reps: int = 2, max_m1: int = 2, log=None,
              row_limit: int = 0, tags=None):
    """Tunes a list of matrices; returns the list of best opts (and appends to the DB)."""
    best = []
    t0 = time.time()
    for i, m in enumerate(mats):
        rec = {} if tags is None else dict(tags[i])
        o, ms, _ = tune_matrix(m, reps=reps, max_m1=max_m1, record=rec, row_limit=row_limit)
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


def training_matrices(count: int = 2000, seed0: int = 100000, n_lo: int = 256, n_hi: int = 40000,
                      d_lo: float = 1e-5, d_hi: float = 1e-2, nnz_cap: int = 2_000_000):
    """The selector's seeded training set: n log-uniform in [n_lo, n_hi], density
    log-uniform in [d_lo, d_hi] (stored nonzeros = density * n^2, at least 16 and at
    most nnz_cap to bound generation and tuning time), even seeds uniform random
    pairs, odd seeds Chung-Lu with degree exponent uniform in [2.0, 2.5].
    Returns (matrices, tags) with tags[i] = {"seed", "kind"}."""
    from . import generate
    mats, tags = [], []
    for i in range(count):
        seed = seed0 + i
        rng = np.random.default_rng(seed)
        n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        dens = math.exp(rng.uniform(math.log(d_lo), math.log(d_hi)))
        nnz = int(min(max(dens * n * n, 16), nnz_cap, n * (n - 1) * 0.5))
        if i % 2 == 0:
            mats.append(generate("uniform", n, max(nnz // 2, 1), seed=seed))
            kind = "uniform"
        else:
            mats.append(generate("chunglu", n, max(nnz, 2), float(rng.uniform(2.0, 2.5)), 0.0, seed=seed))
            kind = "chunglu"
        tags.append({"seed": seed, "kind": kind})
    return mats, tags
