"""Trains and evaluates the learned plan selector.

    python -m cuspfft_amd.train_selector --dir OUT [--count 2000] [--heldout 256] [--epochs 300]
                                         [--per-rader-width 16]

1. Tunes the seeded training set (autotune.training_specs: COUNT random matrices
   plus PER_RADER_WIDTH (default 16) at each width with a Rader transform) exhaustively, timing each plan in
   full (tile-sampled above an estimated 16 ms per transform), appending to
   OUT/train.jsonl (resumes where it stopped).  Matrices are
   generated 8 at a time, as they are tuned.
2. Trains the MLP on the GPU with the library's HIP kernels; writes OUT/selector.json.
3. Tunes a fresh held-out batch (disjoint seeds, 4 per Rader width added) and
   reports the regret of the network's choice against exhaustive tuning, next to
   the regret of the planner's default plan (the method without the network).
4. The named cases c3 (16706^2, a Rader width) and c5 (40421^2): the network's plan
   against exhaustive tuning, timed in full, with the tile candidates.
Prints one JSON report on stdout; progress on stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time


# matrices whose transform is estimated above this many ms are timed on row_limit
# sampled rows (scaled); the others in full
SAMPLE_ABOVE_MS = 16.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def tune_set(specs, path, row_limit, reps, deadline=None):
    """Tunes the specs' matrices missing from the DB at `path`, generating them 8 at
    a time; stops at `deadline` (time.time() value) with the DB holding everything
    tuned so far."""
    from . import autotune
    have = {r.get("seed") for r in autotune.load_db(path)}
    todo = [sp for sp in specs if sp["seed"] not in have]
    if todo:
        log(f"[selector] tuning {len(todo)} matrices into {path}")
        for k in range(0, len(todo), 8):
            if deadline and time.time() > deadline:
                break
            chunk = todo[k:k + 8]
            mats = [autotune.make_matrix(sp) for sp in chunk]
            autotune.tune_many(mats, db_path=path, reps=reps, max_m1=2, row_limit=row_limit,
                               tags=[{"seed": sp["seed"], "kind": sp["kind"]} for sp in chunk],
                               sample_above_ms=SAMPLE_ABOVE_MS)
            log(f"[selector] tuned {min(k + 8, len(todo))}/{len(todo)} "
                f"(largest nnz {max(m.nnz for m in mats)})")
    want = {sp["seed"] for sp in specs}
    return [r for r in autotune.load_db(path) if r.get("seed") in want]


def named_case(name, sel, reps, row_limit):
    """The network's plan against exhaustive tuning on a named configuration."""
    import torch

    from . import autotune, config_matrix
    m = config_matrix(name)
    rec = autotune.features_of(m)
    cands = [dict(length_rule=3)] + autotune.candidates(m.rows, m.cols, 2)
    cands += [dict(c, colpass_rows=r) for c in cands[1:] for r in autotune.COLPASS_ROWS]
    pick = sel.choose(rec, cands)
    best_opts, best_ms, trials = autotune.tune_matrix(m, reps=reps, max_m1=2, record={}, row_limit=row_limit)
    d_r = torch.as_tensor(m.row_idx).cuda()
    d_c = torch.as_tensor(m.col_idx).cuda()
    pick_ms, _, _ = autotune.time_plan(m.rows, m.cols, d_r, d_c, pick, reps, row_limit)
    dflt = [t for t in trials if t.get("default")][0]["ms"]
    log(f"[selector] {name}: selector {autotune.plan_key(pick)} {pick_ms:.3f} ms, exhaustive "
        f"{autotune.plan_key(best_opts)} {best_ms:.3f} ms, default {dflt:.3f} ms")
    return {"selector_plan": pick, "selector_ms": pick_ms, "exhaustive_plan": best_opts,
            "exhaustive_ms": best_ms, "default_ms": dflt, "trials": len(trials),
            "row_limit": row_limit if 0 < row_limit < m.rows // 2 + 1 else 0,
            "regret": pick_ms / best_ms - 1.0, "default_regret": dflt / best_ms - 1.0}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--dir", required=True)
    ap.add_argument("--count", type=int, default=2000)
    ap.add_argument("--heldout", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--row-limit", type=int, default=512)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--skip-c5", action="store_true")
    ap.add_argument("--per-rader-width", type=int, default=16,
                    help="training matrices at each width with a Rader transform (4 per width held out)")
    ap.add_argument("--max-seconds", type=float, default=0,
                    help="tuning budget of this run; the DB resumes in the next run")
    args = ap.parse_args(argv)
    os.makedirs(args.dir, exist_ok=True)

    from . import autotune, selector
    t0 = time.time()
    deadline = t0 + args.max_seconds if args.max_seconds > 0 else None
    specs = autotune.training_specs(args.count, per_rader_width=args.per_rader_width)
    train = tune_set(specs, os.path.join(args.dir, "train.jsonl"), args.row_limit, args.reps, deadline)
    if len(train) < len(specs):
        log(f"[selector] tuning budget spent with {len(train)}/{len(specs)} matrices; run again to resume")
        print(json.dumps({"tuned": len(train), "of": len(specs)}), flush=True)
        return None
    t1 = time.time()
    sel = selector.Selector.fit(train, epochs=args.epochs)
    t_train = time.time() - t1
    sel.save(os.path.join(args.dir, "selector.json"))
    log(f"[selector] trained on {len(train)} matrices in {t_train:.1f} s, loss {sel.train_loss:.4f}")
    fit = selector.regret(sel, train)

    hspecs = autotune.training_specs(args.heldout, seed0=900000, per_rader_width=4)
    held = tune_set(hspecs, os.path.join(args.dir, "heldout.jsonl"), args.row_limit, args.reps, deadline)
    if len(held) < len(hspecs):
        log(f"[selector] tuning budget spent with {len(held)}/{len(hspecs)} held-out matrices; "
            "run again to resume")
        print(json.dumps({"tuned": len(train), "heldout_tuned": len(held)}), flush=True)
        return None
    ho = selector.regret(sel, held)
    nnz = [r["nnz"] for r in train]
    report = {"train_matrices": len(train), "train_rows": int(sum(len(r["trials"]) for r in train)),
              "train_max_nnz": int(max(nnz)), "train_over_2M_nnz": int(sum(x > 2_000_000 for x in nnz)),
              "train_s": round(t_train, 1), "train_loss": sel.train_loss, "fit": fit, "heldout": ho}
    report["c3"] = named_case("c3", sel, args.reps, 0)
    if not args.skip_c5:
        report["c5"] = named_case("c5", sel, args.reps, 0)
    print(json.dumps(report), flush=True)
    return report


if __name__ == "__main__":
    main()
