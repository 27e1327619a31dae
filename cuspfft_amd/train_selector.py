"""Trains and evaluates the learned plan selector.

    python -m cuspfft_amd.train_selector --dir OUT [--count 2000] [--heldout 256] [--epochs 300]

1. Tunes the seeded training set (autotune.training_matrices) exhaustively with
   tile-sampled timing, appending to OUT/train.jsonl (resumes where it stopped).
2. Trains the MLP on the GPU with the library's HIP kernels; writes OUT/selector.json.
3. Tunes a fresh held-out batch (disjoint seeds) and reports the regret of the
   network's choice against exhaustive tuning, next to the regret of the
   planner's default plan (the method without the network).
4. The large case (c5, 40421 x 40421): the network's plan against exhaustive tuning.
Prints one JSON report on stdout; progress on stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def tune_set(mats, tags, path, row_limit, reps, deadline=None):
    """Tunes the matrices missing from the DB at `path`; stops at `deadline`
    (time.time() value) with the DB holding everything tuned so far."""
    from . import autotune
    have = {r.get("seed") for r in autotune.load_db(path)}
    todo = [i for i, t in enumerate(tags) if t["seed"] not in have]
    if todo:
        log(f"[selector] tuning {len(todo)} matrices into {path}")
        for k in range(0, len(todo), 16):
            if deadline and time.time() > deadline:
                break
            chunk = todo[k:k + 16]
            autotune.tune_many([mats[i] for i in chunk], db_path=path, reps=reps, max_m1=2,
                               row_limit=row_limit, tags=[tags[i] for i in chunk])
            log(f"[selector] tuned {min(k + 16, len(todo))}/{len(todo)}")
    want = {t["seed"] for t in tags}
    return [r for r in autotune.load_db(path) if r.get("seed") in want]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--dir", required=True)
    ap.add_argument("--count", type=int, default=2000)
    ap.add_argument("--heldout", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--row-limit", type=int, default=512)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--skip-c5", action="store_true")
    ap.add_argument("--max-seconds", type=float, default=0,
                    help="tuning budget of this run; the DB resumes in the next run")
    args = ap.parse_args(argv)
    os.makedirs(args.dir, exist_ok=True)

    import numpy as np

    from . import autotune, config_matrix, selector
    t0 = time.time()
    deadline = t0 + args.max_seconds if args.max_seconds > 0 else None
    mats, tags = autotune.training_matrices(args.count)
    log(f"[selector] generated {len(mats)} training matrices in {time.time() - t0:.0f} s")
    train = tune_set(mats, tags, os.path.join(args.dir, "train.jsonl"), args.row_limit, args.reps,
                     deadline)
    del mats
    if len(train) < args.count:
        log(f"[selector] tuning budget spent with {len(train)}/{args.count} matrices; run again to resume")
        print(json.dumps({"tuned": len(train), "of": args.count}), flush=True)
        return None
    t1 = time.time()
    sel = selector.Selector.fit(train, epochs=args.epochs)
    t_train = time.time() - t1
    sel.save(os.path.join(args.dir, "selector.json"))
    log(f"[selector] trained on {len(train)} matrices in {t_train:.1f} s, loss {sel.train_loss:.4f}")
    fit = selector.regret(sel, train)

    hmats, htags = autotune.training_matrices(args.heldout, seed0=900000)
    held = tune_set(hmats, htags, os.path.join(args.dir, "heldout.jsonl"), args.row_limit, args.reps,
                    deadline)
    del hmats
    if len(held) < args.heldout:
        log(f"[selector] tuning budget spent with {len(held)}/{args.heldout} held-out matrices; "
            "run again to resume")
        print(json.dumps({"tuned": len(train), "heldout_tuned": len(held)}), flush=True)
        return None
    ho = selector.regret(sel, held)
    report = {"train_matrices": len(train), "train_rows": int(sum(len(r["trials"]) for r in train)),
              "train_s": round(t_train, 1), "train_loss": sel.train_loss, "fit": fit, "heldout": ho}

    if not args.skip_c5:
        m = config_matrix("c5")
        rec = autotune.features_of(m)
        cands = [dict(length_rule=3)] + autotune.candidates(m.rows, m.cols, 2)
        cands += [dict(c, colpass_rows=r) for c in cands[1:] for r in autotune.COLPASS_ROWS]
        pick = sel.choose(rec, cands)
        exh = {}
        best_opts, best_ms, trials = autotune.tune_matrix(m, reps=args.reps, max_m1=2, record=exh,
                                                          row_limit=args.row_limit)
        import torch
        d_r = torch.as_tensor(m.row_idx).cuda()
        d_c = torch.as_tensor(m.col_idx).cuda()
        pick_ms, _, _ = autotune.time_plan(m.rows, m.cols, d_r, d_c, pick, args.reps, args.row_limit)
        dflt = [t for t in trials if t.get("default")][0]["ms"]
        report["c5"] = {"selector_plan": pick, "selector_ms": pick_ms, "exhaustive_plan": best_opts,
                        "exhaustive_ms": best_ms, "default_ms": dflt,
                        "regret": pick_ms / best_ms - 1.0, "default_regret": dflt / best_ms - 1.0}
    print(json.dumps(report), flush=True)
    return report


if __name__ == "__main__":
    main()
