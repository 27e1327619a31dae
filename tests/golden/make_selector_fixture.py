"""Writes tests/golden/selector_timings.jsonl: a fixed subset of a timing DB
measured on an MI355X by the autotuner (python -m cuspfft_amd.train_selector
--dir DIR writes DIR/train.jsonl), for the selector test that must not tune on
the GPU itself.

    python tests/golden/make_selector_fixture.py DIR/train.jsonl [--count 240]

Every k-th record is kept (k = records // count, so the subset spans the seeded
training set's sizes and densities); the per-stage times are dropped to keep
the file small.  The records are measurements, not outputs of this repository's
code: the test trains the network on three quarters of them and checks its
choices on the rest.
"""
import argparse
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--count", type=int, default=240)
    ap.add_argument("--out", default=os.path.join(HERE, "selector_timings.jsonl"))
    args = ap.parse_args()
    with open(args.db) as f:
        recs = [json.loads(line) for line in f if line.strip()]
    recs.sort(key=lambda r: r["seed"])
    step = max(1, len(recs) // args.count)
    keep = recs[::step][: args.count]
    with open(args.out, "w") as f:
        for r in keep:
            r = dict(r)
            r["trials"] = [{k: v for k, v in t.items() if k != "stages"} for t in r["trials"]]
            f.write(json.dumps(r, separators=(",", ":")) + "\n")
    print(f"wrote {len(keep)} of {len(recs)} records to {args.out}")


if __name__ == "__main__":
    main()
