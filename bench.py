"""Benchmark: 2-D sparse FFTs/s and HBM roofline fraction on one MI355X.

One step = one full binary-sparse 2-D R2C FFT of the workload matrix, from a
device-resident int32 COO to the device-resident complex64 half spectrum:
COO->CSC build + column sparse-DFT pass + row pass (all on the GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3] [--no-cpu]

Prints ONE JSON line.  value = FFTs/s over the K timed transforms (no per-kernel
events in them) of the plan under test: by default the planner's plan (for c3 the
Rader row transform; --plan selector: the shipped selector's choice).  Beside it,
`forms` times the other forms of the metric on the same input, each with its
rel-L2 error against the FP64 oracle: the plain FP32 Bluestein plan at the reference's
length (the reference's algorithm) and the BF16x3 MFMA precision form, and their
ratios.  Roofline: algorithmic bytes per FFT B_alg = 8*nnz (COO int32 read) +
8*rows*(cols/2+1) (complex64 output write); achieved = B_alg / (the dominant kernel's
average launch time, from HIP events on the launch stream over K more transforms
after the timed ones); peak = 8 TB/s (MI355X HBM3E).  cpu_baseline: scipy.fft.rfft2
(pocketfft) in float32 on the dense matrix, 1 thread and all host threads (FFTW, the
reference's CPU path, is not installed).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # defaults: 50 untimed transforms (about 0.1 s on c3) let the GPU clocks settle
    # before the timed ones; fewer measured 3-8% slow on the first plan of a process
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--plan", default="default", choices=("selector", "default"),
                    help="plan of the timed transform when no plan option is given: the planner's "
                         "default, or the shipped selector's choice among the plain FP32 candidates")
    ap.add_argument("--rule", type=int, default=3)
    ap.add_argument("--tile", type=int, default=0, help="u-tile rows (0 = one tile of every computed row)")
    ap.add_argument("--row-threads", type=int, default=0,
                    help="on-chip row kernels: threads per workgroup of the shape (0 = the first listed)")
    ap.add_argument("--t-rows", type=int, default=0,
                    help="four-step kernels: rows of the tile buffer T per sub-tile (0 = automatic)")
    ap.add_argument("--m1", type=int, default=0)
    ap.add_argument("--fft-len", type=int, default=0, help="explicit Bluestein length M (0: the rule's)")
    ap.add_argument("--precision", type=int, default=0, help="0 plain FP32, 1 BF16x3 MFMA")
    ap.add_argument("--colpass-rows", type=int, default=0)
    ap.add_argument("--colpass-table", type=int, default=0)
    ap.add_argument("--colpass-pair", type=int, default=0,
                    help="0 auto (rows u and R/2 - u share the column pass when R is even), 1 off")
    ap.add_argument("--row-transform", type=int, default=0,
                    help="0 auto (direct length-cols FFT, else Rader for cols = 2p, when there is a "
                         "shape), 1 Bluestein, 2 direct, 3 Rader")
    ap.add_argument("--row-persist", type=int, default=0,
                    help="fused row kernel: 0 persistent workgroups, 1 one workgroup per row group")
    ap.add_argument("--row-kernel", type=int, default=0,
                    help="0 auto (fused row kernel when M has a shape), 1 four-step kernels, 2 fused")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--batch", type=int, default=512, help="c4: number of matrices")
    ap.add_argument("--tune-db", default="", help="c4: timing DB (JSON lines) to read or append")
    ap.add_argument("--streams", type=int, default=4, help="c4: worker streams")
    ap.add_argument("--c4-plans", default="selector", choices=("selector", "tune", "default"),
                    help="c4: per-matrix plans from the shipped selector, exhaustive autotuning, or the planner")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the other forms (plain Bluestein, BF16x3 MFMA) beside the timed plan")
    ap.add_argument("--to-host", type=int, default=0, metavar="TILE",
                    help="also time the output streamed to pinned host memory with u-tiles of TILE rows")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    return ap.parse_args()


def copy_calibration(mib: int):
    """A plain device-to-device copy of known size, so that the FETCH_SIZE / WRITE_SIZE
    counters of the profiled kernels can be scaled to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_each_way": n * 4, "launches": 3}), flush=True)


def run_batch_bench(args):
    """c4: the batch of uneven matrices -- untuned eager execution vs autotuned
    plans captured as one hipGraph."""
    import numpy as np
    import torch

    import cuspfft_amd as sp
    from cuspfft_amd import autotune
    t_gen = time.perf_counter()
    mats = autotune.batch_matrices(args.batch)
    t_gen = time.perf_counter() - t_gen
    print(f"[c4] generated {len(mats)} matrices in {t_gen:.1f} s, total nnz "
          f"{sum(m.nnz for m in mats)}", file=sys.stderr, flush=True)
    d_row, d_col, off = sp.Batch.pack(mats)
    shapes = [(m.rows, m.cols) for m in mats]
    hint = [m.nnz for m in mats]

    def run(batch, graph: bool):
        arena = batch.alloc_output()
        if graph:
            g = batch.capture(d_row, d_col, off, arena)
            go = g.launch
        else:
            g = None
            go = lambda: batch.execute(d_row, d_col, off, arena)  # noqa: E731
        for _ in range(args.warmup):
            go()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            go()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        batch.check_input()
        return dt, arena

    base = sp.Batch(shapes, opts=[dict(u_tile=m.rows) for m in mats], nnz_hint=hint,
                    streams=args.streams)
    dt_base, _ = run(base, graph=False)
    print(f"[c4] untuned eager: {dt_base * 1e3:.2f} ms per batch", file=sys.stderr, flush=True)
    del base
    t_tune = time.perf_counter()
    if args.c4_plans == "tune":
        # exhaustive autotuning: from the DB when it has every matrix, else tune now
        db = autotune.load_db(args.tune_db) if args.tune_db else []
        by_shape = {(r["rows"], r["cols"], r["nnz"]): r["best"] for r in db}
        missing = [i for i, m in enumerate(mats) if (m.rows, m.cols, m.nnz) not in by_shape]
        if missing:
            tuned = autotune.tune_many([mats[i] for i in missing], db_path=args.tune_db or None,
                                       log=lambda x: print(x, file=sys.stderr, flush=True))
            for i, o in zip(missing, tuned):
                by_shape[(mats[i].rows, mats[i].cols, mats[i].nnz)] = o
        opts = [dict(by_shape[(m.rows, m.cols, m.nnz)], u_tile=m.rows) for m in mats]
    elif args.c4_plans == "selector":
        # the shipped network's choice per matrix (the autotuner's plans without the tuning)
        from cuspfft_amd import selector
        opts = [dict(selector.choose_plan(m, fp32_only=True), u_tile=m.rows) for m in mats]
    else:
        opts = [dict(u_tile=m.rows) for m in mats]
    t_tune = time.perf_counter() - t_tune
    tuned = sp.Batch(shapes, opts=opts, nnz_hint=hint, streams=args.streams)
    _, members = tuned.mode()
    # the same plans, every matrix on its own launches (worker streams), as one graph
    tuned.set_mode(sp.BATCH_PER_MATRIX)
    dt_per, _ = run(tuned, graph=True)
    print(f"[c4] per-matrix launches, graph: {dt_per * 1e3:.2f} ms per batch", file=sys.stderr, flush=True)
    tuned.set_mode(sp.BATCH_AUTO)
    dt_tuned, arena = run(tuned, graph=True)
    print(f"[c4] {args.c4_plans} plans, graph: {dt_tuned * 1e3:.2f} ms per batch (planning {t_tune:.0f} s)",
          file=sys.stderr, flush=True)
    rel = None
    if not args.no_check:
        import oracle
        oracle.build()
        worst = 0.0
        for i in range(0, len(mats), max(1, len(mats) // 16)):
            m = mats[i]
            rows = np.array([0, 1, m.rows // 3, m.rows - 1], np.int32)
            ref = oracle.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, u_list=rows, nthreads=8)
            X = tuned.output_of(arena, i)[torch.as_tensor(rows).cuda(), : m.cols // 2 + 1].cpu().numpy()
            worst = max(worst, oracle.rel_l2(X, ref))
        rel = worst
    forms = {}
    for o in opts:
        k = ("bf16x3" if o.get("precision") else "fp32") + f"-rule{o.get('length_rule', 3)}"
        forms[k] = forms.get(k, 0) + 1
    b_alg = sum(8.0 * m.nnz + 8.0 * m.rows * (m.cols // 2 + 1) for m in mats)
    line = {
        "metric": "2-D sparse FFTs/s + %HBM roofline, MFMA-split vs plain FP32 at rel-L2\u22641e-5",
        "value": round(len(mats) / dt_tuned, 2), "unit": "FFTs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt_tuned * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {"workload": f"c4: batch of {len(mats)} seeded binary symmetric matrices, n "
                               "log-uniform in [500, 12000], average degree U[4, 30], half uniform "
                               "and half Chung-Lu (exponent U[2.0, 2.5])",
                   "streams": args.streams, "plans": forms, "plan_source": args.c4_plans,
                   "shared_launch_members": members},
        "per_matrix_graph_ms_per_batch": round(dt_per * 1e3, 3),
        "untuned_eager_ms_per_batch": round(dt_base * 1e3, 3),
        "graph_speedup_over_untuned_eager": round(dt_base / dt_tuned, 3),
        "planning_s": round(t_tune, 1),
        "worst_rel_l2_vs_fp64_sampled": rel,
        "step_hbm_frac": round(b_alg / dt_tuned / 1e9 / HBM_PEAK_GBS, 5),
    }
    print(json.dumps(line), flush=True)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(m):
    """The reference's CPU path is FFTW's single-precision dense R2C 2-D transform with
    only fftwf_execute timed (src/cpu_reference.cpp:8-64).  FFTW is not installed, so the
    labelled substitute is scipy.fft.rfft2 (pocketfft) on the dense float32 matrix: the
    dense scatter is untimed, one transform is timed with 1 worker (the reference's
    single thread) and one with every host thread."""
    import numpy as np
    import scipy
    import scipy.fft
    dense = np.zeros((m.rows, m.cols), np.float32)
    dense[m.row_idx, m.col_idx] = 1.0
    # every host thread this process may use: the CPU affinity set, capped by
    # OMP_NUM_THREADS where the machine sets it to this job's share
    nproc = len(os.sched_getaffinity(0)) or 1
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        nproc = max(1, min(nproc, int(os.environ["OMP_NUM_THREADS"])))
    out = {}
    for workers in (1, nproc):
        t0 = time.perf_counter()
        X = scipy.fft.rfft2(dense, workers=workers)
        out[workers] = time.perf_counter() - t0
        del X
    del dense
    kind = "scipy-pocketfft-f32 (FFTW absent)"
    return {"value": round(1.0 / out[1], 5), "unit": "FFTs/s", "cores": 1, "kind": kind,
            "sample": f"one full {m.rows}x{m.cols} float32 rfft2, transform only "
                      f"(scipy {scipy.__version__}, workers=1)",
            "all_cores": {"value": round(1.0 / out[nproc], 5), "cores": nproc, "workers": nproc},
            "host": {"threads_used": nproc, "os_cpu_count": os.cpu_count(), "cpu_model": _cpu_model()}}


def time_eager(plan, d_row, d_col, out, steps: int, warmup: int) -> float:
    """Seconds per transform over `steps` eager executions after `warmup` untimed ones."""
    import torch
    for _ in range(warmup):
        plan.execute(d_row, d_col, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.execute(d_row, d_col, out)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the row-tile sharding across GPUs is a design note, not implemented", file=sys.stderr)
        sys.exit(2)
    if args.copy_calib:
        copy_calibration(args.copy_calib)
        return
    if args.config == "c4":
        run_batch_bench(args)
        return
    import numpy as np
    import torch

    import cuspfft_amd as sp
    sp.load()
    dev = torch.device("cuda:0")
    m = sp.config_matrix(args.config)
    explicit = (args.tile > 0 or args.fft_len or args.m1 or args.precision or args.rule != 3
                or args.colpass_rows or args.colpass_table or args.row_kernel or args.colpass_pair
                or args.row_transform or args.row_persist or args.t_rows or args.row_threads)
    opts = dict(length_rule=args.rule, m1=args.m1, fft_len=args.fft_len, precision=args.precision,
                row_kernel=args.row_kernel, colpass_rows=args.colpass_rows,
                colpass_table=args.colpass_table, colpass_pair=args.colpass_pair,
                row_transform=args.row_transform, row_persist=args.row_persist, t_rows=args.t_rows,
                row_threads=args.row_threads)
    if explicit or args.plan == "default":
        plan_source = "options" if explicit else "planner"
    else:
        # the shipped network's choice among the autotuner's plain FP32 candidates: the
        # same transform, the plan predicted fastest
        from cuspfft_amd import selector
        opts = selector.choose_plan(m, fp32_only=True)
        plan_source = "selector"
    plan = sp.Plan(m.rows, m.cols, u_tile=args.tile, **opts)
    info = plan.info
    transform = ("direct" if info["direct"] else "rader" if info["rader"]
                 else "split" if info["row_transform"] == sp.TRANSFORM_SPLIT else "bluestein")
    if plan_source == "planner":
        plan_source = f"planner ({transform}; selector not used)"
    d_row = torch.as_tensor(m.row_idx).to(dev)
    d_col = torch.as_tensor(m.col_idx).to(dev)
    out = plan.alloc_output()
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        plan.execute(d_row, d_col, out)
    torch.cuda.synchronize()

    # the timed region: K transforms, no per-kernel events
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.execute(d_row, d_col, out)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    print(f"[bench] {args.config} {transform} M={info['fft_len']}: {(t1 - t0) * 1e3 / args.steps:.4f} ms",
          file=sys.stderr, flush=True)
    # then K more with HIP events around every kernel on the launch stream: the
    # per-kernel times and the dominant kernel's launch time for the roofline
    plan.profile(True)
    for _ in range(args.steps):
        plan.execute(d_row, d_col, out)
    prof = plan.profile_read(stream)
    plan.profile(False)
    plan.check_input()
    if plan.info["row_transform"] == sp.TRANSFORM_SPLIT:
        # split transform: the width-m2 row kernel (timed as the fft1 stage) and
        # k_splitcols (timed as ifft1_finalize)
        names = {"fft1": "split_rows", "ifft1_finalize": "split_cols"}
        prof = {names.get(k, k): v for k, v in prof.items() if k != "conv"}
    elif plan.info["row_kernel"] == sp.ROW_FUSED:
        # one kernel (k_rowfused / k_rowrader) does fft1 + conv + ifft1: reported as row_fused
        prof = {("row_fused" if k == "conv" else k): v for k, v in prof.items()
                if k not in ("fft1", "ifft1_finalize")}
    elif 0 < info["t_rows"] < info["u_tile"]:
        # sub-tiles of T: the three four-step kernels are timed together
        prof = {("row_four_step" if k == "conv" else k): v for k, v in prof.items()
                if k not in ("fft1", "ifft1_finalize")}

    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    value = 1000.0 / ms_per_step
    V = m.cols // 2 + 1
    b_alg = 8.0 * m.nnz + 8.0 * m.rows * V
    dom = max(prof, key=lambda k: prof[k][0])
    dom_ms, dom_n = prof[dom]
    achieved = b_alg * args.steps / (dom_ms * 1e-3) / 1e9
    step_gbs = b_alg / (ms_per_step * 1e-3) / 1e9

    # HBM bytes per launch of the dominant kernel: not measured in this run -- the
    # rocprofv3 FETCH_SIZE / WRITE_SIZE record of the same kernel and plan shape on this
    # workload, kept in profiles/traffic.json (null when there is none)
    traffic, traffic_source = None, None
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            recs = json.load(f).get(args.config, {}).get(dom) or []
        for rec in recs if isinstance(recs, list) else [recs]:
            if (rec.get("fft_len") == info["fft_len"] and rec.get("m1") == info["m1"]
                    and rec.get("row_kernel") == info["row_kernel"]
                    and rec.get("precision", 0) == info["precision"] and args.tile <= 0):
                traffic = rec["bytes"]
                traffic_source = (f"rocprofv3 FETCH_SIZE + WRITE_SIZE of {rec['kernel']} on this "
                                  f"workload and plan shape, from profiles/traffic.json ({rec['source']}); "
                                  "not measured in this run")
                break
    except (OSError, ValueError):
        traffic = None

    rel, rows, ref = None, None, None
    if not args.no_check:
        import oracle
        oracle.build()
        rng = np.random.default_rng(0)
        rows = np.unique(np.concatenate([[0, m.rows - 1], rng.choice(m.rows, 14, replace=False)]))
        ref = oracle.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, u_list=rows,
                                      nthreads=8)
        got = out[torch.as_tensor(rows).to(dev), :V].cpu().numpy()
        rel = oracle.rel_l2(got, ref)

    def plan_desc(i):
        return {"fft_len": i["fft_len"], "m1": i["m1"], "m2": i["m2"],
                "row_kernel": {1: "four-step", 2: "fused"}[i["row_kernel"]],
                "transform": ("direct" if i["direct"] else "rader" if i["rader"]
                              else "split" if i["row_transform"] == sp.TRANSFORM_SPLIT else "bluestein"),
                "precision": "bf16x3-mfma" if i["precision"] else "fp32"}

    forms = None
    if not args.no_forms:
        # the forms of the metric on the same input, timed like `value` (eager, K steps
        # after W warm-up ones): the timed plan; the plain FP32 Bluestein plan at the
        # reference's length (the reference's algorithm, hand-written Stockham stages);
        # the BF16x3 MFMA precision form (its planner's plan)
        forms = {"timed_plan": dict(plan_desc(info), ms=round(ms_per_step, 4), fft_per_s=round(value, 3),
                                    rel_l2_vs_fp64=rel)}
        for name, kw in (("plain_fp32_bluestein", dict(length_rule=args.rule, row_transform=1)),
                         ("bf16x3_mfma", dict(length_rule=args.rule, precision=1))):
            try:
                fplan = sp.Plan(m.rows, m.cols, **kw)
            except sp.SpfftError as e:
                forms[name] = {"error": str(e)}
                continue
            fout = fplan.alloc_output()
            dt = time_eager(fplan, d_row, d_col, fout, args.steps, args.warmup)
            fplan.check_input()
            err = None
            if ref is not None:
                err = oracle.rel_l2(fout[torch.as_tensor(rows).to(dev), :V].cpu().numpy(), ref)
            forms[name] = dict(plan_desc(fplan.info), ms=round(dt * 1e3, 4), fft_per_s=round(1.0 / dt, 3),
                               rel_l2_vs_fp64=err)
            print(f"[bench] form {name}: {dt * 1e3:.4f} ms rel_l2={err}", file=sys.stderr, flush=True)
            del fplan, fout
        plain = forms["timed_plan"] if not info["precision"] else forms["plain_fp32_bluestein"]
        if "fft_per_s" in forms.get("bf16x3_mfma", {}):
            mf = forms["bf16x3_mfma"]["fft_per_s"]
            forms["mfma_over_plain"] = round(mf / plain["fft_per_s"], 3)
            if "fft_per_s" in forms["plain_fp32_bluestein"]:
                forms["mfma_over_plain_bluestein"] = round(mf / forms["plain_fp32_bluestein"]["fft_per_s"], 3)

    streamed = None
    if args.to_host:
        # the output streamed to pinned host memory (device-to-host copies included)
        splan = sp.Plan(m.rows, m.cols, length_rule=args.rule, u_tile=args.to_host)
        host = splan.alloc_host_output()
        for _ in range(args.warmup):
            splan.execute_to_host(d_row, d_col, host)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            splan.execute_to_host(d_row, d_col, host)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        err = None
        if ref is not None:
            err = oracle.rel_l2(host[torch.as_tensor(rows), :V].numpy(), ref)
        streamed = {"fft_per_s": round(1.0 / dt, 3), "ms": round(dt * 1e3, 4), "u_tile": args.to_host,
                    "d2h_gb_per_s": round(m.rows * V * 8 / dt / 1e9, 2),
                    "device_chunk_bytes": 2 * 2 * min(args.to_host, m.rows // 2 + 1) * splan.pitch * 8,
                    "rel_l2_vs_fp64": err}
        del splan, host

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(m)

    line = {
        "metric": "2-D sparse FFTs/s + %HBM roofline, MFMA-split vs plain FP32 at rel-L2\u22641e-5",
        "value": round(value, 3),
        "unit": "FFTs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3" if info["precision"] else "fp32",
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {
            "workload": f"{args.config}: {sp.CONFIG_TEXT[args.config]}",
            "rows": m.rows, "cols": m.cols, "nnz": m.nnz,
            "form": "bf16x3-mfma" if info["precision"] else "plain-fp32",
            "length_rule": args.rule, "fft_len": info["fft_len"], "m1": info["m1"],
            "m2": info["m2"], "u_tile": info["u_tile"], "t_rows": info["t_rows"],
            "row_threads": info["row_threads"],
            "conv_rows": info["conv_rows"],
            "conv_radix": info["conv_radix"], "colpass_rows": info["colpass_rows"],
            "colpass_full_table": info["colpass_full_table"],
            "colpass_paired": info["colpass_paired"],
            "row_kernel": {1: "four-step", 2: "fused"}[info["row_kernel"]],
            "row_transform": transform,
            "plan": plan_source,
        },
        "rel_l2_vs_fp64": rel,
        "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 5),
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in prof.items()},
        "roofline": {
            "bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
            "traffic_source": traffic_source, "b_alg": int(b_alg),
            "avg_launch_ms": round(dom_ms / max(dom_n, 1), 4),
        },
        "forms": forms,
        "streamed_to_host": streamed,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
