"""Learned plan selector.

A small MLP (features -> 64 -> 64 -> 1, ReLU, FP32), trained and evaluated on
the GPU by the library's own HIP kernels (csrc/mlp.hip), predicts log(GPU time)
of a (matrix, candidate plan) pair.  For a new matrix the selector scores every
candidate of the autotuner's search space and returns the one with the
smallest prediction.  It only chooses among plans that all compute the same
transform within the accuracy bar; exhaustive autotuning remains the fallback
and the reference for its regret.

Features
  matrix: log2 rows, log2 cols, log2 nnz, log2 average column degree, the
          fraction of active columns and the 8-bin log2 column-degree histogram
  plan:   log2 M, M / (2 cols - 1), the exponents of 2, 3, 5, 7 in M, log2 M1,
          log2 M2, fast convolution kernel, convolution stages, arithmetic form,
          column-pass kernel (one-hot rows per lane 1 / 2 / 4 / 8 and the matrix-core
          kernel, auto = all zero), fused
          row kernel, direct transform, Rader transform, log2 threads per workgroup
          of an on-chip row kernel (0 for the four-step kernels), log2 of the tile's
          share of the R/2 + 1 computed rows (0: one tile)
"""
from __future__ import annotations

import ctypes
import json
import math
import os

import numpy as np

from . import TRANSFORM_DIRECT, TRANSFORM_RADER, TRANSFORM_SPLIT, _check, _vp, describe_plan, load

N_FEATURES = 34
# the network trained on the autotuner timings of the seeded training set (train_selector)
DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "selector.json")


def _exponents(M: int):
    out = []
    for p in (2, 3, 5, 7):
        e = 0
        while M % p == 0:
            M //= p
            e += 1
        out.append(float(e))
    return out


def features(rec: dict, opts: dict, info: dict | None = None, trial: bool = False) -> np.ndarray:
    """Feature row of matrix `rec` (autotune.features_of) with plan `opts` (trial: the
    options of a timed trial, trial_opts)."""
    if info is None:
        o = {k: v for k, v in opts.items() if k != "u_tile"}
        if trial and o.get("row_transform") not in (TRANSFORM_DIRECT, TRANSFORM_RADER):
            # trials recorded before the fused row kernel existed carry no row_kernel:
            # they ran the four-step kernels
            o.setdefault("row_kernel", 1)
        info = describe_plan(rec["rows"], rec["cols"], **o)
    rows, cols, nnz = rec["rows"], rec["cols"], rec["nnz"]
    act = max(rec.get("active_cols", cols), 1)
    M, m1, m2 = info["fft_len"], info["m1"], info["m2"]
    cpr = opts.get("colpass_rows", 0)
    f = [math.log2(rows), math.log2(cols), math.log2(nnz + 1), math.log2(nnz / act + 1),
         act / cols]
    f += list(rec["deg_hist"])
    f += [math.log2(M), M / (2.0 * cols - 1.0)] + _exponents(M)
    f += [math.log2(m1), math.log2(max(m2, 1)), float(info["conv_fast"]), float(info["conv_stages"]),
          float(info["precision"])]
    f += [1.0 if cpr == r else 0.0 for r in (1, 2, 4, 8, 32)]
    f += [1.0 if info.get("row_kernel") == 2 else 0.0]
    nt = info.get("row_threads", 0)
    full = rows // 2 + 1
    tile = min(opts.get("u_tile") or full, opts.get("t_rows") or full, full)
    f += [float(bool(info.get("direct"))), float(bool(info.get("rader"))), math.log2(nt) if nt > 0 else 0.0,
          math.log2(tile / full)]
    assert len(f) == N_FEATURES
    return np.asarray(f, np.float32)


def trial_opts(t: dict, rec: dict | None = None) -> dict:
    """The options of a timed trial with its Bluestein length pinned to the one it ran;
    a trial that ran the direct length-cols transform, Rader's convolutions or the split
    transform (flagged, or, in records written before the flags, a length of cols, of
    p - 1 for cols = 2p or p, or below 2 cols - 1) is planned as that again."""
    o = dict(t["opts"])
    n = t.get("fft_len")
    cols = rec["cols"] if rec is not None else None
    direct = t.get("direct") or (cols is not None and n == cols and n < 2 * cols - 1)
    rader = t.get("rader") or (cols is not None and n == (cols // 2 if cols % 2 == 0 else cols) - 1)
    # the split transform reports the length of its width-n2 pass, below any Bluestein length
    split = t.get("split") or (not direct and not rader and cols is not None and n and n < 2 * cols - 1)
    if split:
        o.pop("fft_len", None)
        o["row_transform"] = TRANSFORM_SPLIT
    elif direct or rader:
        o.pop("fft_len", None)
        o["row_transform"] = TRANSFORM_DIRECT if direct else TRANSFORM_RADER
    elif n:
        o["fft_len"] = n
    return o


def dataset(records: list[dict]):
    """(X, y, groups): one row per timed trial; y = log(ms); groups[i] = record index."""
    X, y, g = [], [], []
    for i, rec in enumerate(records):
        for t in rec["trials"]:
            X.append(features(rec, trial_opts(t, rec), trial=True))
            y.append(math.log(max(t["ms"], 1e-4)))
            g.append(i)
    return np.asarray(X, np.float32), np.asarray(y, np.float32), np.asarray(g, np.int64)


class MLP:
    """Handle of the HIP network (spfft_mlp_*)."""

    def __init__(self, in_dim: int = N_FEATURES, hidden: int = 64, seed: int = 0):
        L = load()
        self._bind(L)
        self.D, self.H = in_dim, hidden
        self._h = _vp()
        _check(L.spfft_mlp_create(in_dim, hidden, seed, ctypes.byref(self._h)))
        self.nparam = int(L.spfft_mlp_num_params(self._h))

    @staticmethod
    def _bind(L):
        if getattr(L, "_mlp_bound", False):
            return
        i, vp, f = ctypes.c_int, ctypes.c_void_p, ctypes.c_float
        sig = {
            "spfft_mlp_create": (i, [i, i, ctypes.c_uint64, ctypes.POINTER(vp)]),
            "spfft_mlp_destroy": (i, [vp]),
            "spfft_mlp_num_params": (ctypes.c_int64, [vp]),
            "spfft_mlp_get_params": (i, [vp, vp]),
            "spfft_mlp_set_params": (i, [vp, vp]),
            "spfft_mlp_predict": (i, [vp, vp, i, vp, vp]),
            "spfft_mlp_loss_grad": (i, [vp, vp, vp, i, ctypes.POINTER(f), vp, vp]),
            "spfft_mlp_train": (i, [vp, vp, vp, i, i, i, f, f, ctypes.c_uint64, ctypes.POINTER(f), vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        L._mlp_bound = True

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load().spfft_mlp_destroy(h)
            self._h = None

    def params(self) -> np.ndarray:
        p = np.zeros(self.nparam, np.float32)
        _check(load().spfft_mlp_get_params(self._h, p.ctypes.data))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert len(p) == self.nparam
        _check(load().spfft_mlp_set_params(self._h, p.ctypes.data))

    def predict(self, X: np.ndarray) -> np.ndarray:
        import torch
        dX = torch.as_tensor(np.ascontiguousarray(X, np.float32)).cuda()
        dy = torch.empty(len(X), dtype=torch.float32, device="cuda")
        _check(load().spfft_mlp_predict(self._h, _vp(dX.data_ptr()), len(X), _vp(dy.data_ptr()),
                                        _vp(torch.cuda.current_stream().cuda_stream)))
        return dy.cpu().numpy()

    def loss_grad(self, X, y):
        import torch
        dX = torch.as_tensor(np.ascontiguousarray(X, np.float32)).cuda()
        dy = torch.as_tensor(np.ascontiguousarray(y, np.float32)).cuda()
        g = np.zeros(self.nparam, np.float32)
        loss = ctypes.c_float()
        _check(load().spfft_mlp_loss_grad(self._h, _vp(dX.data_ptr()), _vp(dy.data_ptr()), len(y),
                                          ctypes.byref(loss), g.ctypes.data,
                                          _vp(torch.cuda.current_stream().cuda_stream)))
        return float(loss.value), g

    def train(self, X, y, epochs=300, batch=256, lr=2e-3, weight_decay=1e-5, seed=0) -> float:
        import torch
        dX = torch.as_tensor(np.ascontiguousarray(X, np.float32)).cuda()
        dy = torch.as_tensor(np.ascontiguousarray(y, np.float32)).cuda()
        loss = ctypes.c_float()
        _check(load().spfft_mlp_train(self._h, _vp(dX.data_ptr()), _vp(dy.data_ptr()), len(y), epochs,
                                      batch, lr, weight_decay, seed, ctypes.byref(loss),
                                      _vp(torch.cuda.current_stream().cuda_stream)))
        return float(loss.value)


class Selector:
    """Trained network + feature standardisation."""

    def __init__(self, mlp: MLP, mean: np.ndarray, std: np.ndarray):
        self.mlp, self.mean, self.std = mlp, mean, std

    @classmethod
    def fit(cls, records, epochs=300, hidden=64, seed=0, lr=2e-3):
        X, y, _ = dataset(records)
        mean, std = X.mean(0), X.std(0) + 1e-6
        mlp = MLP(X.shape[1], hidden, seed)
        loss = mlp.train((X - mean) / std, y, epochs=epochs, lr=lr, seed=seed)
        sel = cls(mlp, mean, std)
        sel.train_loss = loss
        return sel

    def predict(self, X):
        return self.mlp.predict((np.asarray(X, np.float32) - self.mean) / self.std)

    def choose(self, rec: dict, cands: list[dict]) -> dict:
        X = np.stack([features(rec, o) for o in cands])
        return cands[int(np.argmin(self.predict(X)))]

    def save(self, path: str):
        with open(path, "w") as f:
            json.dump({"D": self.mlp.D, "H": self.mlp.H, "mean": self.mean.tolist(),
                       "std": self.std.tolist(), "params": self.mlp.params().tolist()}, f)

    @classmethod
    def load(cls, path: str):
        with open(path) as f:
            d = json.load(f)
        mlp = MLP(d["D"], d["H"])
        mlp.set_params(np.asarray(d["params"], np.float32))
        return cls(mlp, np.asarray(d["mean"], np.float32), np.asarray(d["std"], np.float32))


_default = None


def choose_plan(m, max_m1: int = 2, fp32_only: bool = False) -> dict:
    """Plan options for matrix m (COOMatrix) chosen by the shipped network among
    the autotuner's candidates (cuspfft_amd.autotune.candidates: Bluestein, and the
    direct or Rader transform where the width has one), for Plan(**opts).
    fp32_only: among the plain FP32 plans only (no BF16x3 precision form)."""
    global _default
    from . import autotune
    # the split transform (widths with no on-chip row pass of their own) has no timings in
    # the training set: the planner's plan is returned as it is
    if describe_plan(m.rows, m.cols)["row_transform"] == TRANSFORM_SPLIT:
        return dict(length_rule=3)
    if _default is None:
        _default = Selector.load(DEFAULT_MODEL)
    rec = autotune.features_of(m)
    cands = [dict(length_rule=3)] + autotune.candidates(m.rows, m.cols, max_m1)
    cands += [dict(c, colpass_rows=r) for c in cands[1:] for r in autotune.COLPASS_ROWS]
    if fp32_only:
        cands = [c for c in cands if not c.get("precision")]
    return _default.choose(rec, cands)


def regret(sel: Selector, records: list[dict]) -> dict:
    """Per matrix: measured time of the network's choice among the timed trials
    over the best measured trial (the exhaustive search), and the same ratio for
    the planner's default plan (the method without the network)."""
    ratios, default_ratios = [], []
    for rec in records:
        trials = rec["trials"]
        best = min(t["ms"] for t in trials)
        X = np.stack([features(rec, trial_opts(t, rec), trial=True) for t in trials])
        pick = trials[int(np.argmin(sel.predict(X)))]
        ratios.append(pick["ms"] / best)
        dflt = [t for t in trials if t.get("default")]
        if dflt:
            default_ratios.append(dflt[0]["ms"] / best)
    r = np.asarray(ratios)
    out = {"matrices": len(r), "mean_regret": float(r.mean() - 1), "median_regret": float(np.median(r) - 1),
           "p90_regret": float(np.quantile(r, 0.9) - 1), "exact_best_frac": float(np.mean(r <= 1.0 + 1e-9))}
    if default_ratios:
        d = np.asarray(default_ratios)
        out["default_mean_regret"] = float(d.mean() - 1)
        out["selector_speedup_over_default"] = float(np.exp(np.mean(np.log(d)) - np.mean(np.log(r))))
    return out
