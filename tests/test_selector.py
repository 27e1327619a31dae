"""The learned plan selector: the FP64 reference network pinned by finite
differences (CPU), the HIP forward / backward / Adam kernels checked against it
(GPU), and the trained network against the planner's default plan."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
DB = os.path.join(HERE, "golden", "selector_timings.jsonl")


def _rand_params(D, H, seed):
    rng = np.random.default_rng(seed)
    n = H * D + H + H * H + H + H + 1
    return rng.normal(0, 0.5, n)


def test_reference_gradient_finite_differences():
    from oracle import mlp_ref
    D, H, B = 5, 7, 9
    rng = np.random.default_rng(1)
    X = rng.normal(size=(B, D))
    y = rng.normal(size=B)
    p = _rand_params(D, H, 2)
    loss, g = mlp_ref.loss_grad(p, X, y, D, H)
    eps = 1e-6
    for i in rng.choice(len(p), 40, replace=False):
        q = p.copy(); q[i] += eps
        lp, _ = mlp_ref.loss_grad(q, X, y, D, H)
        q[i] -= 2 * eps
        lm, _ = mlp_ref.loss_grad(q, X, y, D, H)
        assert abs((lp - lm) / (2 * eps) - g[i]) < 1e-6 * max(1.0, abs(g[i]))
    # a planted mistake (dropping the ReLU mask of layer 2) changes the gradient
    W1, b1, W2, b2, W3, b3 = mlp_ref.unpack(p, D, H)
    Z2 = np.maximum(X @ W1.T + b1, 0) @ W2.T + b2
    assert (Z2 <= 0).any()


def test_features_shape_and_determinism():
    import cuspfft_amd as sp
    from cuspfft_amd import autotune, selector
    m = sp.config_matrix("c1")
    rec = autotune.features_of(m)
    assert abs(sum(rec["deg_hist"]) - 1) < 1e-9
    cands = autotune.candidates(m.rows, m.cols, 3)
    assert any(c["precision"] == 1 for c in cands)
    f = np.stack([selector.features(rec, c) for c in cands])
    assert f.shape == (len(cands), selector.N_FEATURES) and np.all(np.isfinite(f))
    np.testing.assert_array_equal(f, np.stack([selector.features(rec, c) for c in cands]))


def test_trial_opts_replans_each_transform():
    """A timed trial is planned again as the transform it ran: Bluestein at its length,
    the direct transform, or Rader's convolutions (flagged, or recognised by their length
    in records written before the flags)."""
    import cuspfft_amd as sp
    from cuspfft_amd import selector
    rec = {"cols": 16706}
    blue = selector.trial_opts({"opts": dict(length_rule=3), "fft_len": 34992}, rec)
    assert blue["fft_len"] == 34992 and "row_transform" not in blue
    for t in ({"opts": dict(length_rule=3), "fft_len": 8352, "rader": 1},
              {"opts": dict(length_rule=3), "fft_len": 8352}):
        o = selector.trial_opts(t, rec)
        assert o["row_transform"] == sp.TRANSFORM_RADER and "fft_len" not in o
        assert sp.describe_plan(16706, 16706, **o)["rader"] == 1
    d = selector.trial_opts({"opts": dict(length_rule=3), "fft_len": 16384}, {"cols": 16384})
    assert d["row_transform"] == sp.TRANSFORM_DIRECT


def test_candidates_cover_transforms_and_tiles():
    """The autotuner's search space holds, beside the Bluestein plans, Rader's
    transform in both arithmetic forms (and both workgroup sizes of the BF16x3 kernel
    at 16706) and the direct transform where the width has them; Bluestein candidates
    say so where AUTO would pick another transform; the tile variants stay below the
    R/2 + 1 computed rows."""
    import cuspfft_amd as sp
    from cuspfft_amd import autotune
    cands = autotune.candidates(16706, 16706, 2)
    infos = [sp.describe_plan(16706, 16706, **c) for c in cands]
    rader = [(i["precision"], i["row_threads"]) for i in infos if i["rader"]]
    assert sorted(rader) == [(0, 704), (1, 576), (1, 704)]
    assert all(c.get("row_transform") == sp.TRANSFORM_BLUESTEIN for c, i in zip(cands, infos) if not i["rader"])
    assert len({(i["fft_len"], i["m1"], i["precision"], i["row_kernel"], i["rader"], i["row_threads"])
                for i in infos}) == len(infos)
    d = [sp.describe_plan(4096, 4096, **c) for c in autotune.candidates(4096, 4096, 2)]
    assert sum(i["direct"] for i in d) == 1
    c1 = autotune.candidates(2310, 2310, 2)
    assert all("row_transform" not in c for c in c1)  # no other transform at 2310
    four = [c for c in c1 if sp.describe_plan(2310, 2310, **c)["row_kernel"] == sp.ROW_FOUR_STEP][0]
    t = autotune.tile_candidates(2310, four, sp.describe_plan(2310, 2310, **four))
    assert [x["t_rows"] for x in t] == [1024]
    fused = [c for c in c1 if sp.describe_plan(2310, 2310, **c)["row_kernel"] == sp.ROW_FUSED][0]
    assert [x["u_tile"] for x in autotune.tile_candidates(2310, fused, {"row_kernel": sp.ROW_FUSED})] == [1024]
    assert autotune.tile_candidates(1000, fused, {"row_kernel": sp.ROW_FUSED}) == []


def test_training_set_is_seeded():
    import cuspfft_amd as sp
    from cuspfft_amd import autotune
    a, ta = autotune.training_matrices(6, nnz_cap=20000, per_rader_width=1)
    b, tb = autotune.training_matrices(6, nnz_cap=20000, per_rader_width=1)
    assert ta == tb and len(ta) == 6 + len(sp.rader_widths())
    for x, y in zip(a, b):
        assert x.rows == y.rows and x.nnz == y.nnz
        np.testing.assert_array_equal(x.col_idx, y.col_idx)
    assert {t["kind"] for t in ta} == {"uniform", "chunglu"}
    assert sorted(x.cols for x in a[6:]) == sorted(sp.rader_widths())
    # the default cap does not bind anywhere in the 2000-matrix set: its densest
    # matrices (density up to 1e-2 at n up to 40000) keep their ~16M nonzeros
    specs = autotune.training_specs(2000)
    big = [s["nnz"] for s in specs if s["nnz"] > 2_000_000]
    assert len(big) > 50 and max(big) > 15_000_000
    assert all(s["nnz"] < 16_000_000 for s in specs)


@pytest.mark.gpu
def test_hip_forward_and_gradient_match_reference():
    from cuspfft_amd import selector
    from oracle import mlp_ref
    D, H, B = 27, 64, 300
    net = selector.MLP(D, H, seed=3)
    p = net.params().astype(np.float64)
    rng = np.random.default_rng(4)
    X = rng.normal(size=(B, D)).astype(np.float32)
    y = rng.normal(size=B).astype(np.float32)
    yhat_ref, _ = mlp_ref.forward(p, X, D, H)
    yhat = net.predict(X)
    assert np.abs(yhat - yhat_ref).max() < 1e-4 * max(1.0, np.abs(yhat_ref).max())
    loss, g = net.loss_grad(X, y)
    loss_ref, g_ref = mlp_ref.loss_grad(p, X, y, D, H)
    assert abs(loss - loss_ref) < 1e-5 * max(1.0, loss_ref)
    assert np.linalg.norm(g - g_ref) < 1e-4 * np.linalg.norm(g_ref)


@pytest.mark.gpu
def test_hip_training_fits_a_smooth_target():
    from cuspfft_amd import selector
    rng = np.random.default_rng(5)
    X = rng.normal(size=(4096, 8)).astype(np.float32)
    y = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.3 * np.abs(X[:, 3])).astype(np.float32)
    net = selector.MLP(8, 64, seed=0)
    l0, _ = net.loss_grad(X[:512], y[:512])
    loss = net.train(X, y, epochs=60, batch=256, lr=3e-3, seed=1)
    assert loss < 0.1 * l0
    # deterministic for a fixed seed
    net2 = selector.MLP(8, 64, seed=0)
    loss2 = net2.train(X, y, epochs=60, batch=256, lr=3e-3, seed=1)
    assert loss2 == loss


@pytest.mark.gpu
def test_selector_beats_default_plan_on_heldout():
    """Trained on measured autotuner timings (committed fixture), the network's
    choice on held-out matrices is at least as fast as the planner's default plan."""
    from cuspfft_amd import autotune, selector
    recs = autotune.load_db(DB)
    assert len(recs) >= 40
    train, held = recs[: int(len(recs) * 0.75)], recs[int(len(recs) * 0.75):]
    sel = selector.Selector.fit(train, epochs=200)
    r = selector.regret(sel, held)
    print(r)
    assert r["mean_regret"] <= r["default_mean_regret"] + 1e-9
    assert r["selector_speedup_over_default"] >= 1.0


@pytest.mark.gpu
def test_shipped_selector_plan_runs(oracle_mod):
    """The shipped network picks a plan for the c1, c2 and c3 shapes that computes the
    transform within the bar of its arithmetic form."""
    import torch
    import cuspfft_amd as sp
    from cuspfft_amd import selector
    for cfg in ("c1", "c2", "c3"):
        m = sp.config_matrix(cfg)
        opts = selector.choose_plan(m)
        plan = sp.Plan(m.rows, m.cols, u_tile=m.rows, **opts)
        out = plan.alloc_output()
        plan.execute(torch.as_tensor(m.row_idx).cuda(), torch.as_tensor(m.col_idx).cuda(), out)
        torch.cuda.synchronize()
        plan.check_input()
        rows = np.array([0, 1, m.rows // 2, m.rows - 1])
        ref = oracle_mod.sparse_rfft2_f64(m.rows, m.cols, m.row_idx, m.col_idx, u_list=rows, nthreads=8)
        got = out[torch.as_tensor(rows).cuda(), : m.cols // 2 + 1].cpu().numpy()
        tol = 1e-5 if opts.get("precision") == sp.PREC_BF16X3 else 9e-6
        print(cfg, opts, oracle_mod.rel_l2(got, ref))
        assert oracle_mod.rel_l2(got, ref) <= tol
